// spectral.h — internal: the host record of a batch spectral plan (the
// resampling, window, twiddle and chirp tables a handle keeps in device
// memory for one waveform grid and spectral configuration).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace spx {

struct SpectralPlan {
  bool valid = false;
  // what the tables were built for
  std::vector<double> times;  // empty: uniformly sampled input
  double sample_rate_in = 0.0;
  int32_t num_points = 0, window = 0, zero_padding = 0;
  int64_t fft_size_cfg = 0;
  // derived sizes
  int64_t num_samples = 0;  // after resampling; 0 = empty spectrum
  int64_t fft_size = 0;     // N, the DFT length the reference runs
  int64_t run_size = 0;     // M, the power-of-two transform run on the device
  int64_t num_bins = 0;     // N / 2 + 1
  int log2_run = 0;
  bool bluestein = false;   // N is not a power of two: chirp-z through M >= 2N - 1
  double sample_rate = 0.0, freq_resolution = 0.0, norm = 0.0;
  // byte offsets of the tables in the handle's "sp_tables" workspace
  // (SIZE_MAX: table absent)
  size_t o_idx = 0, o_alpha = 0, o_oma = 0, o_win = 0, o_tw = 0, o_chirp = 0, o_bfft = 0;
};

}  // namespace spx
