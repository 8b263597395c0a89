// spectral.hip — batch spectral analysis of sweep waveforms on the device:
// windowed FFT spectra (compute_fft / compute_fft_from_samples,
// crates/spicier-solver/src/spectral/fft.rs:98-367), THD with its harmonic
// peak search (spectral/thd.rs:60-312), and the batch forms over a sweep
// (crates/spicier-batched-sweep/src/spectral.rs:46-330).
//
// One workgroup runs one waveform at a time and keeps it in LDS from the
// first load to the last output: gather (and linear resampling) of the probe
// column, window, a radix-4 Stockham DFT (a radix-2 first pass for odd
// log2 sizes) with twiddles from a table, the chirp-z (Bluestein) wrap for
// lengths that are not powers of two, then the single-sided spectrum, the
// harmonic search, and Welford accumulators for the per-bin ensemble
// statistics, held in registers across the waveforms a workgroup visits.
// HBM traffic is one read of each waveform plus whatever outputs are asked
// for; the transform itself is LDS-bound.
//
// The plan (resampling indices and weights, window, twiddles, chirp and the
// chirp's transform) depends only on the time grid and the configuration,
// and is built once per grid on the host and kept in the handle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "device_common.h"
#include "handle.h"

namespace spx {
namespace {

// Transforms of up to 8192 complex points run in one workgroup's LDS (128 KB
// for the vector, within gfx950's 160 KB per workgroup): real waveforms up
// to 16384 points, packed two samples per complex point, and chirp-z
// transforms up to 8192; longer ones (up to 2^22) take the four-step path
// through HBM.
constexpr int kMaxLog2 = 13;
constexpr int kMaxLog2Global = 22;
constexpr int kMaxHarmonics = 255;    // harmonic rows per waveform in LDS
constexpr int kStatsGrid = 4096;      // workgroups when per-bin statistics are accumulated
constexpr int kPlainGrid = 8192;      // workgroups otherwise (grid-stride over the batch)
constexpr size_t kAbsent = SIZE_MAX;

constexpr int threads_for(int log2m) {
  return log2m >= 13 ? 512 : (log2m >= 10 ? 256 : (log2m == 9 ? 128 : 64));
}

int check_handle(spx_handle* h) { return h ? SPX_OK : fail(SPX_ERR_INVALID_ARGUMENT, "null handle"); }

struct SpecArgs {
  const double* data;
  int64_t batch, bstride;
  int32_t pstride, num_samples, fft_size, num_bins, bluestein;
  const int32_t* idx;    // [num_samples] resampling: >= 0 interval start, < 0 exact point -idx-1
  const double* alpha;   // [num_samples] weight of the interval's right point
  const double* oma;     // [num_samples] 1 - alpha, as the reference rounds it
  const double* win;     // [num_samples] window weights, NULL = rectangular
  const double2* tw;     // [M] exp(-2 pi i m / M)
  const double2* chirp;  // [N] exp(-i pi m^2 / N)
  const double2* bfft;   // [M] transform of the conjugate chirp, wrapped
  double norm;           // samples * coherent gain
  double inv_norm;       // 1 / norm: bins are scaled by it (within an ulp of dividing)
  double *mag, *db, *phase, *psd;
  double* part;          // [gridDim.x][num_bins][5] Welford partials (count, mean, M2, min, max)
  int32_t thd, found, hstride;
  double f0, fr, search;  // fr: bin spacing; search: 2 x (frequencies[1] - frequencies[0])
  double *thd_pct, *thd_db, *fund_freq, *fund_mag, *hrms, *harm;
};

__device__ __forceinline__ double db_of(double m) { return m > 1e-20 ? 20.0 * log10(m) : -400.0; }

// In-LDS forward DFT of M = 2^LOG2M complex points (re[], im[]), natural
// order in and out.  Stockham autosort: pass with stride Ns reads the
// radix-R butterfly inputs j + r*M/R, applies W_M^(r*k*M/(R*Ns)) with
// k = j mod Ns, and writes the outputs at (j - k)*R + k + r*Ns.  Every pass
// reads all its inputs into registers before any thread writes.
//
// NT threads (tid = 0..NT-1) work on one vector; several vectors can share a
// workgroup, each with its own NT threads, since every vector runs the same
// passes and barriers.  tw holds W_L^m for some L = M * tws (tws >= 1).
template <int LOG2M, int NT>
__device__ __forceinline__ void fft_lds(double* __restrict__ re, double* __restrict__ im,
                                        const double2* __restrict__ tw, int tid, int tws) {
  constexpr int M = 1 << LOG2M;
  int ns = 1;
  if constexpr (LOG2M & 1) {
    constexpr int NB = M / 2, PER = (NB + NT - 1) / NT;
    double ar[PER], ai[PER], br[PER], bi[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int j = tid + p * NT;
      if (j < NB) {
        ar[p] = re[j]; ai[p] = im[j];
        br[p] = re[j + NB]; bi[p] = im[j + NB];
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int j = tid + p * NT;
      if (j < NB) {
        re[2 * j] = ar[p] + br[p];     im[2 * j] = ai[p] + bi[p];
        re[2 * j + 1] = ar[p] - br[p]; im[2 * j + 1] = ai[p] - bi[p];
      }
    }
    __syncthreads();
    ns = 2;
  }
  if constexpr (LOG2M >= 2) {
    constexpr int NB = M / 4, PER = (NB + NT - 1) / NT;
    for (; ns < M; ns <<= 2) {
      const int stride = M / (4 * ns);
      double vr[PER][4], vi[PER][4];
#pragma unroll
      for (int p = 0; p < PER; ++p) {
        const int j = tid + p * NT;
        if (j < NB) {
          const int k = j & (ns - 1);
          double xr[4], xi[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            xr[r] = re[j + r * NB];
            xi[r] = im[j + r * NB];
          }
#pragma unroll
          for (int r = 1; r < 4; ++r) {
            const double2 w = tw[(int64_t)(r * k * stride) * tws];
            const double a = xr[r], b = xi[r];
            xr[r] = a * w.x - b * w.y;
            xi[r] = a * w.y + b * w.x;
          }
          const double t0r = xr[0] + xr[2], t0i = xi[0] + xi[2];
          const double t1r = xr[0] - xr[2], t1i = xi[0] - xi[2];
          const double t2r = xr[1] + xr[3], t2i = xi[1] + xi[3];
          const double t3r = xr[1] - xr[3], t3i = xi[1] - xi[3];
          vr[p][0] = t0r + t2r; vi[p][0] = t0i + t2i;
          vr[p][2] = t0r - t2r; vi[p][2] = t0i - t2i;
          vr[p][1] = t1r + t3i; vi[p][1] = t1i - t3r;  // t1 - i t3
          vr[p][3] = t1r - t3i; vi[p][3] = t1i + t3r;  // t1 + i t3
        }
      }
      __syncthreads();
#pragma unroll
      for (int p = 0; p < PER; ++p) {
        const int j = tid + p * NT;
        if (j < NB) {
          const int k = j & (ns - 1);
          const int d = (j - k) * 4 + k;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            re[d + r * ns] = vr[p][r];
            im[d + r * ns] = vi[p][r];
          }
        }
      }
      __syncthreads();
    }
  }
}

// Transform of one complex vector in place (the plan's chirp filter).
template <int LOG2M>
__global__ __launch_bounds__(threads_for(LOG2M)) void k_fft_inplace(double2* v,
                                                                   const double2* tw) {
  constexpr int M = 1 << LOG2M, NT = threads_for(LOG2M);
  __shared__ double re[M], im[M];
  for (int s = threadIdx.x; s < M; s += NT) { re[s] = v[s].x; im[s] = v[s].y; }
  __syncthreads();
  fft_lds<LOG2M, NT>(re, im, tw, threadIdx.x, 1);
  for (int s = threadIdx.x; s < M; s += NT) v[s] = make_double2(re[s], im[s]);
}

// Bin k of a waveform's DFT, X_k, read from a transform in LDS or HBM.
// Direct: the transform is X itself.  Packed: a real waveform of even
// length N was transformed as the N/2 complex points z_n = x_2n + i x_2n+1;
// with W = Z_(N/2-k), E = (Z_k + conj W)/2 and O = (Z_k - conj W)/(2i),
// X_k = E + W_N^k O for 0 <= k <= N/2 (Z_(N/2) = Z_0).
struct DirectView {
  const double* re;
  const double* im;
  __device__ __forceinline__ void get(int k, double& xr, double& xi) const {
    xr = re[k];
    xi = im[k];
  }
};
struct PackedView {
  const double* re;
  const double* im;
  const double2* tw;  // W_N^m
  int half;           // N / 2
  __device__ __forceinline__ void get(int k, double& xr, double& xi) const {
    const int kk = k == half ? 0 : k, j = k == 0 ? 0 : half - k;
    const double zr = re[kk], zi = im[kk], wr = re[j], wi = im[j];
    const double er = 0.5 * (zr + wr), ei = 0.5 * (zi - wi);
    const double orr = 0.5 * (zi + wi), oi = 0.5 * (wr - zr);
    const double2 t = tw[k];
    xr = er + (orr * t.x - oi * t.y);
    xi = ei + (orr * t.y + oi * t.x);
  }
};

// The spectrum of bin k, as fft.rs:228-258 forms it.
struct Bin {
  double nr, ni, unscaled, scale;
};
template <class View>
__device__ __forceinline__ Bin bin_of(const View& v, int k, int N, int nbins, double inv_norm) {
  double xr, xi;
  v.get(k, xr, xi);
  Bin b;
  b.nr = xr * inv_norm;
  b.ni = xi * inv_norm;
  b.unscaled = hypot(b.nr, b.ni);
  b.scale = (k == 0 || ((N & 1) == 0 && k == nbins - 1)) ? 1.0 : 2.0;
  return b;
}

// Sample s of waveform w as compute_fft forms it: gathered (resampled,
// interpolate_at: v0*(1-alpha) + v1*alpha, unfused), windowed, and for
// chirp-z multiplied by the chirp.  Samples past num_samples are the zero
// padding.
__device__ __forceinline__ void load_sample(const SpecArgs& a, const double* w, int s, double& xr,
                                            double& xi) {
  xr = 0.0;
  xi = 0.0;
  if (s >= a.num_samples) return;
  double x;
  if (!a.idx) {
    x = w[(int64_t)s * a.pstride];
  } else {
    const int k = a.idx[s];
    x = k < 0 ? w[(int64_t)(-k - 1) * a.pstride]
              : __dadd_rn(__dmul_rn(w[(int64_t)k * a.pstride], a.oma[s]),
                          __dmul_rn(w[(int64_t)(k + 1) * a.pstride], a.alpha[s]));
  }
  if (a.win) x = __dmul_rn(x, a.win[s]);
  if (a.bluestein) {
    const double2 c = a.chirp[s];
    xr = x * c.x;
    xi = x * c.y;
  } else {
    xr = x;
  }
}

// Chirp-z, between the two transforms: X_k = w_k * (a (*) conj(w))_k is the
// product with the filter's transform, inverse-transformed as
// conj(F(conj(.)))/M, then multiplied by the output chirp.
__device__ __forceinline__ void chirp_mul(double& re, double& im, double2 f) {
  const double xr = re, xi = im;
  re = xr * f.x - xi * f.y;
  im = -(xr * f.y + xi * f.x);
}
__device__ __forceinline__ void chirp_out(double& re, double& im, double2 c, double inv_m) {
  const double yr = re * inv_m, yi = -im * inv_m;
  re = yr * c.x - yi * c.y;
  im = yr * c.y + yi * c.x;
}

// Single-sided spectrum outputs of bin k (fft.rs:228-259); returns the magnitude.
template <class View>
__device__ __forceinline__ double write_bin(const SpecArgs& a, const View& v, int64_t b, int k) {
  const Bin bn = bin_of(v, k, a.fft_size, a.num_bins, a.inv_norm);
  const double m = bn.unscaled * bn.scale;
  const int64_t o = b * a.num_bins + k;
  if (a.mag) a.mag[o] = m;
  if (a.db) a.db[o] = db_of(m);
  if (a.phase) a.phase[o] = atan2(bn.ni, bn.nr);
  if (a.psd) a.psd[o] = bn.unscaled * bn.unscaled * bn.scale;
  return m;
}

// THD of the spectrum in re/im (LDS or global) for waveform b; every
// thread of the workgroup calls it.  Harmonic g (g = 0: the fundamental) is
// searched by eight lanes around the target's bin: the peak is the largest
// magnitude > 0 within the search range (ties: lowest bin, thd.rs:132-153),
// else the nearest bin (thd.rs:101-127).  Both lie within [c - 3, c + 4],
// c the rounded (clamped) bin of the target.  Then thd.rs:212-311.
template <int NT, class View>
__device__ void thd_stage(const SpecArgs& a, const View& v, int64_t b, double* hf, double* hm,
                          double* hp) {
  const int tid = threadIdx.x, nb = a.num_bins;
  for (int g = tid >> 3; g <= a.found; g += NT >> 3) {
    const int lane8 = tid & 7;
    const double target = g == 0 ? a.f0 : a.f0 * (double)(g + 1);
    double cen = floor(target / a.fr + 0.5);
    cen = fmin(fmax(cen, 0.0), (double)(nb - 1));
    const int i = (int)cen - 3 + lane8;
    const bool inr = i >= 0 && i < nb;
    double fi = 0.0, m = 0.0, ph = 0.0;
    if (inr) {
      const Bin bn = bin_of(v, i, a.fft_size, nb, a.inv_norm);
      fi = (double)i * a.fr;
      m = bn.unscaled * bn.scale;
      ph = atan2(bn.ni, bn.nr);
    }
    const bool cand = inr && fabs(fi - target) <= a.search && m > 0.0;
    double key = cand ? m : -1.0;
    int ki = cand ? i : INT_MAX;
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
      const double ok = __shfl_xor(key, off, 8);
      const int oi = __shfl_xor(ki, off, 8);
      if (ok > key || (ok == key && oi < ki)) { key = ok; ki = oi; }
    }
    if (key < 0.0) {  // no peak: nearest bin
      double dk = inr ? fabs(fi - target) : INFINITY;
      ki = inr ? i : INT_MAX;
#pragma unroll
      for (int off = 1; off < 8; off <<= 1) {
        const double od = __shfl_xor(dk, off, 8);
        const int oi = __shfl_xor(ki, off, 8);
        if (od < dk || (od == dk && oi < ki)) { dk = od; ki = oi; }
      }
    }
    if (inr && i == ki) { hf[g] = fi; hm[g] = m; hp[g] = ph; }
  }
  __syncthreads();
  const double fund = hm[0];
  if (tid == 0) {
    double sq = 0.0;
    for (int h = 1; h <= a.found; ++h) sq += hm[h] * hm[h];
    const double rms = sqrt(sq);
    const double thd = fund > 1e-20 ? (rms / fund) * 100.0 : 0.0;
    if (a.thd_pct) a.thd_pct[b] = thd;
    if (a.thd_db) a.thd_db[b] = thd > 1e-20 ? 20.0 * log10(thd / 100.0) : -INFINITY;
    if (a.fund_freq) a.fund_freq[b] = hf[0];
    if (a.fund_mag) a.fund_mag[b] = fund;
    if (a.hrms) a.hrms[b] = rms;
  }
  if (a.harm) {
    for (int h = tid; h <= a.found; h += NT) {
      double* row = a.harm + (b * a.hstride + h) * 5;
      row[0] = hf[h];
      row[1] = hm[h];
      row[2] = db_of(hm[h]);
      row[3] = hp[h];
      row[4] = h == 0 ? 100.0 : (fund > 1e-20 ? (hm[h] / fund) * 100.0 : 0.0);
    }
  }
}

enum { kModeSpectra = 0, kModeStats = 1, kModeThd = 2 };

// One waveform per workgroup iteration, all in LDS.  PACKED: a real
// waveform of power-of-two length N = 2L as an L-point complex transform
// (twiddles W_N^2m from the length-N table), separated per bin; otherwise
// the L-point complex transform directly, with the chirp-z wrap when the
// plan needs it.  kModeStats also folds every magnitude into per-thread
// Welford accumulators for the bins the thread owns.
template <int LOG2L, int MODE, bool PACKED>
__global__ __launch_bounds__(threads_for(LOG2L)) void k_spectrum(SpecArgs a) {
  constexpr int L = 1 << LOG2L, NT = threads_for(LOG2L);
  constexpr int NBMAX = PACKED ? L + 1 : L / 2 + 1;
  constexpr int KB = (NBMAX + NT - 1) / NT;  // bins per thread
  constexpr int KS = MODE == kModeStats ? KB : 1;
  __shared__ double re[L], im[L];
  __shared__ double hf[MODE == kModeThd ? kMaxHarmonics + 1 : 1];
  __shared__ double hm[MODE == kModeThd ? kMaxHarmonics + 1 : 1];
  __shared__ double hp[MODE == kModeThd ? kMaxHarmonics + 1 : 1];
  const int tid = threadIdx.x;
  const int nb = a.num_bins;

  double wm[KS], w2[KS], wmin[KS], wmax[KS];
#pragma unroll
  for (int q = 0; q < KS; ++q) { wm[q] = 0.0; w2[q] = 0.0; wmin[q] = INFINITY; wmax[q] = -INFINITY; }
  double cnt = 0.0;

  for (int64_t b = blockIdx.x; b < a.batch; b += gridDim.x) {
    const double* w = a.data + b * a.bstride;
    for (int s = tid; s < L; s += NT) {
      if constexpr (PACKED) {
        double xr, xi, yr, yi;
        load_sample(a, w, 2 * s, xr, xi);
        load_sample(a, w, 2 * s + 1, yr, yi);
        re[s] = xr;
        im[s] = yr;
      } else {
        load_sample(a, w, s, re[s], im[s]);
      }
    }
    __syncthreads();
    fft_lds<LOG2L, NT>(re, im, a.tw, tid, PACKED ? 2 : 1);
    if (!PACKED && a.bluestein) {
      for (int s = tid; s < L; s += NT) chirp_mul(re[s], im[s], a.bfft[s]);
      __syncthreads();
      fft_lds<LOG2L, NT>(re, im, a.tw, tid, 1);
      for (int s = tid; s < a.fft_size; s += NT) chirp_out(re[s], im[s], a.chirp[s], 1.0 / (double)L);
      __syncthreads();
    }
    using View = typename std::conditional<PACKED, PackedView, DirectView>::type;
    View v;
    if constexpr (PACKED) v = PackedView{re, im, a.tw, L};
    else v = DirectView{re, im};
    if constexpr (MODE == kModeThd) {
      thd_stage<NT>(a, v, b, hf, hm, hp);
    } else if constexpr (MODE == kModeStats) {
      cnt += 1.0;
      const double inv = 1.0 / cnt;
#pragma unroll
      for (int q = 0; q < KB; ++q) {
        const int k = tid + q * NT;
        if (k < nb) {
          const double m = write_bin(a, v, b, k);
          const double d = m - wm[q];
          wm[q] += d * inv;
          w2[q] += d * (m - wm[q]);
          wmin[q] = fmin(wmin[q], m);
          wmax[q] = fmax(wmax[q], m);
        }
      }
    } else {
#pragma unroll 1
      for (int k = tid; k < nb; k += NT) (void)write_bin(a, v, b, k);
    }
    __syncthreads();  // LDS is refilled by the next waveform
  }
  if constexpr (MODE == kModeStats) {
#pragma unroll
    for (int q = 0; q < KB; ++q) {
      const int k = tid + q * NT;
      if (k < nb) {
        double* o = a.part + ((int64_t)blockIdx.x * nb + k) * 5;
        o[0] = cnt; o[1] = wm[q]; o[2] = w2[q]; o[3] = wmin[q]; o[4] = wmax[q];
      }
    }
  }
}

// ------------------------------------------------ transforms past 2^13 points
// Four-step: M = M1 * M2 with the input viewed as [n1][n2] (n = n1*M2 + n2).
// Column pass: length-M1 transforms down each column n2, then the twiddle
// W_M^(n2*k1); row pass: length-M2 transforms along each row k1, giving
// X[k1 + M1*k2].  Each pass stages V vectors per workgroup through LDS; the
// waveforms of a chunk live in a workspace as [chunk][re plane | im plane].
constexpr int kGlobalThreads = 256;
constexpr int vec_threads(int log2m) { return log2m >= 10 ? 256 : (1 << log2m) / 4; }

__global__ __launch_bounds__(kGlobalThreads) void k_gather(SpecArgs a, int64_t b0, int64_t M,
                                                           double* __restrict__ z) {
  const int64_t bl = blockIdx.y, b = b0 + bl;
  double* zr = z + bl * 2 * M;
  const double* w = a.data + b * a.bstride;
  for (int64_t s = (int64_t)blockIdx.x * kGlobalThreads + threadIdx.x; s < M;
       s += (int64_t)gridDim.x * kGlobalThreads) {
    double xr, xi;
    load_sample(a, w, (int)s, xr, xi);
    zr[s] = xr;
    zr[M + s] = xi;
  }
}

// COLS: column pass (reads [n1][n2], writes twiddled [k1][n2]); else row
// pass (reads rows [k1][n], writes X[k1 + M1*k2]).
template <int LOG2V, bool COLS>
__global__ __launch_bounds__(kGlobalThreads) void k_four_step(const double* __restrict__ in,
                                                              double* __restrict__ out,
                                                              const double2* __restrict__ tw,
                                                              int log2m) {
  constexpr int L = 1 << LOG2V, TV = vec_threads(LOG2V), V = kGlobalThreads / TV;
  __shared__ double re[V * L], im[V * L];
  const int64_t M = int64_t{1} << log2m;
  const int64_t other = M >> LOG2V;  // M2 for columns, M1 for rows
  const int64_t bl = blockIdx.y;
  const double* ir = in + bl * 2 * M;
  double* orr = out + bl * 2 * M;
  const int64_t v0 = (int64_t)blockIdx.x * V;  // first column n2 / row k1 of this workgroup
  const int tid = threadIdx.x;
  for (int e = tid; e < V * L; e += kGlobalThreads) {
    int v, j;
    int64_t src;
    if (COLS) { v = e % V; j = e / V; src = (int64_t)j * other + v0 + v; }
    else { v = e / L; j = e % L; src = (v0 + v) * L + j; }
    re[v * L + j] = ir[src];
    im[v * L + j] = ir[M + src];
  }
  __syncthreads();
  const int v = tid / TV;
  fft_lds<LOG2V, TV>(re + v * L, im + v * L, tw, tid % TV, (int)other);
  if (COLS) {
    for (int k1 = tid % TV; k1 < L; k1 += TV) {
      const double2 w = tw[(v0 + v) * k1];
      const double xr = re[v * L + k1], xi = im[v * L + k1];
      re[v * L + k1] = xr * w.x - xi * w.y;
      im[v * L + k1] = xr * w.y + xi * w.x;
    }
    __syncthreads();
  }
  for (int e = tid; e < V * L; e += kGlobalThreads) {
    const int vv = e % V, j = e / V;  // consecutive lanes: consecutive columns / rows
    const int64_t dst = COLS ? (int64_t)j * other + v0 + vv : v0 + vv + other * (int64_t)j;
    orr[dst] = re[vv * L + j];
    orr[M + dst] = im[vv * L + j];
  }
}

__global__ void k_chirp_mul_global(SpecArgs a, int64_t M, double* __restrict__ z) {
  const int64_t bl = blockIdx.y;
  double* zr = z + bl * 2 * M;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < M;
       s += (int64_t)gridDim.x * blockDim.x)
    chirp_mul(zr[s], zr[M + s], a.bfft[s]);
}

__global__ void k_chirp_out_global(SpecArgs a, int64_t M, double* __restrict__ z) {
  const int64_t bl = blockIdx.y;
  double* zr = z + bl * 2 * M;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < a.fft_size;
       s += (int64_t)gridDim.x * blockDim.x)
    chirp_out(zr[s], zr[M + s], a.chirp[s], 1.0 / (double)M);
}

// Spectrum outputs and THD of a chunk's transforms (one workgroup per waveform).
template <bool THD>
__global__ __launch_bounds__(kGlobalThreads) void k_post_global(SpecArgs a, int64_t b0,
                                                                int64_t M,
                                                                const double* __restrict__ z) {
  __shared__ double hf[THD ? kMaxHarmonics + 1 : 1], hm[THD ? kMaxHarmonics + 1 : 1],
      hp[THD ? kMaxHarmonics + 1 : 1];
  const int64_t bl = blockIdx.x, b = b0 + bl;
  const DirectView v{z + bl * 2 * M, z + bl * 2 * M + M};
  if constexpr (THD) {
    thd_stage<kGlobalThreads>(a, v, b, hf, hm, hp);
  } else {
    for (int k = threadIdx.x; k < a.num_bins; k += kGlobalThreads) (void)write_bin(a, v, b, k);
  }
}

// Per-bin Welford partials over magnitudes [batch][nb]: workgroup column g
// visits waveforms g, g + G, ... (the same split as k_spectrum's).
__global__ void k_bin_stats_cols(const double* __restrict__ mag, int64_t batch, int nb,
                                 double* __restrict__ part) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y, G = gridDim.y;
  if (k >= nb) return;
  double mean = 0.0, m2 = 0.0, mn = INFINITY, mx = -INFINITY, cnt = 0.0;
  for (int64_t b = g; b < batch; b += G) {
    const double m = mag[b * nb + k];
    cnt += 1.0;
    const double d = m - mean;
    mean += d / cnt;
    m2 += d * (m - mean);
    mn = fmin(mn, m);
    mx = fmax(mx, m);
  }
  double* o = part + ((int64_t)g * nb + k) * 5;
  o[0] = cnt; o[1] = mean; o[2] = m2; o[3] = mn; o[4] = mx;
}

// Per-bin merge of Welford partials [groups][nb][5] (count, mean, M2, min,
// max) in group order, Chan et al.'s pairwise update: each level folds runs
// of kMergeRun consecutive groups, until one group is left, which the last
// kernel turns into the mean and population deviation.  The fold order is
// fixed, so results do not depend on scheduling.
constexpr int kMergeRun = 16;

struct Welford {
  double n = 0.0, mean = 0.0, m2 = 0.0, mn = INFINITY, mx = -INFINITY;
  __device__ void fold(const double* p) {
    if (p[0] <= 0.0) return;
    const double c = p[0], tot = n + c, d = p[1] - mean;
    mean += d * (c / tot);
    m2 += p[2] + d * d * (n * c / tot);
    n = tot;
    mn = fmin(mn, p[3]);
    mx = fmax(mx, p[4]);
  }
};

__global__ void k_bin_stats_fold(const double* __restrict__ part, int groups, int nb,
                                 double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const int g0 = blockIdx.y * kMergeRun, g1 = min(groups, g0 + kMergeRun);
  Welford w;
  for (int g = g0; g < g1; ++g) w.fold(part + ((int64_t)g * nb + k) * 5);
  double* o = out + ((int64_t)blockIdx.y * nb + k) * 5;
  o[0] = w.n; o[1] = w.mean; o[2] = w.m2; o[3] = w.mn; o[4] = w.mx;
}

__global__ void k_bin_stats_finish(const double* __restrict__ part, int64_t batch, int nb,
                                   double* mean_o, double* std_o, double* min_o, double* max_o) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const double* p = part + (int64_t)k * 5;
  if (mean_o) mean_o[k] = p[1];
  if (std_o) std_o[k] = sqrt(p[2] / (double)batch);
  if (min_o) min_o[k] = p[3];
  if (max_o) max_o[k] = p[4];
}

template <typename T>
__global__ void k_fill_rows(T* p, int64_t count, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ThdStatistics (spectral.rs:117-159) in one workgroup, without sorting:
// sums in a fixed tree (the result does not depend on the launch), min and
// max by reduction, and the median's order statistics (n/2 - 1 and n/2, or
// n/2) by radix selection on order-preserving 64-bit keys (sign bit set for
// positives, all bits inverted for negatives: -0 ranks below +0, as in a
// radix sort of the values): eight passes of eight bits, each counting the
// candidates that still match the selected prefix into 256 LDS bins.
constexpr int kStatThreads = 1024;

__device__ __forceinline__ uint64_t order_key(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}
__device__ __forceinline__ double key_value(uint64_t k) {
  return __longlong_as_double((long long)((k >> 63) ? (k & ~(1ull << 63)) : ~k));
}

// The k-th smallest (0-based) of v[0..n); every thread returns it.
__device__ double select_kth(const double* __restrict__ v, int64_t n, int64_t k,
                             unsigned* hist, uint64_t* s_prefix, int64_t* s_k) {
  const int t = threadIdx.x;
  uint64_t prefix = 0, mask = 0;
  for (int shift = 56; shift >= 0; shift -= 8) {
    if (t < 256) hist[t] = 0;
    __syncthreads();
    for (int64_t i = t; i < n; i += kStatThreads) {
      const uint64_t key = order_key(v[i]);
      if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (t == 0) {
      int64_t below = 0;
      int d = 0;
      for (; d < 255; ++d) {
        if (below + (int64_t)hist[d] > k) break;
        below += hist[d];
      }
      *s_prefix = prefix | ((uint64_t)d << shift);
      *s_k = k - below;
    }
    __syncthreads();
    prefix = *s_prefix;
    k = *s_k;
    mask |= 0xFFull << shift;
    __syncthreads();
  }
  return key_value(prefix);
}

__global__ __launch_bounds__(kStatThreads) void k_thd_stats(const double* __restrict__ v,
                                                            int64_t n, double spec,
                                                            double* __restrict__ out) {
  __shared__ double s0[kStatThreads], s1[kStatThreads], s2[kStatThreads], s3[kStatThreads];
  __shared__ unsigned hist[256];
  __shared__ uint64_t s_prefix;
  __shared__ int64_t s_k;
  const int t = threadIdx.x;
  double sum = 0.0, pass = 0.0, nan = 0.0, lo = INFINITY, hi = -INFINITY;
  for (int64_t i = t; i < n; i += kStatThreads) {
    const double x = v[i];
    sum += x;
    pass += x <= spec ? 1.0 : 0.0;
    nan += x != x ? 1.0 : 0.0;
    lo = fmin(lo, x);
    hi = fmax(hi, x);
  }
  s0[t] = sum;
  s1[t] = pass;
  s2[t] = lo;
  s3[t] = hi;
  __syncthreads();
  for (int w = kStatThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
      s0[t] += s0[t + w];
      s1[t] += s1[t + w];
      s2[t] = fmin(s2[t], s2[t + w]);
      s3[t] = fmax(s3[t], s3[t + w]);
    }
    __syncthreads();
  }
  const double mean = s0[0] / (double)n, passing = s1[0], vmin = s2[0], vmax = s3[0];
  __syncthreads();
  double var = 0.0;
  for (int64_t i = t; i < n; i += kStatThreads) {
    const double d = v[i] - mean;
    var += d * d;
  }
  s0[t] = var;
  s1[t] = nan;
  __syncthreads();
  for (int w = kStatThreads / 2; w > 0; w >>= 1) {
    if (t < w) { s0[t] += s0[t + w]; s1[t] += s1[t + w]; }
    __syncthreads();
  }
  const double sq = s0[0], nans = s1[0];
  __syncthreads();
  double med = 0.0;
  if (nans == 0.0) {  // NaNs have no order; the host reports them
    const double upper = select_kth(v, n, n / 2, hist, &s_prefix, &s_k);
    med = (n % 2 == 0) ? (select_kth(v, n, n / 2 - 1, hist, &s_prefix, &s_k) + upper) / 2.0
                       : upper;
  }
  if (t == 0) {
    out[0] = mean;
    out[1] = sqrt(sq / (double)n);
    out[2] = vmin;
    out[3] = vmax;
    out[4] = med;
    out[5] = (passing / (double)n) * 100.0;
    out[6] = nans;
  }
}

// ---------------------------------------------------------------- host plan
double window_weight(int window, int64_t i, int64_t n) {
#pragma clang fp contract(off)
  const double phase = 2.0 * M_PI * (double)i / (double)(n - 1);
  switch (window) {
    case SPX_WINDOW_HANNING: return 0.5 * (1.0 - std::cos(phase));
    case SPX_WINDOW_HAMMING: return 0.54 - 0.46 * std::cos(phase);
    case SPX_WINDOW_BLACKMAN: return 0.42 - 0.5 * std::cos(phase) + 0.08 * std::cos(2.0 * phase);
    default: return 1.0;
  }
}

double coherent_gain(int window, int64_t n) {
#pragma clang fp contract(off)
  if (n == 0) return 0.0;
  if (window == SPX_WINDOW_RECTANGULAR) return 1.0;
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += window_weight(window, i, n);
  return s / (double)n;
}

int64_t next_pow2(int64_t n) {
  int64_t p = 1;
  while (p < n) p *= 2;
  return p;
}

bool same_key(const SpectralPlan& p, const spx_waveform_batch& wf, const spx_spectral_config& c) {
  if (!p.valid || p.num_points != wf.num_points || p.window != c.window ||
      p.zero_padding != (c.zero_padding != 0) || p.fft_size_cfg != c.fft_size)
    return false;
  if (!wf.times) return p.times.empty() && p.sample_rate_in == wf.sample_rate;
  return (int64_t)p.times.size() == wf.num_points &&
         std::memcmp(p.times.data(), wf.times, sizeof(double) * wf.num_points) == 0;
}

int check_config(const spx_spectral_config& c) {
  if (c.window < SPX_WINDOW_RECTANGULAR || c.window > SPX_WINDOW_BLACKMAN)
    return fail(SPX_ERR_INVALID_ARGUMENT, "unknown window " + std::to_string(c.window));
  if (c.fft_size < 0) return fail(SPX_ERR_INVALID_ARGUMENT, "negative fft_size");
  return SPX_OK;
}

int check_batch(const spx_waveform_batch* wf) {
  if (!wf) return fail(SPX_ERR_INVALID_ARGUMENT, "null waveform batch");
  if (wf->batch < 0 || wf->num_points < 0 || wf->point_stride < 1 || wf->batch_stride < 0)
    return fail(SPX_ERR_INVALID_DIMENSION, "bad waveform batch shape");
  if (wf->batch > 0 && wf->num_points > 0 && !wf->data)
    return fail(SPX_ERR_INVALID_ARGUMENT, "null waveform data");
  if (wf->times) {
    for (int32_t i = 0; i < wf->num_points; ++i) {
      if (!std::isfinite(wf->times[i]))
        return fail(SPX_ERR_INVALID_ARGUMENT, "non-finite time at point " + std::to_string(i));
      if (i > 0 && !(wf->times[i] > wf->times[i - 1]))
        return fail(SPX_ERR_INVALID_ARGUMENT,
                    "times must be strictly increasing (point " + std::to_string(i) + ")");
    }
  }
  return SPX_OK;
}

// Sizes of the spectrum (fft.rs:98-131, 191-207, 296-311).
void plan_sizes(const spx_waveform_batch& wf, const spx_spectral_config& c, SpectralPlan& p) {
#pragma clang fp contract(off)
  p.num_samples = 0;
  p.sample_rate = 0.0;
  if (wf.times) {
    if (wf.num_points > 0) {
      const double duration = wf.times[wf.num_points - 1] - wf.times[0];
      if (duration > 0.0) {
        const int64_t n = wf.num_points;
        p.sample_rate = n > 1 ? (double)(n - 1) / duration : 1.0;
        p.num_samples = (int64_t)std::ceil(duration * p.sample_rate) + 1;
      }
    }
  } else if (wf.num_points > 0 && wf.sample_rate > 0.0) {
    p.num_samples = wf.num_points;
    p.sample_rate = wf.sample_rate;
  }
  if (p.num_samples == 0) {
    p.sample_rate = 0.0;
    p.fft_size = p.num_bins = p.run_size = 0;
    return;
  }
  const int64_t want = c.fft_size > 0 ? c.fft_size : next_pow2(p.num_samples);
  p.fft_size = (c.zero_padding && p.num_samples < want) ? want : p.num_samples;
  p.num_bins = p.fft_size / 2 + 1;
  p.freq_resolution = p.sample_rate / (double)p.fft_size;
  p.bluestein = next_pow2(p.fft_size) != p.fft_size;
  p.run_size = p.bluestein ? next_pow2(2 * p.fft_size - 1) : p.fft_size;
  p.log2_run = 0;
  while ((int64_t{1} << p.log2_run) < p.run_size) ++p.log2_run;
}

void launch_fft_inplace(int log2m, double2* v, const double2* tw, hipStream_t st) {
  switch (log2m) {
#define SPX_FFT_CASE(L) \
  case L: hipLaunchKernelGGL(k_fft_inplace<L>, dim3(1), dim3(threads_for(L)), 0, st, v, tw); break;
    SPX_FFT_CASE(0) SPX_FFT_CASE(1) SPX_FFT_CASE(2) SPX_FFT_CASE(3) SPX_FFT_CASE(4)
    SPX_FFT_CASE(5) SPX_FFT_CASE(6) SPX_FFT_CASE(7) SPX_FFT_CASE(8) SPX_FFT_CASE(9)
    SPX_FFT_CASE(10) SPX_FFT_CASE(11) SPX_FFT_CASE(12) SPX_FFT_CASE(13)
#undef SPX_FFT_CASE
  }
}

template <int MODE, bool PACKED>
void launch_spectrum_mode(int log2l, unsigned grid, const SpecArgs& a, hipStream_t st) {
  switch (log2l) {
#define SPX_SPEC_CASE(L)                                                                    \
  case L:                                                                                   \
    hipLaunchKernelGGL((k_spectrum<L, MODE, PACKED>), dim3(grid), dim3(threads_for(L)), 0, \
                       st, a);                                                              \
    break;
    SPX_SPEC_CASE(0) SPX_SPEC_CASE(1) SPX_SPEC_CASE(2) SPX_SPEC_CASE(3) SPX_SPEC_CASE(4)
    SPX_SPEC_CASE(5) SPX_SPEC_CASE(6) SPX_SPEC_CASE(7) SPX_SPEC_CASE(8) SPX_SPEC_CASE(9)
    SPX_SPEC_CASE(10) SPX_SPEC_CASE(11) SPX_SPEC_CASE(12) SPX_SPEC_CASE(13)
#undef SPX_SPEC_CASE
  }
}

// Power-of-two lengths from 4 up to 2^14 points run packed (half-length
// transform); the chirp-z plans and N <= 2 run the complex transform up to
// 2^13 points; everything else takes the four-step path.
bool packed_path(const SpectralPlan& p) {
  return !p.bluestein && p.fft_size >= 4 && p.log2_run - 1 <= kMaxLog2;
}
bool lds_path(const SpectralPlan& p) { return packed_path(p) || p.log2_run <= kMaxLog2; }

template <int MODE>
void launch_spectrum_m(const SpectralPlan& p, unsigned grid, const SpecArgs& a, hipStream_t st) {
  if (packed_path(p)) launch_spectrum_mode<MODE, true>(p.log2_run - 1, grid, a, st);
  else launch_spectrum_mode<MODE, false>(p.log2_run, grid, a, st);
}

void launch_spectrum(int mode, const SpectralPlan& p, unsigned grid, const SpecArgs& a,
                     hipStream_t st) {
  if (mode == kModeStats) launch_spectrum_m<kModeStats>(p, grid, a, st);
  else if (mode == kModeThd) launch_spectrum_m<kModeThd>(p, grid, a, st);
  else launch_spectrum_m<kModeSpectra>(p, grid, a, st);
}

template <int LOG2V, bool COLS>
void launch_pass(const double* in, double* out, const double2* tw, int log2m, int64_t count,
                 hipStream_t st) {
  constexpr int V = kGlobalThreads / vec_threads(LOG2V);
  const int64_t vectors = int64_t{1} << (log2m - LOG2V);
  hipLaunchKernelGGL((k_four_step<LOG2V, COLS>), dim3((unsigned)(vectors / V), (unsigned)count),
                     dim3(kGlobalThreads), 0, st, in, out, tw, log2m);
}

template <bool COLS>
void launch_pass_n(int log2v, const double* in, double* out, const double2* tw, int log2m,
                   int64_t count, hipStream_t st) {
  switch (log2v) {
    case 7: launch_pass<7, COLS>(in, out, tw, log2m, count, st); break;
    case 8: launch_pass<8, COLS>(in, out, tw, log2m, count, st); break;
    case 9: launch_pass<9, COLS>(in, out, tw, log2m, count, st); break;
    case 10: launch_pass<10, COLS>(in, out, tw, log2m, count, st); break;
    case 11: launch_pass<11, COLS>(in, out, tw, log2m, count, st); break;
  }
}

// z (count vectors as re|im planes of 2^log2m) -> its transform, in natural
// order, back in z; y is scratch of the same size.
void four_step(double* z, double* y, const double2* tw, int log2m, int64_t count,
               hipStream_t st) {
  const int l1 = log2m / 2, l2 = log2m - l1;
  launch_pass_n<true>(l1, z, y, tw, log2m, count, st);
  launch_pass_n<false>(l2, y, z, tw, log2m, count, st);
}

__global__ void k_planes_to_d2(const double* __restrict__ z, int64_t M, double2* __restrict__ out) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < M;
       s += (int64_t)gridDim.x * blockDim.x)
    out[s] = make_double2(z[s], z[M + s]);
}

__global__ void k_d2_to_planes(const double2* __restrict__ in, int64_t M, double* __restrict__ z) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < M;
       s += (int64_t)gridDim.x * blockDim.x) {
    z[s] = in[s].x;
    z[M + s] = in[s].y;
  }
}

// Waveforms per chunk of the four-step path: each of its two scratch buffers
// stays within 1 GiB.
int64_t global_chunk(int64_t M, int64_t batch) {
  return std::max<int64_t>(1, std::min<int64_t>(batch, (int64_t{1} << 30) / (16 * M)));
}

// Builds (or reuses) the handle's plan for this grid and configuration.
// Caller holds the handle's mutex and has checked batch and config.
int get_plan(spx_handle* h, const spx_waveform_batch& wf, const spx_spectral_config& c,
             SpectralPlan** out) {
#pragma clang fp contract(off)
  SpectralPlan& p = h->spectral;
  if (same_key(p, wf, c)) {
    *out = &p;
    return SPX_OK;
  }
  p = SpectralPlan{};
  p.num_points = wf.num_points;
  p.window = c.window;
  p.zero_padding = c.zero_padding != 0;
  p.fft_size_cfg = c.fft_size;
  if (wf.times) p.times.assign(wf.times, wf.times + wf.num_points);
  p.sample_rate_in = wf.times ? 0.0 : wf.sample_rate;
  plan_sizes(wf, c, p);
  if (p.num_samples == 0) {
    p.valid = true;
    *out = &p;
    return SPX_OK;
  }
  if (p.num_samples > INT32_MAX)
    return fail(SPX_ERR_INVALID_DIMENSION, "too many samples after resampling");
  if (p.log2_run > kMaxLog2Global)
    return fail(SPX_ERR_INVALID_DIMENSION,
                "FFT length " + std::to_string(p.fft_size) +
                    " is beyond this library's limit (powers of two up to 2^22, other lengths "
                    "up to 2^21)");
  const int64_t ns = p.num_samples, N = p.fft_size, M = p.run_size;
  // resampling table (resample_uniform + interpolate_at)
  std::vector<int32_t> idx;
  std::vector<double> alpha, oma;
  if (wf.times) {
    const double* t = wf.times;
    const int64_t T = wf.num_points;
    const double duration = t[T - 1] - t[0];
    const double dt = ns > 1 ? duration / (double)(ns - 1) : duration;
    idx.resize(ns);
    alpha.resize(ns);
    oma.resize(ns);
    for (int64_t i = 0; i < ns; ++i) {
      const double ti = t[0] + (double)i * dt;
      if (ti <= t[0]) {
        idx[i] = -1;
      } else if (ti >= t[T - 1]) {
        idx[i] = -(int32_t)T;
      } else {
        const int64_t j = std::lower_bound(t + 1, t + T, ti) - t;  // first t[j] >= ti, j >= 1
        const double a = (ti - t[j - 1]) / (t[j] - t[j - 1]);
        idx[i] = (int32_t)(j - 1);
        alpha[i] = a;
        oma[i] = 1.0 - a;
      }
    }
  }
  std::vector<double> win;
  if (c.window != SPX_WINDOW_RECTANGULAR) {
    win.resize(ns);
    for (int64_t i = 0; i < ns; ++i) win[i] = window_weight(c.window, i, ns);
  }
  p.norm = (double)ns * coherent_gain(c.window, ns);
  // twiddles W_M^m and, for chirp-z, w_m = exp(-i pi (m^2 mod 2N) / N) and
  // the wrapped conjugate chirp; long double arguments keep them to an ulp
  const long double pi_l = 3.141592653589793238462643383279502884L;
  std::vector<double2> tw(M), chirp, bvec;
  for (int64_t m = 0; m < M; ++m) {
    const long double ang = -2.0L * pi_l * (long double)m / (long double)M;
    tw[m] = make_double2((double)cosl(ang), (double)sinl(ang));
  }
  if (p.bluestein) {
    chirp.resize(N);
    bvec.assign(M, make_double2(0.0, 0.0));
    for (int64_t m = 0; m < N; ++m) {
      const int64_t r = (m * m) % (2 * N);
      const long double ang = -pi_l * (long double)r / (long double)N;
      chirp[m] = make_double2((double)cosl(ang), (double)sinl(ang));
      const double2 cc = make_double2(chirp[m].x, -chirp[m].y);
      bvec[m] = cc;
      if (m > 0) bvec[M - m] = cc;
    }
  }
  // one workspace: [idx][alpha][oma][win][tw][chirp][bfft], 256-byte aligned pieces
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  p.o_idx = idx.empty() ? kAbsent : take(sizeof(int32_t) * ns);
  p.o_alpha = idx.empty() ? kAbsent : take(sizeof(double) * ns);
  p.o_oma = idx.empty() ? kAbsent : take(sizeof(double) * ns);
  p.o_win = win.empty() ? kAbsent : take(sizeof(double) * ns);
  p.o_tw = take(sizeof(double2) * M);
  p.o_chirp = p.bluestein ? take(sizeof(double2) * N) : kAbsent;
  p.o_bfft = p.bluestein ? take(sizeof(double2) * M) : kAbsent;
  void* ws = nullptr;
  if (int rc = workspace(h, "sp_tables", off, &ws)) return rc;
  char* base = static_cast<char*>(ws);
  auto up = [&](size_t o, const void* src, size_t bytes) -> int {
    if (o == kAbsent || bytes == 0) return SPX_OK;
    SPX_HIP(hipMemcpyAsync(base + o, src, bytes, hipMemcpyHostToDevice, h->stream),
            SPX_ERR_TRANSFER);
    return SPX_OK;
  };
  if (int rc = up(p.o_idx, idx.data(), sizeof(int32_t) * idx.size())) return rc;
  if (int rc = up(p.o_alpha, alpha.data(), sizeof(double) * alpha.size())) return rc;
  if (int rc = up(p.o_oma, oma.data(), sizeof(double) * oma.size())) return rc;
  if (int rc = up(p.o_win, win.data(), sizeof(double) * win.size())) return rc;
  if (int rc = up(p.o_tw, tw.data(), sizeof(double2) * tw.size())) return rc;
  if (int rc = up(p.o_chirp, chirp.data(), sizeof(double2) * chirp.size())) return rc;
  if (int rc = up(p.o_bfft, bvec.data(), sizeof(double2) * bvec.size())) return rc;
  if (p.bluestein) {
    auto* bf = reinterpret_cast<double2*>(base + p.o_bfft);
    const auto* twd = reinterpret_cast<const double2*>(base + p.o_tw);
    if (p.log2_run <= kMaxLog2) {
      launch_fft_inplace(p.log2_run, bf, twd, h->stream);
    } else {
      void *z = nullptr, *y = nullptr;
      if (int rc = workspace(h, "sp_z", sizeof(double) * 2 * M, &z)) return rc;
      if (int rc = workspace(h, "sp_y", sizeof(double) * 2 * M, &y)) return rc;
      const unsigned g = (unsigned)std::min<int64_t>((M + 255) / 256, 4096);
      hipLaunchKernelGGL(k_d2_to_planes, dim3(g), dim3(256), 0, h->stream, bf, M,
                         static_cast<double*>(z));
      four_step(static_cast<double*>(z), static_cast<double*>(y), twd, p.log2_run, 1, h->stream);
      hipLaunchKernelGGL(k_planes_to_d2, dim3(g), dim3(256), 0, h->stream,
                         static_cast<const double*>(z), M, bf);
    }
    SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  }
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);  // host tables go out of scope
  p.valid = true;
  *out = &p;
  return SPX_OK;
}

SpecArgs base_args(spx_handle* h, const spx_waveform_batch& wf, const SpectralPlan& p) {
  char* base = static_cast<char*>(h->ws["sp_tables"].first);
  auto at = [&](size_t o) -> void* { return o == kAbsent ? nullptr : base + o; };
  SpecArgs a{};
  a.data = wf.data;
  a.batch = wf.batch;
  a.bstride = wf.batch_stride;
  a.pstride = wf.point_stride;
  a.num_samples = (int32_t)p.num_samples;
  a.fft_size = (int32_t)p.fft_size;
  a.num_bins = (int32_t)p.num_bins;
  a.bluestein = p.bluestein ? 1 : 0;
  a.idx = static_cast<const int32_t*>(at(p.o_idx));
  a.alpha = static_cast<const double*>(at(p.o_alpha));
  a.oma = static_cast<const double*>(at(p.o_oma));
  a.win = static_cast<const double*>(at(p.o_win));
  a.tw = static_cast<const double2*>(at(p.o_tw));
  a.chirp = static_cast<const double2*>(at(p.o_chirp));
  a.bfft = static_cast<const double2*>(at(p.o_bfft));
  a.norm = p.norm;
  a.inv_norm = 1.0 / p.norm;
  return a;
}

void fill_info(const SpectralPlan& p, spx_spectrum_info* info) {
  if (!info) return;
  info->num_samples = p.num_samples;
  info->fft_size = p.fft_size;
  info->num_bins = p.num_bins;
  info->sample_rate = p.sample_rate;
  info->freq_resolution = p.num_samples ? p.freq_resolution : 0.0;
}

unsigned grid_of(int64_t batch, int cap) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(batch, cap));
}

// Transforms past 2^13 points: the batch in chunks through two HBM scratch
// buffers (gather, four-step, chirp-z if needed, then the spectrum or THD
// outputs per waveform).
int run_global(spx_handle* h, const SpectralPlan& p, const SpecArgs& a, bool thd) {
  const int64_t M = p.run_size, B = a.batch;
  const int64_t chunk = global_chunk(M, B);
  void *zv = nullptr, *yv = nullptr;
  if (int rc = workspace(h, "sp_z", sizeof(double) * 2 * M * chunk, &zv)) return rc;
  if (int rc = workspace(h, "sp_y", sizeof(double) * 2 * M * chunk, &yv)) return rc;
  auto* z = static_cast<double*>(zv);
  auto* y = static_cast<double*>(yv);
  const unsigned gx = (unsigned)std::min<int64_t>((M + kGlobalThreads - 1) / kGlobalThreads, 1024);
  for (int64_t b0 = 0; b0 < B; b0 += chunk) {
    const int64_t cb = std::min(chunk, B - b0);
    hipLaunchKernelGGL(k_gather, dim3(gx, (unsigned)cb), dim3(kGlobalThreads), 0, h->stream, a,
                       b0, M, z);
    four_step(z, y, a.tw, p.log2_run, cb, h->stream);
    if (p.bluestein) {
      hipLaunchKernelGGL(k_chirp_mul_global, dim3(gx, (unsigned)cb), dim3(kGlobalThreads), 0,
                         h->stream, a, M, z);
      four_step(z, y, a.tw, p.log2_run, cb, h->stream);
      hipLaunchKernelGGL(k_chirp_out_global, dim3(gx, (unsigned)cb), dim3(kGlobalThreads), 0,
                         h->stream, a, M, z);
    }
    if (thd)
      hipLaunchKernelGGL(k_post_global<true>, dim3((unsigned)cb), dim3(kGlobalThreads), 0,
                         h->stream, a, b0, M, static_cast<const double*>(z));
    else
      hipLaunchKernelGGL(k_post_global<false>, dim3((unsigned)cb), dim3(kGlobalThreads), 0,
                         h->stream, a, b0, M, static_cast<const double*>(z));
    SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  }
  return SPX_OK;
}

}  // namespace
}  // namespace spx

using namespace spx;

extern "C" {

void spx_spectral_config_default(spx_spectral_config* cfg) {
  if (!cfg) return;
  cfg->window = SPX_WINDOW_HANNING;
  cfg->zero_padding = 1;
  cfg->fft_size = 0;
}

int spx_window_coefficients(int32_t window, int64_t n, double* out) {
  if (window < SPX_WINDOW_RECTANGULAR || window > SPX_WINDOW_BLACKMAN)
    return fail(SPX_ERR_INVALID_ARGUMENT, "unknown window " + std::to_string(window));
  if (n < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative length");
  if (n > 0 && !out) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  for (int64_t i = 0; i < n; ++i) out[i] = window_weight(window, i, n);
  return SPX_OK;
}

double spx_window_coherent_gain(int32_t window, int64_t n) {
  if (window < SPX_WINDOW_RECTANGULAR || window > SPX_WINDOW_BLACKMAN || n < 0) return NAN;
  return coherent_gain(window, n);
}

int spx_spectrum_info_query(spx_handle* h, const spx_waveform_batch* wf,
                            const spx_spectral_config* cfg, spx_spectrum_info* info) {
  if (int rc = check_handle(h)) return rc;
  if (int rc = check_batch(wf)) return rc;
  spx_spectral_config c;
  spx_spectral_config_default(&c);
  if (cfg) c = *cfg;
  if (int rc = check_config(c)) return rc;
  if (!info) return fail(SPX_ERR_INVALID_ARGUMENT, "null info");
  SpectralPlan p;
  plan_sizes(*wf, c, p);
  fill_info(p, info);
  return SPX_OK;
}

int spx_batch_fft_device(spx_handle* h, const spx_waveform_batch* wf,
                         const spx_spectral_config* cfg, const spx_spectrum_out* out,
                         spx_spectrum_info* info) {
  if (int rc = check_handle(h)) return rc;
  if (int rc = check_batch(wf)) return rc;
  spx_spectral_config c;
  spx_spectral_config_default(&c);
  if (cfg) c = *cfg;
  if (int rc = check_config(c)) return rc;
  if (!out) return fail(SPX_ERR_INVALID_ARGUMENT, "null output");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  SpectralPlan* p = nullptr;
  if (int rc = get_plan(h, *wf, c, &p)) return rc;
  fill_info(*p, info);
  if (p->num_samples == 0 || wf->batch == 0) return SPX_OK;
  SpecArgs a = base_args(h, *wf, *p);
  a.mag = out->magnitude;
  a.db = out->magnitude_db;
  a.phase = out->phase;
  a.psd = out->psd;
  const bool stats =
      out->mean_magnitude || out->std_magnitude || out->min_magnitude || out->max_magnitude;
  const int64_t B = wf->batch;
  const int nb = (int)p->num_bins;
  const unsigned sgrid = grid_of(B, kStatsGrid);
  void* part = nullptr;
  void* part2 = nullptr;
  if (stats) {
    const int64_t g2 = (sgrid + kMergeRun - 1) / kMergeRun;
    if (int rc = workspace(h, "sp_part", sizeof(double) * 5 * sgrid * nb, &part)) return rc;
    if (int rc = workspace(h, "sp_part2", sizeof(double) * 5 * g2 * nb, &part2)) return rc;
  }
  if (lds_path(*p)) {
    if (stats) a.part = static_cast<double*>(part);
    launch_spectrum(stats ? kModeStats : kModeSpectra, *p, stats ? sgrid : grid_of(B, kPlainGrid),
                    a, h->stream);
  } else {
    if (stats && !a.mag) {  // the statistics read every magnitude back
      void* mags = nullptr;
      if (int rc = workspace(h, "sp_mags", sizeof(double) * B * nb, &mags)) return rc;
      a.mag = static_cast<double*>(mags);
    }
    if (int rc = run_global(h, *p, a, false)) return rc;
    if (stats)
      hipLaunchKernelGGL(k_bin_stats_cols, dim3((nb + 255) / 256, sgrid), dim3(256), 0, h->stream,
                         a.mag, B, nb, static_cast<double*>(part));
  }
  if (stats) {
    // fold levels alternate between the two partial buffers; the second is
    // never smaller than a level's output
    double* src = static_cast<double*>(part);
    double* dst = static_cast<double*>(part2);
    int groups = (int)sgrid;
    while (groups > 1) {
      const int next = (groups + kMergeRun - 1) / kMergeRun;
      hipLaunchKernelGGL(k_bin_stats_fold, dim3((nb + 255) / 256, next), dim3(256), 0, h->stream,
                         static_cast<const double*>(src), groups, nb, dst);
      std::swap(src, dst);
      groups = next;
    }
    hipLaunchKernelGGL(k_bin_stats_finish, dim3((nb + 255) / 256), dim3(256), 0, h->stream,
                       static_cast<const double*>(src), B, nb, out->mean_magnitude,
                       out->std_magnitude, out->min_magnitude, out->max_magnitude);
  }
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_batch_thd_device(spx_handle* h, const spx_waveform_batch* wf, double fundamental_freq,
                         int32_t num_harmonics, const spx_thd_out* out, int32_t* found_harmonics) {
  if (int rc = check_handle(h)) return rc;
  if (int rc = check_batch(wf)) return rc;
  if (!out) return fail(SPX_ERR_INVALID_ARGUMENT, "null output");
  if (num_harmonics < 0 || num_harmonics > kMaxHarmonics)
    return fail(SPX_ERR_INVALID_ARGUMENT, "num_harmonics must be within 0..255");
  if (std::isnan(fundamental_freq)) return fail(SPX_ERR_INVALID_ARGUMENT, "NaN fundamental");
  spx_spectral_config c;
  spx_spectral_config_default(&c);  // compute_thd's configuration (thd.rs:66-71)
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  SpectralPlan* p = nullptr;
  if (int rc = get_plan(h, *wf, c, &p)) return rc;
  const int64_t B = wf->batch;
  if (p->num_samples == 0 || !(fundamental_freq > 0.0)) {
    // the empty ThdResult (thd.rs:161-171)
    if (found_harmonics) *found_harmonics = 0;
    if (B == 0) return SPX_OK;
    const unsigned g = (unsigned)std::min<int64_t>((B + 255) / 256, 65536);
    auto fill = [&](double* d, double v) {
      if (d) hipLaunchKernelGGL(k_fill_rows<double>, dim3(g), dim3(256), 0, h->stream, d, B, v);
    };
    fill(out->thd_percent, 0.0);
    fill(out->thd_db, -INFINITY);
    fill(out->fundamental_freq, fundamental_freq);
    fill(out->fundamental_magnitude, 0.0);
    fill(out->harmonic_rms, 0.0);
    SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
    return SPX_OK;
  }
  // harmonics at or below Nyquist (thd.rs:230-236); the same for every waveform
  int found = 0;
  for (int n = 2; n <= num_harmonics + 1; ++n) {
    if (fundamental_freq * (double)n > p->sample_rate / 2.0) break;
    ++found;
  }
  if (found_harmonics) *found_harmonics = found;
  if (B == 0) return SPX_OK;
  SpecArgs a = base_args(h, *wf, *p);
  a.thd = 1;
  a.found = found;
  a.hstride = num_harmonics + 1;
  a.f0 = fundamental_freq;
  a.fr = p->freq_resolution;
  // thd.rs:174-179: frequencies[1] - frequencies[0] is the bin spacing itself
  a.search = (p->num_bins > 1 ? p->freq_resolution : fundamental_freq * 0.01) * 2.0;
  a.thd_pct = out->thd_percent;
  a.thd_db = out->thd_db;
  a.fund_freq = out->fundamental_freq;
  a.fund_mag = out->fundamental_magnitude;
  a.hrms = out->harmonic_rms;
  a.harm = out->harmonics;
  if (lds_path(*p)) {
    launch_spectrum(kModeThd, *p, grid_of(B, kPlainGrid), a, h->stream);
  } else if (int rc = run_global(h, *p, a, true)) {
    return rc;
  }
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_thd_statistics_device(spx_handle* h, const double* d_thd_percent, int64_t count,
                              double thd_spec, spx_thd_statistics* stats) {
  if (int rc = check_handle(h)) return rc;
  if (count < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative count");
  if (!stats || (count > 0 && !d_thd_percent)) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  if (count > INT32_MAX) return fail(SPX_ERR_BATCH_TOO_LARGE, "too many THD values (32-bit selection counts)");
  *stats = spx_thd_statistics{};
  stats->thd_spec = thd_spec;
  if (count == 0) {  // spectral.rs:104-115
    stats->yield_percent = 100.0;
    return SPX_OK;
  }
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  void* res = nullptr;
  if (int rc = workspace(h, "sp_stats", 8 * sizeof(double), &res)) return rc;
  hipLaunchKernelGGL(k_thd_stats, dim3(1), dim3(kStatThreads), 0, h->stream, d_thd_percent,
                     count, thd_spec, static_cast<double*>(res));
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  double v[8];
  SPX_HIP(hipMemcpyAsync(v, res, sizeof(double) * 7, hipMemcpyDeviceToHost, h->stream),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  if (v[6] > 0.0)
    return fail(SPX_ERR_INVALID_ARGUMENT,
                std::to_string((int64_t)v[6]) + " THD value(s) are NaN; they have no order");
  stats->mean_thd = v[0];
  stats->std_thd = v[1];
  stats->min_thd = v[2];
  stats->max_thd = v[3];
  stats->median_thd = v[4];
  stats->yield_percent = v[5];
  stats->num_samples = count;
  return SPX_OK;
}

}  // extern "C"
