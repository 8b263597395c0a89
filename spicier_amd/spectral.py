"""Host mirror of the reference's spectral analysis of sweep waveforms, over
the C ABI (include/spicier_hip.h).  Spectra, THD and their statistics run in
libspicier_hip.so on device buffers; torch tensors are only the device
memory the calls read and write.

Names, argument meaning and results follow:
  * ``WindowFunction`` (crates/spicier-solver/src/spectral/window.rs:8-108),
  * ``SpectralConfig`` / ``SpectralResult`` / ``compute_fft`` /
    ``compute_fft_from_samples`` (spectral/fft.rs:10-367),
  * ``HarmonicInfo`` / ``ThdResult`` / ``compute_thd`` /
    ``compute_thd_from_samples`` (spectral/thd.rs:14-312),
  * ``ThdStatistics`` / ``compute_batch_thd`` / ``compute_batch_thd_with_stats``
    / ``compute_thd_statistics`` / ``BatchSpectralResult`` /
    ``compute_batch_fft_stats`` (crates/spicier-batched-sweep/src/spectral.rs:12-330).

A waveform here is one node of a transient: ``TransientWaveform`` (times and
a [points][nodes] solution array, the reference's TransientResult) for one
instance, or a ``WaveformBatch`` over every instance of a sweep (host or
device data, read in place with strides).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from enum import IntEnum

import numpy as np

from . import _abi as A
from .engine import Device, InvalidArgument, InvalidDimension, _check


def _torch():
    import torch
    return torch


class WindowFunction(IntEnum):
    """window.rs:8-21; Hanning is the default."""
    RECTANGULAR = A.SPX_WINDOW_RECTANGULAR
    HANNING = A.SPX_WINDOW_HANNING
    HAMMING = A.SPX_WINDOW_HAMMING
    BLACKMAN = A.SPX_WINDOW_BLACKMAN

    def coefficients(self, n: int) -> np.ndarray:
        out = np.empty(max(int(n), 0))
        _check(A.load_library().spx_window_coefficients(int(self), int(n), A.ptr(out)))
        return out

    def apply(self, signal) -> np.ndarray:
        """WindowFunction::apply (window.rs:24-67)."""
        x = np.asarray(signal, dtype=np.float64)
        if self == WindowFunction.RECTANGULAR:
            return x.copy()
        return x * self.coefficients(x.size)

    def coherent_gain(self, n: int) -> float:
        """WindowFunction::coherent_gain (window.rs:73-107)."""
        return float(A.load_library().spx_window_coherent_gain(int(self), int(n)))


@dataclass
class SpectralConfig:
    """fft.rs:10-29."""
    window: WindowFunction = WindowFunction.HANNING
    fft_size: int | None = None
    zero_padding: bool = True

    def to_c(self) -> A.SpxSpectralConfig:
        return A.SpxSpectralConfig(int(self.window), int(bool(self.zero_padding)),
                                   int(self.fft_size or 0))


@dataclass
class SpectralResult:
    """fft.rs:31-85."""
    frequencies: np.ndarray
    magnitude: np.ndarray
    magnitude_db: np.ndarray
    phase: np.ndarray
    power_spectral_density: np.ndarray
    sample_rate: float
    fft_size: int

    def peak_frequency(self):
        if self.magnitude.size == 0:
            return None
        # max_by keeps the last of equal maxima (fft.rs:52-58)
        m = self.magnitude
        i = m.size - 1 - int(np.argmax(m[::-1]))
        return float(self.frequencies[i])

    def _bin(self, freq: float):
        if self.frequencies.size == 0 or freq < 0.0:
            return None
        b = int(round_half_away(freq / (self.sample_rate / self.fft_size)))
        return b if b < self.frequencies.size else None

    def magnitude_at(self, freq: float):
        b = self._bin(freq)
        return None if b is None else float(self.magnitude[b])

    def magnitude_db_at(self, freq: float):
        b = self._bin(freq)
        return None if b is None else float(self.magnitude_db[b])


def round_half_away(x: float) -> float:
    """f64::round: halves away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


@dataclass
class HarmonicInfo:
    """thd.rs:14-29."""
    harmonic_number: int
    frequency: float
    magnitude: float
    magnitude_db: float
    phase: float
    relative_percent: float


@dataclass
class ThdResult:
    """thd.rs:31-48."""
    thd_percent: float
    thd_db: float
    fundamental_freq: float
    fundamental_magnitude: float
    harmonics: list
    harmonic_rms: float
    num_harmonics: int


@dataclass
class ThdStatistics:
    """spectral.rs:12-31."""
    mean_thd: float
    std_thd: float
    min_thd: float
    max_thd: float
    median_thd: float
    yield_percent: float
    num_samples: int
    thd_spec: float


@dataclass
class BatchSpectralResult:
    """spectral.rs:162-177."""
    frequencies: np.ndarray
    mean_magnitude: np.ndarray
    std_magnitude: np.ndarray
    min_magnitude: np.ndarray
    max_magnitude: np.ndarray
    num_samples: int


@dataclass
class TransientWaveform:
    """One instance's transient (TransientResult, transient/result.rs:15-21):
    times [points] and solutions [points][num_nodes]."""
    times: np.ndarray
    solutions: np.ndarray

    def voltage_waveform(self, node_idx: int):
        return list(zip(self.times.tolist(), self.solutions[:, node_idx].tolist()))


@dataclass
class WaveformBatch:
    """One node's waveform for every instance of a sweep.  ``data`` is a host
    array or device tensor [batch][points] or [batch][points][probes] (the
    transient sweep's layout); ``node_idx`` picks the probe column.  ``times``
    (strictly increasing) makes the analysis resample to uniform steps as
    compute_fft does; without it the points are uniform at ``sample_rate``."""
    data: object
    times: np.ndarray | None = None
    sample_rate: float | None = None
    node_idx: int = 0
    device_id: int = 0
    _dev: object = field(default=None, repr=False)
    _times_c: object = field(default=None, repr=False)

    @classmethod
    def from_transient(cls, result, node_idx: int, device_id: int = 0) -> "WaveformBatch":
        """From a transient sweep (engine.TransientSweepResult or the device
        variant's tensors): node_idx is the probe column."""
        return cls(result.waveforms, times=np.asarray(result.times, dtype=np.float64),
                   node_idx=node_idx, device_id=device_id)

    def _prepare(self):
        torch = _torch()
        if self._dev is None:
            d = self.data
            if isinstance(d, torch.Tensor) and d.device.type == "cuda":
                t = d if d.dtype == torch.float64 else d.to(torch.float64)
            else:
                t = torch.as_tensor(np.asarray(d, dtype=np.float64))
                t = t.to(torch.device(f"cuda:{self.device_id}"))
            if t.dim() == 2:
                t = t.unsqueeze(-1)
            if t.dim() != 3:
                raise InvalidDimension(
                    "waveforms must be [batch][points] or [batch][points][probes]")
            self._dev = t.contiguous()
        t = self._dev
        B, T, P = t.shape
        if not 0 <= self.node_idx < P:
            raise InvalidDimension(f"node_idx {self.node_idx} outside 0..{P - 1}")
        c = A.SpxWaveformBatch()
        c.data = C.c_void_p(t.data_ptr() + 8 * self.node_idx) if t.numel() else None
        c.batch, c.batch_stride, c.num_points, c.point_stride = B, T * P, T, P
        if self.times is not None:
            tt = np.ascontiguousarray(self.times, dtype=np.float64)
            if tt.size != T:
                raise InvalidDimension(f"{tt.size} times for {T} points")
            self._times_c = tt
            c.times = tt.ctypes.data_as(C.POINTER(C.c_double))
            c.sample_rate = 0.0
        else:
            if self.sample_rate is None:
                raise InvalidArgument("give times or sample_rate")
            c.sample_rate = float(self.sample_rate)
        return c


def _info(dev, wb_c, cfg_c) -> A.SpxSpectrumInfo:
    info = A.SpxSpectrumInfo()
    _check(dev.lib.spx_spectrum_info_query(dev.handle, C.byref(wb_c), C.byref(cfg_c),
                                           C.byref(info)))
    return info


@dataclass
class BatchSpectra:
    """Every waveform's SpectralResult arrays, [batch][num_bins] device tensors."""
    frequencies: np.ndarray
    magnitude: object
    magnitude_db: object
    phase: object
    power_spectral_density: object
    sample_rate: float
    fft_size: int

    def __len__(self):
        return int(self.magnitude.shape[0])

    def __getitem__(self, i) -> SpectralResult:
        return SpectralResult(self.frequencies.copy(), self.magnitude[i].cpu().numpy(),
                              self.magnitude_db[i].cpu().numpy(), self.phase[i].cpu().numpy(),
                              self.power_spectral_density[i].cpu().numpy(), self.sample_rate,
                              self.fft_size)


def _run_fft(wb: WaveformBatch, config: SpectralConfig | None, spectra: bool, stats: bool):
    torch = _torch()
    cfg = (config or SpectralConfig()).to_c()
    dev = Device.get(wb.device_id)
    c = wb._prepare()
    info = _info(dev, c, cfg)
    B, nb = c.batch, int(info.num_bins)
    freqs = np.arange(nb, dtype=np.float64) * info.freq_resolution
    d = torch.device(f"cuda:{wb.device_id}")
    per = [torch.empty((B, nb), dtype=torch.float64, device=d) for _ in range(4)] if spectra \
        else [None] * 4
    agg = [torch.empty(nb, dtype=torch.float64, device=d) for _ in range(4)] if stats \
        else [None] * 4
    out = A.SpxSpectrumOut(*[A.ptr(x) for x in per + agg])
    torch.cuda.synchronize(wb.device_id)
    _check(dev.lib.spx_batch_fft_device(dev.handle, C.byref(c), C.byref(cfg), C.byref(out),
                                        C.byref(info)))
    dev.synchronize()
    return info, freqs, per, agg


def compute_batch_fft(waveforms: WaveformBatch, config: SpectralConfig | None = None) -> BatchSpectra:
    """compute_fft (fft.rs:171-270) for every waveform of a batch."""
    info, freqs, per, _ = _run_fft(waveforms, config, True, False)
    return BatchSpectra(freqs, per[0], per[1], per[2], per[3], float(info.sample_rate),
                        int(info.fft_size))


def compute_batch_fft_stats(waveforms: WaveformBatch,
                            config: SpectralConfig | None = None) -> BatchSpectralResult:
    """compute_batch_fft_stats (spectral.rs:190-260): per-bin mean, population
    std, min and max of the magnitude over the batch."""
    B = waveforms._prepare().batch
    if B == 0:
        e = np.zeros(0)
        return BatchSpectralResult(e, e, e, e, e, 0)
    info, freqs, _, agg = _run_fft(waveforms, config, False, True)
    if info.fft_size == 0:
        e = np.zeros(0)
        return BatchSpectralResult(e, e, e, e, e, B)
    return BatchSpectralResult(freqs, *[a.cpu().numpy() for a in agg], B)


def _single(samples_or_wave, sample_rate, node_idx, device_id) -> WaveformBatch:
    if isinstance(samples_or_wave, TransientWaveform):
        v = np.asarray(samples_or_wave.solutions, dtype=np.float64)[:, node_idx]
        return WaveformBatch(v.reshape(1, -1), times=np.asarray(samples_or_wave.times),
                             device_id=device_id)
    v = np.asarray(samples_or_wave, dtype=np.float64).reshape(1, -1)
    return WaveformBatch(v, sample_rate=sample_rate, device_id=device_id)


def compute_fft(waveform: TransientWaveform, node_idx: int,
                config: SpectralConfig | None = None, device_id: int = 0) -> SpectralResult:
    """compute_fft (fft.rs:171-270): resampled to uniform steps first."""
    return _one_spectrum(_single(waveform, None, node_idx, device_id), config)


def compute_fft_from_samples(samples, sample_rate: float, config: SpectralConfig | None = None,
                             device_id: int = 0) -> SpectralResult:
    """compute_fft_from_samples (fft.rs:279-367)."""
    return _one_spectrum(_single(samples, sample_rate, 0, device_id), config)


def _one_spectrum(wb, config) -> SpectralResult:
    if np.size(wb.data) == 0:
        e = np.zeros(0)
        return SpectralResult(e, e, e, e, e, 0.0, 0)
    bs = compute_batch_fft(wb, config)
    if bs.fft_size == 0:
        e = np.zeros(0)
        return SpectralResult(e, e, e, e, e, 0.0, 0)
    return bs[0]


@dataclass
class BatchThdResult:
    """compute_batch_thd's Vec<ThdResult> as device arrays: per waveform
    thd_percent, thd_db, fundamental_freq, fundamental_magnitude,
    harmonic_rms [batch] and harmonics [batch][num_harmonics+1][5]."""
    thd_percent: object
    thd_db: object
    fundamental_freq: object
    fundamental_magnitude: object
    harmonic_rms: object
    harmonics: object
    num_harmonics: int

    def __len__(self):
        return int(self.thd_percent.shape[0])

    def __getitem__(self, i) -> ThdResult:
        if self.num_harmonics < 0:  # the empty ThdResult (thd.rs:161-171)
            return ThdResult(0.0, -math.inf, float(self.fundamental_freq[i]), 0.0, [], 0.0, 0)
        h = self.harmonics[i, :self.num_harmonics + 1].cpu().numpy()
        infos = [HarmonicInfo(k + 1, *map(float, h[k])) for k in range(h.shape[0])]
        return ThdResult(float(self.thd_percent[i]), float(self.thd_db[i]),
                         float(self.fundamental_freq[i]), float(self.fundamental_magnitude[i]),
                         infos, float(self.harmonic_rms[i]), self.num_harmonics)

    def __iter__(self):
        return (self[i] for i in range(len(self)))


def compute_batch_thd(waveforms: WaveformBatch, fundamental_freq: float,
                      num_harmonics: int) -> BatchThdResult:
    """compute_batch_thd (spectral.rs:46-71): compute_thd (thd.rs:60-75) for
    every waveform, Hanning window, zero padding to a power of two."""
    torch = _torch()
    dev = Device.get(waveforms.device_id)
    c = waveforms._prepare()
    B = c.batch
    d = torch.device(f"cuda:{waveforms.device_id}")
    cols = [torch.empty(B, dtype=torch.float64, device=d) for _ in range(5)]
    harm = torch.full((B, num_harmonics + 1, 5), float("nan"), dtype=torch.float64, device=d)
    out = A.SpxThdOut(*[A.ptr(x) for x in cols + [harm]])
    found = C.c_int32(0)
    torch.cuda.synchronize(waveforms.device_id)
    _check(dev.lib.spx_batch_thd_device(dev.handle, C.byref(c), float(fundamental_freq),
                                        int(num_harmonics), C.byref(out), C.byref(found)))
    dev.synchronize()
    empty = not (fundamental_freq > 0.0) or _info(dev, c, SpectralConfig().to_c()).fft_size == 0
    # num_harmonics -1: every waveform has the empty ThdResult
    return BatchThdResult(*cols, harm, -1 if empty else int(found.value))


def compute_thd(waveform: TransientWaveform, node_idx: int, fundamental_freq: float,
                num_harmonics: int, device_id: int = 0) -> ThdResult:
    """compute_thd (thd.rs:60-75)."""
    return _one_thd(_single(waveform, None, node_idx, device_id), fundamental_freq, num_harmonics)


def compute_thd_from_samples(samples, sample_rate: float, fundamental_freq: float,
                             num_harmonics: int, device_id: int = 0) -> ThdResult:
    """compute_thd_from_samples (thd.rs:84-98)."""
    return _one_thd(_single(samples, sample_rate, 0, device_id), fundamental_freq, num_harmonics)


def _one_thd(wb, f0, h) -> ThdResult:
    if np.size(wb.data) == 0:
        return ThdResult(0.0, -math.inf, float(f0), 0.0, [], 0.0, 0)
    return compute_batch_thd(wb, f0, h)[0]


def compute_thd_statistics(results, thd_spec: float, device_id: int = 0) -> ThdStatistics:
    """compute_thd_statistics (spectral.rs:103-160) over THD results, a
    BatchThdResult, or a sequence / tensor of THD percentages."""
    torch = _torch()
    if isinstance(results, BatchThdResult):
        v = results.thd_percent
    elif isinstance(results, (list, tuple)) and results and isinstance(results[0], ThdResult):
        v = [r.thd_percent for r in results]
    else:
        v = results
    if not (isinstance(v, torch.Tensor) and v.device.type == "cuda"):
        v = torch.as_tensor(np.asarray(v, dtype=np.float64)).to(torch.device(f"cuda:{device_id}"))
    v = v.to(torch.float64).contiguous()
    dev = Device.get(device_id)
    s = A.SpxThdStatistics()
    torch.cuda.synchronize(device_id)
    _check(dev.lib.spx_thd_statistics_device(dev.handle, A.ptr(v) if v.numel() else None,
                                             int(v.numel()), float(thd_spec), C.byref(s)))
    return ThdStatistics(s.mean_thd, s.std_thd, s.min_thd, s.max_thd, s.median_thd,
                         s.yield_percent, int(s.num_samples), s.thd_spec)


def compute_batch_thd_with_stats(waveforms: WaveformBatch, fundamental_freq: float,
                                 num_harmonics: int, thd_spec: float):
    """compute_batch_thd_with_stats (spectral.rs:86-100)."""
    r = compute_batch_thd(waveforms, fundamental_freq, num_harmonics)
    return r, compute_thd_statistics(r, thd_spec, waveforms.device_id)
