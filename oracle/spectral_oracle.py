"""CPU restatement of the reference's spectral analysis of transient
waveforms: uniform resampling and linear interpolation
(crates/spicier-solver/src/spectral/fft.rs:98-147,
crates/spicier-solver/src/transient/result.rs:41-68), window functions
(spectral/window.rs:24-108), the windowed FFT spectrum (fft.rs:171-367), THD
with harmonic peak search (spectral/thd.rs:60-312), and the batch reductions
over a sweep's waveforms (crates/spicier-batched-sweep/src/spectral.rs:33-330).

TEST INFRASTRUCTURE ONLY: imported by tests/ — never by the engine.

The reference's DFT is rustfft 6.4.1 (a crates.io dependency, not in the
reference tree): its forward transform is the unnormalised DFT
X_k = sum_n x_n exp(-2 pi i k n / N) for any N.  numpy.fft.fft computes the
same published definition; tests pin it against a brute-force DFT.  Every
other step is a plain Python loop in the reference's order.
"""
from __future__ import annotations

import math

import numpy as np

RECTANGULAR, HANNING, HAMMING, BLACKMAN = 0, 1, 2, 3


def interpolate_at(times, values, t):
    """TransientResult::interpolate_at (result.rs:41-68) for one node.
    Returns None where the reference does."""
    if len(times) == 0:
        return None
    if t <= times[0]:
        return values[0]
    if t >= times[-1]:
        return values[-1]
    for i in range(len(times) - 1):
        t0, t1 = times[i], times[i + 1]
        if t0 <= t <= t1:
            alpha = (t - t0) / (t1 - t0)
            return values[i] * (1.0 - alpha) + values[i + 1] * alpha
    return None


def resample_uniform(times, values, sample_rate=None):
    """resample_uniform (fft.rs:98-147).  Returns (samples, sample_rate)."""
    times = [float(t) for t in times]
    values = [float(v) for v in values]
    if not times:
        return [], 0.0
    t_start, t_stop = times[0], times[-1]
    duration = t_stop - t_start
    if duration <= 0.0:
        return [], 0.0
    if sample_rate is None:
        n = len(times)
        sr = (n - 1) / duration if n > 1 else 1.0
    else:
        sr = float(sample_rate)
    num_samples = int(math.ceil(duration * sr)) + 1
    dt = duration / (num_samples - 1) if num_samples > 1 else duration
    out = []
    for i in range(num_samples):
        t = t_start + float(i) * dt
        v = interpolate_at(times, values, t)
        if v is not None:
            out.append(v)
        elif out:
            out.append(out[-1])
        else:
            out.append(0.0)
    return out, sr


def window_coefficients(window, n):
    """WindowFunction::apply's weights (window.rs:24-67), reference expression order."""
    w = []
    for i in range(n):
        if window == RECTANGULAR:
            w.append(1.0)
            continue
        d = float(n - 1)
        phase = 2.0 * math.pi * float(i) / d if d != 0.0 else float("nan")
        if window == HANNING:
            w.append(0.5 * (1.0 - math.cos(phase)))
        elif window == HAMMING:
            w.append(0.54 - 0.46 * math.cos(phase))
        else:
            w.append(0.42 - 0.5 * math.cos(phase) + 0.08 * math.cos(2.0 * phase))
    return w


def window_apply(window, signal):
    """WindowFunction::apply (window.rs:24-67)."""
    if window == RECTANGULAR:
        return [float(x) for x in signal]
    w = window_coefficients(window, len(signal))
    return [float(x) * wi for x, wi in zip(signal, w)]


def coherent_gain(window, n):
    """WindowFunction::coherent_gain (window.rs:73-107)."""
    if n == 0:
        return 0.0
    if window == RECTANGULAR:
        return 1.0
    s = 0.0
    for wi in window_coefficients(window, n):
        s += wi
    return s / float(n)


def next_power_of_2(n):
    """fft.rs:150-159."""
    p = 1
    while p < n:
        p *= 2
    return p


def dft(x):
    """The forward DFT that rustfft computes, via numpy's FFT."""
    return np.fft.fft(np.asarray(x, dtype=complex))


def empty_spectrum():
    return dict(frequencies=[], magnitude=[], magnitude_db=[], phase=[],
                power_spectral_density=[], sample_rate=0.0, fft_size=0)


def _spectrum(samples, sample_rate, window, fft_size, zero_padding):
    # the shared body of compute_fft (fft.rs:191-269) and
    # compute_fft_from_samples (fft.rs:296-366)
    size = fft_size if fft_size else next_power_of_2(len(samples))
    windowed = window_apply(window, samples)
    gain = coherent_gain(window, len(windowed))
    x = list(windowed)
    if zero_padding and len(x) < size:
        x += [0.0] * (size - len(x))
    n_fft = len(x)
    X = dft(x)
    n_pos = n_fft // 2 + 1
    res = sample_rate / float(n_fft)
    norm = float(len(windowed)) * gain
    out = empty_spectrum()
    for i in range(n_pos):
        out["frequencies"].append(float(i) * res)
        # IEEE division as the reference's f64: a zero coherent gain (a
        # two-point Hanning window weighs both points zero) gives 0 / 0 = NaN
        with np.errstate(divide="ignore", invalid="ignore"):
            re = float(np.float64(X[i].real) / norm)
            im = float(np.float64(X[i].imag) / norm)
        unscaled = math.hypot(re, im)
        scale = 1.0 if (i == 0 or (n_fft % 2 == 0 and i == n_pos - 1)) else 2.0
        mag = unscaled * scale
        out["magnitude"].append(mag)
        out["magnitude_db"].append(20.0 * math.log10(mag) if mag > 1e-20 else -400.0)
        out["phase"].append(math.atan2(im, re))
        out["power_spectral_density"].append(unscaled * unscaled * scale)
    out["sample_rate"] = sample_rate
    out["fft_size"] = n_fft
    return out


def compute_fft_from_samples(samples, sample_rate, window=HANNING, fft_size=None,
                             zero_padding=True):
    """compute_fft_from_samples (fft.rs:279-367)."""
    if len(samples) == 0 or sample_rate <= 0.0:
        return empty_spectrum()
    return _spectrum([float(s) for s in samples], float(sample_rate), window, fft_size,
                     zero_padding)


def compute_fft(times, values, window=HANNING, fft_size=None, zero_padding=True):
    """compute_fft (fft.rs:171-270) for one node's waveform."""
    samples, sr = resample_uniform(times, values, None)
    if not samples:
        return empty_spectrum()
    return _spectrum(samples, sr, window, fft_size, zero_padding)


def _find_bin_for_frequency(freqs, target):
    # thd.rs:101-127
    if not freqs or target < 0.0:
        return None
    best, best_diff = 0, abs(freqs[0] - target)
    for i, f in enumerate(freqs):
        d = abs(f - target)
        if d < best_diff:
            best, best_diff = i, d
    return best


def _find_peak_near_frequency(freqs, mags, target, search_range):
    # thd.rs:132-153
    best, best_mag = None, 0.0
    for i, (f, m) in enumerate(zip(freqs, mags)):
        if abs(f - target) <= search_range and m > best_mag:
            best, best_mag = i, m
    return best


def _db(m):
    return 20.0 * math.log10(m) if m > 1e-20 else -400.0


def thd_from_spectrum(spec, fundamental_freq, num_harmonics):
    """compute_thd_from_spectrum (thd.rs:156-312).  harmonics: list of
    (number, frequency, magnitude, magnitude_db, phase, relative_percent)."""
    freqs, mags, phases = spec["frequencies"], spec["magnitude"], spec["phase"]
    if not freqs or fundamental_freq <= 0.0:
        return dict(thd_percent=0.0, thd_db=-math.inf, fundamental_freq=fundamental_freq,
                    fundamental_magnitude=0.0, harmonics=[], harmonic_rms=0.0, num_harmonics=0)
    res = freqs[1] - freqs[0] if len(freqs) > 1 else fundamental_freq * 0.01
    search = res * 2.0
    idx = _find_peak_near_frequency(freqs, mags, fundamental_freq, search)
    if idx is None:
        idx = _find_bin_for_frequency(freqs, fundamental_freq) or 0
    fund, fphase, ffreq = mags[idx], phases[idx], freqs[idx]
    harmonics = [(1, ffreq, fund, _db(fund), fphase, 100.0)]
    sum_sq, found = 0.0, 0
    for n in range(2, num_harmonics + 2):
        target = fundamental_freq * float(n)
        if target > spec["sample_rate"] / 2.0:
            break
        idx = _find_peak_near_frequency(freqs, mags, target, search)
        if idx is None:
            idx = _find_bin_for_frequency(freqs, target)
        m, p, f = mags[idx], phases[idx], freqs[idx]
        rel = (m / fund) * 100.0 if fund > 1e-20 else 0.0
        harmonics.append((n, f, m, _db(m), p, rel))
        sum_sq += m * m
        found += 1
    rms = math.sqrt(sum_sq)
    thd = (rms / fund) * 100.0 if fund > 1e-20 else 0.0
    thd_db = 20.0 * math.log10(thd / 100.0) if thd > 1e-20 else -math.inf
    return dict(thd_percent=thd, thd_db=thd_db, fundamental_freq=ffreq, fundamental_magnitude=fund,
                harmonics=harmonics, harmonic_rms=rms, num_harmonics=found)


def compute_thd(times, values, fundamental_freq, num_harmonics):
    """compute_thd (thd.rs:60-75): Hanning, zero-padded to a power of two."""
    return thd_from_spectrum(compute_fft(times, values, HANNING, None, True), fundamental_freq,
                             num_harmonics)


def compute_thd_from_samples(samples, sample_rate, fundamental_freq, num_harmonics):
    """compute_thd_from_samples (thd.rs:84-98)."""
    return thd_from_spectrum(compute_fft_from_samples(samples, sample_rate, HANNING, None, True),
                             fundamental_freq, num_harmonics)


def thd_statistics(values, thd_spec):
    """compute_thd_statistics (spectral.rs:103-160)."""
    if len(values) == 0:
        return dict(mean_thd=0.0, std_thd=0.0, min_thd=0.0, max_thd=0.0, median_thd=0.0,
                    yield_percent=100.0, num_samples=0, thd_spec=thd_spec)
    v = [float(x) for x in values]
    n = float(len(v))
    s = 0.0
    for x in v:
        s += x
    mean = s / n
    var = 0.0
    for x in v:
        var += (x - mean) ** 2
    std = math.sqrt(var / n)
    mn, mx = math.inf, -math.inf
    for x in v:
        mn = mn if x != x else min(mn, x)
        mx = mx if x != x else max(mx, x)
    srt = sorted(v)
    k = len(srt)
    median = (srt[k // 2 - 1] + srt[k // 2]) / 2.0 if k % 2 == 0 else srt[k // 2]
    passing = sum(1 for x in v if x <= thd_spec)
    return dict(mean_thd=mean, std_thd=std, min_thd=mn, max_thd=mx, median_thd=median,
                yield_percent=(passing / n) * 100.0, num_samples=len(v), thd_spec=thd_spec)


def batch_stats_from_spectra(spectra):
    """compute_batch_stats_from_spectra (spectral.rs:262-330): two passes."""
    if not spectra:
        return dict(frequencies=[], mean_magnitude=[], std_magnitude=[], min_magnitude=[],
                    max_magnitude=[], num_samples=0)
    nb = len(spectra[0]["frequencies"])
    n = float(len(spectra))
    mean = [0.0] * nb
    mn = [math.inf] * nb
    mx = [-math.inf] * nb
    for sp in spectra:
        for i, m in enumerate(sp["magnitude"]):
            if i < nb:
                mean[i] += m
                mn[i] = mn[i] if m != m else min(mn[i], m)
                mx[i] = mx[i] if m != m else max(mx[i], m)
    mean = [m / n for m in mean]
    std = [0.0] * nb
    for sp in spectra:
        for i, m in enumerate(sp["magnitude"]):
            if i < nb:
                std[i] += (m - mean[i]) ** 2
    std = [math.sqrt(s / n) for s in std]
    return dict(frequencies=list(spectra[0]["frequencies"]), mean_magnitude=mean,
                std_magnitude=std, min_magnitude=mn, max_magnitude=mx, num_samples=len(spectra))
