"""Pins the spectral oracle (oracle/spectral_oracle.py) with the reference's
own known-answer tests (spectral/window.rs:110-163, spectral/fft.rs:369-491,
spectral/thd.rs:314-433, spicier-batched-sweep/src/spectral.rs:332-441) and
with what the mathematics fixes: a brute-force DFT, the window sums in closed
form, Parseval's identity for the rectangular window, exact amplitudes of
on-bin sinusoids, and exact interpolation of linear waveforms."""
import math

import numpy as np
import pytest

from oracle import spectral_oracle as S


def _sin(freq, amp, sr, dur):
    n = int(sr * dur)
    return [amp * math.sin(2 * math.pi * freq * (i / sr)) for i in range(n)]


def _harmonics(f0, amps, sr, dur):
    n = int(sr * dur)
    return [sum(a * math.sin(2 * math.pi * f0 * (h + 1) * (i / sr)) for h, a in enumerate(amps))
            for i in range(n)]


# ---------------------------------------------------------------- DFT
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 17, 64])
def test_dft_matches_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    k = np.arange(n)
    brute = np.array([np.sum(x * np.exp(-2j * np.pi * kk * k / n)) for kk in k])
    np.testing.assert_allclose(S.dft(x), brute, rtol=0, atol=1e-12 * n)


# ------------------------------------------------------------- windows
def test_reference_window_tests():
    # window.rs:114-162
    assert S.window_apply(S.RECTANGULAR, [1.0, 2.0, 3.0, 4.0]) == [1.0, 2.0, 3.0, 4.0]
    h = S.window_apply(S.HANNING, [1.0] * 100)
    assert abs(h[0]) < 1e-10 and abs(h[99]) < 1e-10 and abs(h[50] - 1.0) < 0.01
    hm = S.window_apply(S.HAMMING, [1.0] * 100)
    assert abs(hm[0] - 0.08) < 0.01 and abs(hm[99] - 0.08) < 0.01
    b = S.window_apply(S.BLACKMAN, [1.0] * 100)
    assert abs(b[0]) < 0.001 and abs(b[99]) < 0.001
    assert abs(S.coherent_gain(S.RECTANGULAR, 1024) - 1.0) < 1e-10
    assert abs(S.coherent_gain(S.HANNING, 1024) - 0.5) < 0.01
    assert S.window_apply(S.HANNING, []) == []


@pytest.mark.parametrize("n", [4, 10, 101, 1024])
def test_coherent_gain_closed_form(n):
    # for n >= 4, sum_{i<n} cos(2 pi m i/(n-1)) = 1 for m = 1, 2 (full
    # periods of n-1 >= 3 terms plus the i = n-1 term), so the sums are
    # a0*n - a1 + a2
    assert S.coherent_gain(S.HANNING, n) == pytest.approx((0.5 * n - 0.5) / n, rel=1e-13)
    assert S.coherent_gain(S.HAMMING, n) == pytest.approx((0.54 * n - 0.46) / n, rel=1e-13)
    assert S.coherent_gain(S.BLACKMAN, n) == pytest.approx((0.42 * n - 0.42) / n, rel=1e-12,
                                                          abs=1e-15)


def test_coherent_gain_two_points():
    # n = 2: both points sit at phase 0 and 2 pi, where the Hanning weight is 0
    assert S.coherent_gain(S.HANNING, 2) == 0.0
    assert S.coherent_gain(S.HAMMING, 2) == pytest.approx(0.08)
    # n = 3: phases 0, pi, 2 pi; the Blackman cos(2 phase) term is 1 at all three
    assert S.coherent_gain(S.HANNING, 3) == pytest.approx(1.0 / 3.0)
    assert S.coherent_gain(S.BLACKMAN, 3) == pytest.approx((0.42 * 3 - 0.5 + 0.08 * 3) / 3)


# ----------------------------------------------------------------- FFT
def test_reference_fft_tests():
    rect = dict(window=S.RECTANGULAR, fft_size=None, zero_padding=False)
    r = S.compute_fft_from_samples(_sin(1000.0, 1.0, 10000.0, 0.1), 10000.0, **rect)
    peak = r["frequencies"][int(np.argmax(r["magnitude"]))]
    assert abs(peak - 1000.0) < 20.0
    bw = r["sample_rate"] / r["fft_size"]
    assert abs(r["magnitude"][int(round(1000.0 / bw))] - 1.0) < 0.1     # fft.rs:384-411
    x = _sin(500.0, 2.0, 8000.0, 0.128)
    r = S.compute_fft_from_samples(x, 8000.0, **rect)
    ratio = sum(r["power_spectral_density"]) / (sum(v * v for v in x) / len(x))
    assert abs(ratio - 1.0) < 0.15                                       # fft.rs:413-440
    r = S.compute_fft_from_samples([5.0] * 1024, 1000.0, **rect)
    assert abs(r["magnitude"][0] - 5.0) < 0.01 and max(r["magnitude"][1:]) < 0.01
    assert [S.next_power_of_2(n) for n in (0, 1, 2, 3, 1000, 1024, 1025)] == \
        [1, 1, 2, 4, 1024, 1024, 2048]                                   # fft.rs:472-481
    r = S.compute_fft_from_samples([], 1000.0)
    assert r["frequencies"] == [] and r["magnitude"] == []


@pytest.mark.parametrize("n", [16, 27, 64, 100])
def test_parseval_exact_rectangular(n):
    # single-sided PSD sums to the mean square exactly when nothing is padded
    rng = np.random.default_rng(n)
    x = list(rng.standard_normal(n))
    r = S.compute_fft_from_samples(x, 1.0, S.RECTANGULAR, None, False)
    assert sum(r["power_spectral_density"]) == pytest.approx(np.mean(np.square(x)), rel=1e-12)


@pytest.mark.parametrize("n,k0", [(64, 5), (100, 7), (256, 1)])
def test_on_bin_sinusoid_amplitude_and_phase(n, k0):
    amp, ph = 1.7, 0.3
    x = [amp * math.cos(2 * math.pi * k0 * i / n + ph) for i in range(n)]
    r = S.compute_fft_from_samples(x, float(n), S.RECTANGULAR, None, False)
    assert r["magnitude"][k0] == pytest.approx(amp, rel=1e-12)
    assert r["phase"][k0] == pytest.approx(ph, abs=1e-12)
    assert r["frequencies"][k0] == pytest.approx(k0, rel=1e-15)
    others = [m for i, m in enumerate(r["magnitude"]) if i != k0]
    assert max(others) < 1e-12


def test_zero_padding_and_sizes():
    x = list(np.random.default_rng(1).standard_normal(100))
    assert S.compute_fft_from_samples(x, 1.0)["fft_size"] == 128
    assert S.compute_fft_from_samples(x, 1.0, fft_size=300)["fft_size"] == 300
    assert S.compute_fft_from_samples(x, 1.0, fft_size=64)["fft_size"] == 100
    assert S.compute_fft_from_samples(x, 1.0, zero_padding=False)["fft_size"] == 100
    assert len(S.compute_fft_from_samples(x, 1.0, zero_padding=False)["magnitude"]) == 51
    assert len(S.compute_fft_from_samples(x[:99], 1.0, zero_padding=False)["magnitude"]) == 50


# ----------------------------------------------------------- resampling
def test_resample_linear_waveform_exact():
    # linear interpolation reproduces a linear waveform; sizes follow fft.rs:116-131
    t = [0.0, 0.1, 0.25, 0.3, 0.7, 1.0]
    v = [2.0 + 3.0 * ti for ti in t]
    s, sr = S.resample_uniform(t, v)
    assert sr == pytest.approx(5.0) and len(s) == int(math.ceil(1.0 * sr)) + 1
    dt = 1.0 / (len(s) - 1)
    np.testing.assert_allclose(s, [2.0 + 3.0 * i * dt for i in range(len(s))], rtol=1e-14)
    assert S.resample_uniform([0.0], [1.0]) == ([], 0.0)
    assert S.resample_uniform([], []) == ([], 0.0)


def test_resample_uniform_grid_keeps_points():
    h = 1e-12
    t = [k * h for k in range(201)]
    v = list(np.sin(np.arange(201) * 0.1))
    s, sr = S.resample_uniform(t, v)
    # ceil(duration * sr) is 200 or 201 depending on rounding; both are valid
    assert len(s) in (201, 202)
    if len(s) == 201:
        np.testing.assert_allclose(s, v, rtol=1e-12, atol=1e-15)


# ----------------------------------------------------------------- THD
def test_reference_thd_tests():
    r = S.compute_thd_from_samples(_harmonics(1000.0, [1.0], 44100.0, 0.1), 44100.0, 1000.0, 10)
    assert r["thd_percent"] < 1.0                                        # thd.rs:338-349
    r = S.compute_thd_from_samples(_harmonics(1000.0, [1.0, 0.1, 0.05], 44100.0, 0.1), 44100.0,
                                   1000.0, 10)
    assert abs(r["thd_percent"] - math.sqrt(0.1 ** 2 + 0.05 ** 2) * 100.0) < 2.0
    assert len(r["harmonics"]) >= 3 and r["harmonics"][0][0] == 1 and r["harmonics"][1][0] == 2
    sq = [1.0 if math.sin(2 * math.pi * 1000.0 * (i / 100000.0)) >= 0 else -1.0
          for i in range(10000)]
    r = S.compute_thd_from_samples(sq, 100000.0, 1000.0, 20)
    assert 30.0 < r["thd_percent"] < 60.0                                # thd.rs:375-400
    r = S.compute_thd_from_samples(_harmonics(1000.0, [1.0, 0.1], 44100.0, 0.1), 44100.0,
                                   1000.0, 5)
    assert abs(r["harmonics"][1][5] - 10.0) < 2.0                        # thd.rs:402-417
    assert abs(r["thd_db"] - 20.0 * math.log10(r["thd_percent"] / 100.0)) < 1.0


def test_thd_on_bin_harmonics_exact():
    # Hanning leaks into the bins either side; with harmonics 8 bins apart
    # the peak bins carry amp * 2 * gain(periodic)/gain(symmetric), and the
    # ratio of two such peaks is the amplitude ratio exactly
    n, sr = 1024, 1024.0
    f0 = 32.0
    x = [math.sin(2 * math.pi * f0 * i / sr) + 0.2 * math.sin(2 * math.pi * 2 * f0 * i / sr)
         + 0.05 * math.sin(2 * math.pi * 3 * f0 * i / sr) for i in range(n)]
    r = S.compute_thd_from_samples(x, sr, f0, 2)
    assert r["num_harmonics"] == 2
    assert [h[1] for h in r["harmonics"]] == [32.0, 64.0, 96.0]
    assert r["harmonics"][1][5] == pytest.approx(20.0, rel=1e-3)
    assert r["thd_percent"] == pytest.approx(math.hypot(0.2, 0.05) * 100.0, rel=1e-3)


def test_thd_edges():
    r = S.compute_thd_from_samples([], 1000.0, 100.0, 5)
    assert r["thd_percent"] == 0.0 and r["thd_db"] == -math.inf and r["num_harmonics"] == 0
    r = S.compute_thd_from_samples(list(np.ones(64)), 64.0, 0.0, 5)
    assert r["thd_percent"] == 0.0 and r["fundamental_freq"] == 0.0
    # harmonics past Nyquist are not counted (thd.rs:233-236)
    x = _sin(10.0, 1.0, 64.0, 1.0)
    r = S.compute_thd_from_samples(x, 64.0, 10.0, 10)
    assert r["num_harmonics"] == 2


# --------------------------------------------------------------- batch
def test_reference_batch_tests():
    thd = [1.0, 5.0, 15.0]
    assert abs(S.thd_statistics(thd, 10.0)["yield_percent"] - 66.67) < 1.0   # spectral.rs:374-411
    assert abs(S.thd_statistics(thd, 1.0)["yield_percent"] - 33.33) < 1.0
    st = S.thd_statistics([], 10.0)
    assert st["num_samples"] == 0 and st["yield_percent"] == 100.0
    waves = [[(1.0 + 0.01 * i) * math.sin(2 * math.pi * 1000.0 * (k / 100000.0))
              for k in range(1000)] for i in range(10)]
    times = [k / 100000.0 for k in range(1000)]
    res = [S.compute_thd(times, w, 1000.0, 5) for w in waves]
    st = S.thd_statistics([r["thd_percent"] for r in res], 10.0)
    assert st["num_samples"] == 10 and st["mean_thd"] < 5.0 and st["yield_percent"] > 90.0
    spectra = [S.compute_fft(times, waves[0]) for _ in range(5)]
    bs = S.batch_stats_from_spectra(spectra)
    assert bs["num_samples"] == 5 and bs["frequencies"] and max(bs["std_magnitude"]) < 0.01


def test_thd_statistics_closed_form():
    v = [3.0, 1.0, 4.0, 1.0, 5.0, 9.0]
    st = S.thd_statistics(v, 4.0)
    assert st["mean_thd"] == pytest.approx(np.mean(v)) and st["std_thd"] == pytest.approx(np.std(v))
    assert st["median_thd"] == 3.5 and st["min_thd"] == 1.0 and st["max_thd"] == 9.0
    assert st["yield_percent"] == pytest.approx(400.0 / 6.0)
    assert S.thd_statistics(v[:5], 4.0)["median_thd"] == 3.0


def test_batch_spectrum_stats_closed_form():
    # magnitudes scale linearly with amplitude: mean/std/min/max follow the amplitudes
    times = [k / 1000.0 for k in range(256)]
    base = [math.sin(2 * math.pi * 50.0 * t) + 0.3 for t in times]
    amps = [0.5, 1.0, 2.0, 4.5]
    spectra = [S.compute_fft(times, [a * b for b in base]) for a in amps]
    bs = S.batch_stats_from_spectra(spectra)
    m1 = np.asarray(S.compute_fft(times, base)["magnitude"])
    tol = 1e-13 * m1.max()   # rounding of noise-level bins scales with the peak
    np.testing.assert_allclose(bs["mean_magnitude"], np.mean(amps) * m1, rtol=1e-12, atol=tol)
    np.testing.assert_allclose(bs["std_magnitude"], np.std(amps) * m1, rtol=1e-10, atol=tol)
    np.testing.assert_allclose(bs["max_magnitude"], 4.5 * m1, rtol=1e-12, atol=tol)
    np.testing.assert_allclose(bs["min_magnitude"], 0.5 * m1, rtol=1e-12, atol=tol)
