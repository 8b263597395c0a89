"""Batch spectral analysis on the GPU (spicier_amd/spectral.py over
spx_batch_fft_device / spx_batch_thd_device / spx_thd_statistics_device)
against the oracle (oracle/spectral_oracle.py) on the same waveforms, and the
reference's own spectral tests run through the GPU path.

Tolerances (floating point): a spectrum bin agrees to 1e-12 of the
waveform's largest magnitude (the DFTs round differently); phase and dB are
compared on bins above 1e-6 of the peak, where they are defined to 1e-9 rad
and 1e-8 dB; THD figures to 1e-9 relative; the bins the harmonic search
picks must be identical."""
import math

import numpy as np
import pytest

from oracle import spectral_oracle as S
from spicier_amd import engine as E
from spicier_amd import spectral as SP
from spicier_amd.spectral import SpectralConfig, WaveformBatch, WindowFunction as W

pytestmark = pytest.mark.gpu


def _check_spectrum(got: SP.SpectralResult, want: dict):
    assert got.fft_size == want["fft_size"]
    assert got.sample_rate == want["sample_rate"]
    np.testing.assert_array_equal(got.frequencies, want["frequencies"])
    m = np.asarray(want["magnitude"])
    if m.size == 0:
        assert got.magnitude.size == 0
        return
    if np.all(np.isnan(m)):  # zero coherent gain: 0 / 0 in every bin, as the reference
        assert np.all(np.isnan(got.magnitude)) and np.all(np.isnan(got.power_spectral_density))
        return
    peak = max(m.max(), 1e-300)
    np.testing.assert_allclose(got.magnitude, m, rtol=0, atol=1e-12 * peak)
    p = np.asarray(want["power_spectral_density"])
    np.testing.assert_allclose(got.power_spectral_density, p, rtol=0, atol=1e-12 * max(p.max(), 1e-300))
    big = m > 1e-6 * peak
    dphi = np.angle(np.exp(1j * (got.phase[big] - np.asarray(want["phase"])[big])))
    assert np.all(np.abs(dphi) < 1e-9)
    np.testing.assert_allclose(got.magnitude_db[big], np.asarray(want["magnitude_db"])[big],
                               rtol=0, atol=1e-8)


CASES = [  # (points, window, fft_size, zero_padding)
    (1, W.RECTANGULAR, None, True), (2, W.HANNING, None, True), (3, W.HAMMING, None, False),
    (64, W.HANNING, None, True), (100, W.BLACKMAN, None, True), (100, W.HANNING, 300, True),
    (100, W.RECTANGULAR, 64, True), (201, W.HANNING, None, False), (1000, W.HAMMING, None, True),
    (1000, W.RECTANGULAR, None, False), (1024, W.HANNING, None, True),
    (3000, W.HANNING, None, True), (2047, W.BLACKMAN, None, False), (4096, W.HANNING, None, True),
    (5000, W.HAMMING, None, True), (4095, W.RECTANGULAR, None, False),
    # past 2^13 points: the four-step path through HBM (chirp-z on top for 9001)
    (10000, W.HANNING, None, True), (9001, W.RECTANGULAR, None, False),
    (70000, W.HAMMING, None, True),
]


@pytest.mark.parametrize("T,window,fft_size,pad", CASES)
def test_batch_fft_uniform_matches_oracle(T, window, fft_size, pad):
    rng = np.random.default_rng(T)
    B = 5
    x = rng.standard_normal((B, T)) + 0.3
    sr = 1000.0
    cfg = SpectralConfig(window, fft_size, pad)
    got = SP.compute_batch_fft(WaveformBatch(x, sample_rate=sr), cfg)
    for b in range(B):
        want = S.compute_fft_from_samples(list(x[b]), sr, int(window), fft_size, pad)
        _check_spectrum(got[b], want)


@pytest.mark.parametrize("kind", ["uniform_grid", "nonuniform"])
def test_batch_fft_resampled_matches_oracle(kind):
    rng = np.random.default_rng(7)
    if kind == "uniform_grid":
        T, h = 201, 1e-12
        times = np.arange(T) * h           # the transient sweep's time points
    else:
        T = 300
        times = np.cumsum(rng.uniform(0.2, 1.8, T)) * 1e-9
    B = 4
    # three probes; analyse the middle one, read in place with strides
    x = rng.standard_normal((B, T, 3))
    got = SP.compute_batch_fft(WaveformBatch(x, times=times, node_idx=1))
    for b in range(B):
        _check_spectrum(got[b], S.compute_fft(list(times), list(x[b, :, 1])))


def _harm_signal(rng, B, n, sr, f0):
    t = np.arange(n) / sr
    amps = rng.uniform(0.01, 0.3, (B, 6))
    amps[:, 0] = rng.uniform(0.8, 1.2, B)
    ph = rng.uniform(0, 2 * np.pi, (B, 6))
    return sum(amps[:, h:h + 1] * np.sin(2 * np.pi * f0 * (h + 1) * t + ph[:, h:h + 1])
               for h in range(6)) + 0.05 * rng.standard_normal((B, n))


def _check_thd(got: SP.ThdResult, want: dict):
    assert got.num_harmonics == want["num_harmonics"]
    for k in ("thd_percent", "fundamental_magnitude", "harmonic_rms"):
        assert getattr(got, k) == pytest.approx(want[k], rel=1e-9, abs=1e-15), k
    assert got.fundamental_freq == want["fundamental_freq"]
    if math.isfinite(want["thd_db"]):
        assert got.thd_db == pytest.approx(want["thd_db"], abs=1e-8)
    else:
        assert got.thd_db == want["thd_db"]
    assert len(got.harmonics) == len(want["harmonics"])
    for g, w in zip(got.harmonics, want["harmonics"]):
        assert g.harmonic_number == w[0] and g.frequency == w[1]
        assert g.magnitude == pytest.approx(w[2], rel=1e-9, abs=1e-15)
        assert g.relative_percent == pytest.approx(w[5], rel=1e-9, abs=1e-13)


@pytest.mark.parametrize("n,sr,f0,H", [(1000, 44100.0, 1000.0, 10), (4410, 44100.0, 997.0, 20),
                                       (256, 1000.0, 31.0, 40), (300, 1000.0, 30.0, 5)])
def test_batch_thd_matches_oracle(n, sr, f0, H):
    rng = np.random.default_rng(n + H)
    B = 6
    x = _harm_signal(rng, B, n, sr, f0)
    got = SP.compute_batch_thd(WaveformBatch(x, sample_rate=sr), f0, H)
    assert len(got) == B
    for b in range(B):
        _check_thd(got[b], S.compute_thd_from_samples(list(x[b]), sr, f0, H))


def test_batch_thd_resampled_matches_oracle():
    rng = np.random.default_rng(3)
    T = 400
    times = np.cumsum(rng.uniform(0.5, 1.5, T)) * 1e-5
    sr = (T - 1) / (times[-1] - times[0])
    x = _harm_signal(rng, 5, T, sr, 500.0)
    got = SP.compute_batch_thd(WaveformBatch(x, times=times), 500.0, 8)
    for b in range(5):
        _check_thd(got[b], S.compute_thd(list(times), list(x[b]), 500.0, 8))


def test_reference_thd_tests_on_gpu():
    # thd.rs:338-433 through the GPU path
    def sig(amps, sr=44100.0, dur=0.1, f0=1000.0):
        n = int(sr * dur)
        return [sum(a * math.sin(2 * math.pi * f0 * (h + 1) * (i / sr)) for h, a in enumerate(amps))
                for i in range(n)]
    assert SP.compute_thd_from_samples(sig([1.0]), 44100.0, 1000.0, 10).thd_percent < 1.0
    r = SP.compute_thd_from_samples(sig([1.0, 0.1, 0.05]), 44100.0, 1000.0, 10)
    assert abs(r.thd_percent - math.sqrt(0.1 ** 2 + 0.05 ** 2) * 100.0) < 2.0
    assert len(r.harmonics) >= 3 and r.harmonics[0].harmonic_number == 1
    sq = [1.0 if math.sin(2 * math.pi * 1000.0 * (i / 100000.0)) >= 0 else -1.0
          for i in range(10000)]
    assert 30.0 < SP.compute_thd_from_samples(sq, 100000.0, 1000.0, 20).thd_percent < 60.0
    r = SP.compute_thd_from_samples(sig([1.0, 0.1]), 44100.0, 1000.0, 5)
    assert abs(r.harmonics[1].relative_percent - 10.0) < 2.0
    assert abs(r.thd_db - 20.0 * math.log10(r.thd_percent / 100.0)) < 1.0


def test_reference_fft_tests_on_gpu():
    # fft.rs:384-470 through the GPU path
    rect = SpectralConfig(W.RECTANGULAR, None, False)
    x = [math.sin(2 * math.pi * 1000.0 * (i / 10000.0)) for i in range(1000)]
    r = SP.compute_fft_from_samples(x, 10000.0, rect)
    assert abs(r.peak_frequency() - 1000.0) < 20.0 and abs(r.magnitude_at(1000.0) - 1.0) < 0.1
    x = [2.0 * math.sin(2 * math.pi * 500.0 * (i / 8000.0)) for i in range(1024)]
    r = SP.compute_fft_from_samples(x, 8000.0, rect)
    assert abs(r.power_spectral_density.sum() / np.mean(np.square(x)) - 1.0) < 0.15
    r = SP.compute_fft_from_samples([5.0] * 1024, 1000.0, rect)
    assert abs(r.magnitude[0] - 5.0) < 0.01 and r.magnitude[1:].max() < 0.01
    assert SP.compute_fft_from_samples([], 1000.0).frequencies.size == 0


def test_batch_fft_stats_matches_oracle():
    rng = np.random.default_rng(11)
    B, T = 2500, 200    # more waveforms than workgroups: exercises the partial merge
    t = np.arange(T) * 1e-3
    amp = rng.uniform(0.5, 1.5, (B, 1))
    x = amp * np.sin(2 * np.pi * 40 * t) + 0.1 * rng.standard_normal((B, T))
    cfg = SpectralConfig(W.HAMMING)
    got = SP.compute_batch_fft_stats(WaveformBatch(x, sample_rate=1e3), cfg)
    spectra = [S.compute_fft_from_samples(list(x[b]), 1e3, S.HAMMING) for b in range(B)]
    want = S.batch_stats_from_spectra(spectra)
    assert got.num_samples == B
    np.testing.assert_array_equal(got.frequencies, want["frequencies"])
    peak = max(want["max_magnitude"])
    for k in ("mean_magnitude", "min_magnitude", "max_magnitude"):
        np.testing.assert_allclose(getattr(got, k), want[k], rtol=1e-12, atol=1e-12 * peak)
    np.testing.assert_allclose(got.std_magnitude, want["std_magnitude"], rtol=1e-9,
                               atol=1e-12 * peak)


def test_batch_fft_stats_four_step_matches_oracle():
    rng = np.random.default_rng(12)
    B, T = 40, 9000
    x = rng.standard_normal((B, T)) * rng.uniform(0.5, 2.0, (B, 1))
    got = SP.compute_batch_fft_stats(WaveformBatch(x, sample_rate=1e4))
    want = S.batch_stats_from_spectra([S.compute_fft_from_samples(list(x[b]), 1e4)
                                       for b in range(B)])
    peak = max(want["max_magnitude"])
    for k in ("mean_magnitude", "min_magnitude", "max_magnitude"):
        np.testing.assert_allclose(getattr(got, k), want[k], rtol=1e-12, atol=1e-12 * peak)
    np.testing.assert_allclose(got.std_magnitude, want["std_magnitude"], rtol=1e-9,
                               atol=1e-12 * peak)


def test_batch_fft_stats_identical_waveforms():
    # spectral.rs:413-431: identical waveforms have (near) zero spread
    t = np.arange(1000) / 10000.0
    x = np.tile(np.sin(2 * np.pi * 1000.0 * t), (5, 1))
    got = SP.compute_batch_fft_stats(WaveformBatch(x, times=t))
    assert got.num_samples == 5 and got.frequencies.size and got.std_magnitude.max() < 1e-12


@pytest.mark.parametrize("n", [1, 2, 10000, 10001])
def test_thd_statistics_matches_oracle(n):
    rng = np.random.default_rng(n)
    v = rng.gamma(2.0, 2.0, n)
    got = SP.compute_thd_statistics(v, 4.0)
    want = S.thd_statistics(list(v), 4.0)
    assert got.num_samples == n and got.thd_spec == 4.0
    for k in ("mean_thd", "std_thd"):
        assert getattr(got, k) == pytest.approx(want[k], rel=1e-12, abs=1e-15)
    for k in ("min_thd", "max_thd", "median_thd", "yield_percent"):
        assert getattr(got, k) == want[k], k


def test_thd_statistics_reference_cases():
    # spectral.rs:374-440
    st = SP.compute_thd_statistics([1.0, 5.0, 15.0], 10.0)
    assert abs(st.yield_percent - 66.67) < 1.0
    assert abs(SP.compute_thd_statistics([1.0, 5.0, 15.0], 1.0).yield_percent - 33.33) < 1.0
    st = SP.compute_thd_statistics([], 10.0)
    assert st.num_samples == 0 and st.yield_percent == 100.0
    with pytest.raises(E.InvalidArgument):
        SP.compute_thd_statistics([1.0, float("nan")], 10.0)
    t = np.arange(1000) / 100000.0
    x = np.stack([(1.0 + 0.01 * i) * np.sin(2 * np.pi * 1000.0 * t) for i in range(10)])
    r, st = SP.compute_batch_thd_with_stats(WaveformBatch(x, times=t), 1000.0, 5, 10.0)
    assert len(r) == 10 and st.num_samples == 10 and st.mean_thd < 5.0 and st.yield_percent > 90.0


def test_spectral_edges():
    # empty batch, empty spectrum, empty THD, invalid grids and sizes
    r = SP.compute_batch_fft(WaveformBatch(np.zeros((0, 16)), sample_rate=1.0))
    assert len(r) == 0
    r = SP.compute_batch_fft(WaveformBatch(np.ones((3, 1)), times=np.array([0.0])))
    assert r.fft_size == 0 and r.magnitude.shape == (3, 0)
    th = SP.compute_batch_thd(WaveformBatch(np.ones((3, 1)), times=np.array([0.0])), 100.0, 5)
    assert th[2].thd_percent == 0.0 and th[2].thd_db == -math.inf and th[2].harmonics == []
    th = SP.compute_batch_thd(WaveformBatch(np.ones((2, 64)), sample_rate=64.0), 0.0, 5)
    assert th[0].fundamental_freq == 0.0 and th[0].num_harmonics == 0
    # fundamental past Nyquist: no harmonics, the fundamental is the nearest bin
    x = np.sin(2 * np.pi * 5.0 * np.arange(64) / 64.0).reshape(1, -1)
    th = SP.compute_thd_from_samples(x[0], 64.0, 40.0, 3)
    want = S.compute_thd_from_samples(list(x[0]), 64.0, 40.0, 3)
    _check_thd(th, want)
    with pytest.raises(E.InvalidArgument):
        SP.compute_batch_fft(WaveformBatch(np.ones((1, 3)), times=np.array([0.0, 1.0, 1.0])))
    with pytest.raises(E.InvalidDimension):   # past 2^22 points
        SP.compute_batch_fft(WaveformBatch(np.ones((1, 16)), sample_rate=1.0),
                             SpectralConfig(W.HANNING, (1 << 22) + 1, True))


def test_transient_sweep_to_thd_matches_oracle():
    # a sweep's probe waveforms analysed where the transient left them
    from spicier_amd import circuit as CK
    ckt = CK.ladder(10, seed=2)
    res = E.transient_sweep(ckt, 16, 5e-12, 200, probes=[0, 8], seed=3)
    assert np.all(res.status == 1)
    wb = WaveformBatch.from_transient(res, 1)
    f0 = 1.0 / 1e-9
    got = SP.compute_batch_thd(wb, f0, 6)
    spec = SP.compute_batch_fft(WaveformBatch.from_transient(res, 1))
    for b in range(16):
        want = S.compute_thd(list(res.times), list(res.waveforms[b, :, 1]), f0, 6)
        _check_thd(got[b], want)
        _check_spectrum(spec[b], S.compute_fft(list(res.times), list(res.waveforms[b, :, 1])))
