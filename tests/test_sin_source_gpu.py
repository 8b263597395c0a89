"""SIN sources on the GPU (waveforms.rs:185-201): a resistive divider
against the closed form on every kernel kind, and a Monte Carlo diode
clipper driven by a sine through the trapezoidal transient and batch THD,
against the oracle on the same sampled parameters (1e-9 relative /
1e-12 absolute; THD to 1e-9)."""
import math

import numpy as np
import pytest

from oracle import spectral_oracle as S
from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E
from spicier_amd import spectral as SP

pytestmark = pytest.mark.gpu


def _close(a, b, rtol=1e-9, atol=1e-12):
    return np.all(np.abs(np.asarray(a) - np.asarray(b)) <= np.maximum(atol, rtol * np.abs(b)))


@pytest.mark.parametrize("kind", [E.KernelKind.Dense, E.KernelKind.Sparse64, E.KernelKind.Lanes])
def test_sin_divider_closed_form(kind):
    vo, va, f, td, theta, ph = 0.2, 1.1, 4e3, 0.05e-3, 300.0, 20.0
    c = CK.Circuit("sin_divider")
    c.vsin("in", "0", vo, va, f, td, theta, ph)
    c.resistor("in", "out", 1e3)
    c.resistor("out", "0", c.uniform(1e3, 3e3, name="R2"))
    h, steps = 2e-6, 250
    cl = E.TranClass(c, 16, h, steps, probes=[0, 1], seed=1, method=E.IntegrationMethod.Trapezoidal)
    E.transient_mixed_device([cl], E.nr_config(), graph=False, kernels=[kind])
    E.Device.get(0).synchronize()
    r = cl.host_result()
    t = np.arange(steps + 1) * h
    v = vo + va * np.exp(-theta * (t - td)) * np.sin(2 * np.pi * f * (t - td) + ph * np.pi / 180)
    v = np.where(t < td, vo, v)
    v[0] = vo
    for i in range(16):
        np.testing.assert_allclose(r.waveforms[i, :, 0], v, rtol=1e-12, atol=1e-14)
    ratio = r.waveforms[:, 1:, 1] / r.waveforms[:, 1:, 0]
    assert np.all(np.ptp(ratio, axis=1) < 1e-12)


def _clipper():
    c = CK.Circuit("sin_clipper")
    c.vsin("in", "0", 0.0, 2.0, 1e3)
    c.resistor("in", "out", c.uniform(0.8e3, 1.2e3, name="R"))
    c.diode("out", "0", c.loguniform(1e-15, 1e-13, name="IS"), c.const(1.0, "N"))
    c.diode("0", "out", c.loguniform(1e-15, 1e-13, name="IS2"), c.const(1.0, "N2"))
    c.capacitor("out", "0", c.const(1e-9, "C"))
    return c


def test_sin_clipper_transient_and_thd_match_oracle(oracle):
    c = _clipper()
    B, h, steps = 24, 1e-5, 1000       # 10 periods of 1 kHz at 100 points per period
    cfg = E.nr_config(limiter=E.Limiter.Pnjlim)
    out = c.node_names["out"] - 1
    got = E.transient_sweep(c, B, h, steps, probes=[0, out], seed=4, config=cfg,
                            method=E.IntegrationMethod.Trapezoidal)
    want = oracle.transient_ex(c, got.params, A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM),
                               h, steps, probes=[0, out], method=A.SPX_TRAN_TRAP, nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    np.testing.assert_array_equal(got.iterations, want["iterations"])
    assert _close(got.waveforms, want["waveforms"])
    # the clipped output's THD, on the GPU and by the oracle from the oracle's waveforms
    thd = SP.compute_batch_thd(SP.WaveformBatch.from_transient(got, 1), 1e3, 9)
    for i in range(B):
        ref = S.compute_thd(list(got.times), list(want["waveforms"][i, :, 1]), 1e3, 9)
        assert thd[i].thd_percent == pytest.approx(ref["thd_percent"], rel=1e-8)
        assert thd[i].fundamental_freq == ref["fundamental_freq"]
    assert min(thd[i].thd_percent for i in range(B)) > 5.0   # the diodes clip the 2 V sine
