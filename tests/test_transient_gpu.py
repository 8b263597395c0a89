"""GPU parity of the Backward-Euler transient sweep against the oracle's
restatement (transient/solver.rs:41-160, companion.rs:41-66, eval_pulse
waveforms.rs:160-180) on identical seeded parameters, through the C ABI:
the register-LU kernel for n <= 64, the shared-symbolic sparse kernel
above.  Bar: every probed waveform sample and final solution within 1e-9
relative / 1e-12 absolute; status, flags, iteration counts equal."""
import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-9, 1e-12


def close(a, b, rtol=RTOL, atol=ATOL):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) <= np.maximum(atol, rtol * np.abs(b))


def test_rc_step_response_on_gpu():
    R, Cap, V = 2e3, 1.5e-12, 1.0
    c = CK.rc_step(R, Cap, V)
    h, steps = 1e-12, 200
    got = E.transient_sweep(c, 4, h, steps, seed=1)
    k = np.arange(steps + 1, dtype=float)
    want = V * (1.0 - (1.0 + h / (R * Cap)) ** (-k))
    for i in range(4):
        np.testing.assert_allclose(got.waveforms[i, :, 1], want, rtol=1e-12, atol=1e-15)
    assert np.all(got.status == A.SPX_CONVERGED) and np.all(got.fail_step == -1)
    assert np.all(got.iterations == steps + 1)


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 256])
def test_ladder_transient_matches_oracle(oracle, n):
    c = CK.ladder(n, seed=n)
    B = 48 if n <= 64 else 16
    h, steps = 1e-12, 200
    cfg = E.nr_config(limiter=E.Limiter.Pnjlim)
    got = E.transient_sweep(c, B, h, steps, seed=7, config=cfg)
    want = oracle.transient(c, got.params, A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM), h,
                            steps, nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    np.testing.assert_array_equal(got.fail_step, want["fail_step"])
    np.testing.assert_array_equal(got.flags, want["flags"])
    np.testing.assert_array_equal(got.iterations, want["iterations"])
    ok = close(got.waveforms, want["waveforms"])
    assert np.all(ok), (n, np.argwhere(~ok)[:5], np.abs(got.waveforms - want["waveforms"]).max())
    assert np.all(close(got.solutions, want["x"]))
    assert np.all(got.accepted_mask())
    # the source node is the PULSE waveform itself
    t = got.times
    np.testing.assert_allclose(got.waveforms[:, :, 0], np.tile(np.clip(t / 10e-12, 0, 1) * 1.8,
                                                               (B, 1)), rtol=1e-12, atol=1e-15)


def test_transient_probes_and_no_steps(oracle):
    c = CK.ladder(24, seed=5)
    probes = [3, 0, 23]
    got = E.transient_sweep(c, 8, 1e-12, 0, probes=probes, seed=2)
    assert got.waveforms.shape == (8, 1, 3)
    np.testing.assert_array_equal(got.waveforms[:, 0, :], got.solutions[:, probes])
    with pytest.raises(E.InvalidArgument):
        E.transient_sweep(c, 2, 1e-12, 5, probes=[24])
    with pytest.raises(E.InvalidArgument):
        E.transient_sweep(c, 2, 0.0, 5)


# --------------------------------------------------------- FP32 LU + FP64 IR
def _ir_cfg():
    return E.nr_config(limiter=E.Limiter.Pnjlim, precision=E.Precision.F32_IR)


def test_rc_step_response_ir_reuses_factors():
    # linear: the matrix is the same at every step, so the FP32 factors of the
    # first step serve all 200; refinement keeps the FP64 recurrence
    R, Cap, V = 2e3, 1.5e-12, 1.0
    c = CK.rc_step(R, Cap, V)
    h, steps = 1e-12, 200
    got = E.transient_sweep(c, 4, h, steps, seed=1, config=_ir_cfg())
    k = np.arange(steps + 1, dtype=float)
    want = V * (1.0 - (1.0 + h / (R * Cap)) ** (-k))
    for i in range(4):
        np.testing.assert_allclose(got.waveforms[i, :, 1], want, rtol=1e-12, atol=1e-15)
    assert np.all(got.status == A.SPX_CONVERGED) and np.all(got.fail_step == -1)
    assert np.all((got.flags & A.SPX_FLAG_FP64_ESCALATED) == 0)


@pytest.mark.parametrize("n", [8, 16, 32, 64])
def test_ladder_transient_ir_matches_fp64_oracle(oracle, n):
    # FP32 factors carried across steps, FP64 residuals: each Newton step is
    # solved to FP64 backward error, so the iterates follow the FP64 oracle's
    c = CK.ladder(n, seed=n)
    B, h, steps = 48, 1e-12, 200
    got = E.transient_sweep(c, B, h, steps, seed=7, config=_ir_cfg())
    want = oracle.transient(c, got.params, A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM), h,
                            steps, nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    np.testing.assert_array_equal(got.fail_step, want["fail_step"])
    np.testing.assert_array_equal(got.flags & ~A.SPX_FLAG_FP64_ESCALATED, want["flags"])
    same = got.iterations == want["iterations"]
    assert same.mean() >= 0.999, (n, np.count_nonzero(~same))
    ok = close(got.waveforms[same], want["waveforms"][same])
    assert np.all(ok), (n, np.abs(got.waveforms[same] - want["waveforms"][same]).max())
    assert np.all(close(got.solutions[same], want["x"][same]))
    # any other instance (a borderline dx test at some step) at the SPICE tolerance
    assert np.all(np.abs(got.waveforms[~same] - want["waveforms"][~same])
                  <= 1e-3 * np.abs(want["waveforms"][~same]) + 1e-6)
    assert np.all(got.accepted_mask())
