"""GPU parity of the sparse (shared-symbolic LU) Newton path, n > 64, against
the CPU oracle on identical seeded inputs, through the C ABI.

The oracle solves each Newton step with its dense nalgebra-style LU
(partial pivoting); the engine factors in a minimum-degree order without
row interchanges.  Both are exact-arithmetic equal, so per-node agreement is
at rounding level; the bar is the north-star 1e-9 relative / 1e-12 V.
The grid workloads use the SPICE junction limiter: with the reference's
memoryless limiter the diodes nearest the 3.3 V corner settle on its
spurious fixed point above vcrit, which the KCL test flags (tested below).
"""
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


def _pnj():
    return E.nr_config(limiter=E.Limiter.Pnjlim)


def _check(got, want, tag):
    np.testing.assert_array_equal(got.status, want["status"], err_msg=tag)
    np.testing.assert_array_equal(got.flags, want["flags"], err_msg=tag)
    np.testing.assert_array_equal(got.iterations, want["iterations"], err_msg=tag)
    ok = close(got.solutions, want["x"])
    assert np.all(ok), (tag, np.argwhere(~ok)[:5], np.abs(got.solutions - want["x"]).max())


def test_grid16_lds_factor_matches_oracle(oracle):
    # n = 260: the filled L+U (4,270 values) fits the workgroup's LDS
    c = CK.grid_diodes(16)
    B = 96
    got = E.sweep_dcop(c, B, seed=42, config=_pnj())
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM),
                             nthreads=16)
    _check(got, want, "grid16")
    assert np.all(got.accepted_mask())


def test_grid32_global_factor_matches_oracle(oracle):
    # n = 1,028: the filled L+U (24,514 values) lives in the global workspace
    c = CK.grid_diodes(32)
    assert c.size == 1028
    B = 8
    got = E.sweep_dcop(c, B, seed=42, config=_pnj())
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM),
                             nthreads=16)
    _check(got, want, "grid32")
    assert np.all(got.accepted_mask())


def test_grid32_reference_limiter_fixed_point_is_flagged(oracle):
    c = CK.grid_diodes(32)
    B = 4
    got = E.sweep_dcop(c, B, seed=42, config=E.nr_config())
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(), nthreads=16)
    _check(got, want, "grid32 reference limiter")
    assert np.all(got.flags & A.SPX_FLAG_KCL_FAIL)


def test_linear_grid_single_solve(oracle):
    # no diodes: one sparse solve per instance (sweep.rs:159-217)
    c = CK.grid_diodes(20, every=10**9)
    assert not c.nonlinear and c.size == 404
    B = 64
    got = E.sweep_dcop(c, B, seed=7)
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(), nthreads=16)
    _check(got, want, "linear grid")
    assert np.all(got.iterations == 1) and np.all(got.accepted_mask())


def test_grid32_full_batch_properties():
    # 4,096 instances: every one converges and passes the FP64 KCL test;
    # statistics agree with a host reduction of the returned solutions
    c = CK.grid_diodes(32)
    B = 4096
    got = E.sweep_dcop(c, B, seed=42, config=_pnj())
    assert got.converged_count == B and np.all(got.accepted_mask())
    x = got.solutions
    np.testing.assert_allclose(got.node_stats[:, 1], x.sum(0), rtol=1e-12)
    np.testing.assert_array_equal(got.node_stats[:, 3], x.min(0))
    np.testing.assert_array_equal(got.node_stats[:, 4], x.max(0))
    # the sources' nodes sit at their voltages; diode nodes clamp below 0.8 V
    N = 32
    np.testing.assert_allclose(x[:, [0, N - 1, N * (N - 1), N * N - 1]],
                               np.tile([1.0, 1.0 + 2.3 / 3, 1.0 + 4.6 / 3, 3.3], (B, 1)),
                               rtol=0, atol=1e-12)
    dn = [k for k in range(N * N) if k % 8 == 4]
    assert x[:, dn].max() < 0.8


def test_sparse_path_rejects_fp32():
    c = CK.grid_diodes(16)
    with pytest.raises(E.InvalidArgument):
        E.sweep_dcop(c, 4, config=E.nr_config(precision=E.Precision.F32_IR))
