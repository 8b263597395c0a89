"""GPU GMRES-Jacobi(30) on the Newton Jacobians of the 32x32 grid + diodes
workload (n = 1,028), against the oracle's restatement of
solve_gmres_real_preconditioned (gmres/real.rs:208-400) and against the
Newton solution itself.

The systems are the oracle's own final Newton systems J(x*) x = b(x*) of
grid instances (limiter off at the solution): the companion models are
linearised at x*, so x* solves them exactly and every solver must return it
to the system's conditioning.  Same Krylov method as the oracle: iteration
counts agree to one and the solutions to far inside the tolerance."""
import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid_systems(oracle):
    c = CK.grid_diodes(32)
    B = 6
    params = oracle.generate_params(c, B, seed=42)
    cfg = A.default_nr_config()
    cfg.limiter = A.SPX_LIMIT_PNJLIM
    ref = oracle.sweep_dcop(c, params, cfg, nthreads=B)
    assert np.all(ref["status"] == A.SPX_CONVERGED)
    mats, rhs = [], []
    for i in range(B):
        J, b = oracle.stamp(c, params[i], ref["x"][i], limiter=A.SPX_LIMIT_NONE)
        mats.append(J)
        rhs.append(b)
    pattern = np.zeros_like(mats[0], dtype=bool)
    for J in mats:
        pattern |= J != 0
    rows, cols = np.nonzero(pattern)
    vals = np.stack([J[rows, cols] for J in mats])
    return c.size, rows.astype(np.int32), cols.astype(np.int32), vals, np.stack(rhs), \
        np.stack(mats), ref["x"]


def test_grid_jacobians_gmres_matches_oracle(oracle, grid_systems):
    n, rows, cols, vals, b, mats, xstar = grid_systems
    assert n == 1028
    B = len(b)
    cfg = E.GmresConfig(restart=30, max_iter=5000, tol=1e-10)
    r = E.HipBatchedGmres(cfg).solve_batch_triplets(rows, cols, vals, b, n)
    assert np.all(r.converged)
    for i in range(B):
        xo, conv, it, res = oracle.gmres(n, rows, cols, vals[i], b[i], restart=30,
                                         max_iter=5000, tol=1e-10)
        assert conv
        assert abs(int(r.iterations[i]) - it) <= 1, (i, r.iterations[i], it)
        x = r.solution(i)
        # the stopping test is the relative true residual
        assert np.linalg.norm(b[i] - mats[i] @ x) / np.linalg.norm(b[i]) < 1e-10
        assert r.residual[i] < 1e-10
        # same Krylov iteration as the oracle's
        np.testing.assert_allclose(x, xo, rtol=1e-6, atol=1e-9)
        # and the Newton solution the system was linearised at (the grid's
        # condition number is ~1e4, so 1e-10 backward error leaves ~1e-6 V)
        np.testing.assert_allclose(x, xstar[i], rtol=1e-5, atol=1e-6)


def test_grid_jacobians_gmres_fp32_ir(grid_systems):
    n, rows, cols, vals, b, mats, xstar = grid_systems
    cfg = E.GmresConfig(restart=30, max_iter=20000, tol=1e-10, precision=E.Precision.F32_IR)
    r = E.HipBatchedGmres(cfg).solve_batch_triplets(rows, cols, vals, b, n)
    assert np.all(r.converged)
    x = r.solutions.reshape(len(b), n)
    resid = np.linalg.norm(b - np.einsum("bij,bj->bi", mats, x), axis=1) / np.linalg.norm(b, axis=1)
    assert np.all(resid < 1e-10), resid.max()
    np.testing.assert_allclose(x, xstar, rtol=1e-5, atol=1e-6)
