"""GPU parity of the batched GMRES-Jacobi kernel against the oracle's
restatement of solve_gmres_real_preconditioned (same algorithm, sequential
dot products) and against LAPACK solutions of the same systems."""
import numpy as np
import pytest

from spicier_amd import engine as E

pytestmark = pytest.mark.gpu


def _family(rng, n, B, extra=3, dom=0.6):
    """B systems with one random pattern (row i: diagonal + `extra` random
    off-diagonals + a duplicated (0, 0) triplet) and per-system values."""
    rows, cols = [], []
    for i in range(n):
        for j in rng.choice(n, size=extra, replace=False):
            if j != i:
                rows.append(i); cols.append(int(j))
        rows.append(i); cols.append(i)
    rows.append(0); cols.append(0)
    rows, cols = np.array(rows), np.array(cols)
    vals = rng.standard_normal((B, rows.size))
    diag = rows == cols
    vals[:, diag] = dom * (extra + 1) + rng.random((B, int(diag.sum())))
    return rows, cols, vals


def _dense(n, rows, cols, v):
    m = np.zeros((n, n))
    np.add.at(m, (rows, cols), v)
    return m


@pytest.mark.parametrize("restart", [30, 7])
def test_gmres_matches_oracle_and_lapack(oracle, restart):
    rng = np.random.default_rng(restart)
    n, B = 240, 24
    rows, cols, vals = _family(rng, n, B)
    b = rng.standard_normal((B, n))
    cfg = E.GmresConfig(restart=restart, tol=1e-10)
    r = E.HipBatchedGmres(cfg).solve_batch_triplets(rows, cols, vals, b, n)
    assert np.all(r.converged)
    same_it = 0
    for i in range(B):
        xo, conv, it, res = oracle.gmres(n, rows, cols, vals[i], b[i], restart=restart,
                                         tol=1e-10)
        assert conv
        same_it += it == r.iterations[i]
        assert abs(int(r.iterations[i]) - it) <= 1, (i, r.iterations[i], it)
        m = _dense(n, rows, cols, vals[i])
        x = r.solution(i)
        assert np.linalg.norm(b[i] - m @ x) / np.linalg.norm(b[i]) < 1e-10
        assert r.residual[i] < 1e-10
        # same Krylov iteration: agreement far inside the tolerance
        np.testing.assert_allclose(x, xo, rtol=1e-7, atol=1e-9)
    assert same_it >= B - 2


def test_gmres_edge_cases(oracle):
    n = 64
    idx = np.arange(n)
    d = np.linspace(1.0, 1e5, n)
    b = np.stack([np.linspace(-1, 2, n), np.zeros(n)])
    vals = np.stack([d, d])
    r = E.HipBatchedGmres(E.GmresConfig(tol=1e-14)).solve_batch_triplets(idx, idx, vals, b, n)
    # diagonal system: Jacobi solves it in one step; zero rhs: no step, x = 0
    assert list(r.iterations) == [1, 0] and np.all(r.converged)
    np.testing.assert_allclose(r.solution(0), b[0] / d, rtol=1e-14)
    assert np.all(r.solution(1) == 0)
    # iteration budget: not converged, count includes the step that trips it
    rng = np.random.default_rng(9)
    rows, cols, vals = _family(rng, 150, 2, dom=0.3)
    bb = rng.standard_normal((2, 150))
    r = E.HipBatchedGmres(E.GmresConfig(restart=4, max_iter=6, tol=1e-14)).solve_batch_triplets(
        rows, cols, vals, bb, 150)
    xo, conv, it, res = oracle.gmres(150, rows, cols, vals[0], bb[0], restart=4, max_iter=6,
                                     tol=1e-14)
    assert not conv and it == 7 and list(r.iterations) == [7, 7] and not r.converged.any()
    np.testing.assert_allclose(r.solution(0), xo, rtol=1e-9, atol=1e-12)


def test_gmres_mna_grid_with_warm_start():
    # 16x16 conductance grid + 4 ideal corner sources: a saddle-point MNA
    # system (zero-diagonal branch rows take Jacobi weight 1)
    from test_sparse_gpu import _dense as dense3, _grid_mna
    rng = np.random.default_rng(4)
    B = 8
    n, rows, cols, vals = _grid_mna(16, rng, B)
    b = np.zeros((B, n))
    b[:, 256:] = [1.0, 1.5, 2.0, 3.3]
    g = E.HipBatchedGmres(E.GmresConfig(tol=1e-12, max_iter=5000))
    r = g.solve_batch_triplets(rows, cols, vals, b, n)
    m = dense3(n, rows, cols, vals)
    want = np.linalg.solve(m, b[..., None])[..., 0]
    assert np.all(r.converged)
    x = r.solutions.reshape(B, n)
    # the stopping test is a backward error; the forward error carries the
    # condition number of the saddle-point system (~1e4 here)
    resid = np.linalg.norm(b - np.einsum("bij,bj->bi", m, x), axis=1) / np.linalg.norm(b, axis=1)
    assert np.all(resid < 1e-12)
    np.testing.assert_allclose(x, want, rtol=1e-7, atol=1e-9)
    # starting from the solution converges at once
    r2 = g.solve_batch_triplets(rows, cols, vals, b, n, x0=want)
    assert np.all(r2.iterations <= 1) and np.all(r2.converged)


# ------------------------------------------- FP32 Krylov inside FP64 refinement
@pytest.mark.parametrize("tol", [1e-8, 1e-12])
def test_gmres_fp32_ir_reaches_fp64_residual(tol):
    # same stopping test on the FP64 true residual as the FP64 kernel; the
    # Krylov vectors, inner products and rotations are FP32
    from test_sparse_gpu import _dense as dense3, _grid_mna
    rng = np.random.default_rng(6)
    B = 12
    n, rows, cols, vals = _grid_mna(16, rng, B)
    b = np.zeros((B, n))
    b[:, 256:] = [1.0, 1.5, 2.0, 3.3]
    b[:, :256] += 1e-4 * rng.standard_normal((B, 256))
    cfg = E.GmresConfig(tol=tol, max_iter=20000, precision=E.Precision.F32_IR)
    r = E.HipBatchedGmres(cfg).solve_batch_triplets(rows, cols, vals, b, n)
    m = dense3(n, rows, cols, vals)
    x = r.solutions.reshape(B, n)
    resid = np.linalg.norm(b - np.einsum("bij,bj->bi", m, x), axis=1) / np.linalg.norm(b, axis=1)
    assert np.all(r.converged) and np.all(resid < tol), resid.max()
    np.testing.assert_allclose(r.residual, resid, rtol=1e-6, atol=1e-3 * tol)
    want = np.linalg.solve(m, b[..., None])[..., 0]
    err = np.abs(x - want).max(axis=1) / np.abs(want).max(axis=1)
    assert np.all(err < 1e5 * tol), err.max()   # condition number ~1e4


def test_gmres_fp32_ir_random_family_and_zero_rhs():
    rng = np.random.default_rng(8)
    n, B = 150, 30
    rows, cols, vals = _family(rng, n, B)
    b = rng.standard_normal((B, n))
    b[3] = 0.0
    cfg = E.GmresConfig(tol=1e-10, precision=E.Precision.F32_IR)
    r = E.HipBatchedGmres(cfg).solve_batch_triplets(rows, cols, vals, b, n)
    assert np.all(r.converged)
    assert np.all(r.solutions.reshape(B, n)[3] == 0) and r.iterations[3] == 0
    for i in range(B):
        m = _dense(n, rows, cols, vals[i])
        x = r.solution(i)
        assert np.linalg.norm(b[i] - m @ x) <= 1e-10 * max(np.linalg.norm(b[i]), 1e-300) or i == 3
