"""The oracle's restatement of the reference's CachedSparseLu
(spicier-solver src/linear.rs:334-409), used by its Newton and transient
solves at n >= SPARSE_THRESHOLD = 50 (linear.rs:22, newton.rs:92-108,
transient/solver.rs:131-141): pinned against LAPACK (numpy.linalg.solve) on
random sparse systems, MNA grids with voltage-source rows (zero diagonals
that need row pivoting), duplicate triplets, singular systems and the
reference's own sparse tests; and the oracle's Newton with the size rule
against the same Newton with dense LU at every size."""
import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK


def _triplets(M):
    r, c = np.nonzero(M)
    return r, c, M[r, c]


def test_reference_sparse_known_answers(oracle):
    # faer_sparse_solver.rs:368-398 (identity), 400-418 ([[2,1],[1,3]] x = [5,10]
    # -> [1,3]) and 420-452 (the triplet batch, second system scaled by 2)
    x, s = oracle.sparse_lu_solve(2, [0, 1], [0, 1], [1.0, 1.0], [3.0, 4.0])
    assert not s and np.allclose(x, [3.0, 4.0], atol=1e-10)
    for k in (1.0, 2.0):
        x, s = oracle.sparse_lu_solve(2, [0, 0, 1, 1], [0, 1, 0, 1],
                                      [2.0 * k, 1.0 * k, 1.0 * k, 3.0 * k], [5.0 * k, 10.0 * k])
        assert not s and np.allclose(x, [1.0, 3.0], atol=1e-10)


@pytest.mark.parametrize("n", [1, 5, 50, 200, 600])
def test_random_sparse_against_lapack(oracle, n):
    rng = np.random.default_rng(n)
    for _ in range(5):
        M = np.zeros((n, n))
        k = max(1, 3 * n)
        M[rng.integers(0, n, k), rng.integers(0, n, k)] = rng.uniform(-1, 1, k)
        M += np.diag(rng.uniform(0.1, 1.0, n)) * rng.choice([-1.0, 1.0], n)
        b = rng.uniform(-1, 1, n)
        x, s = oracle.sparse_lu_solve(n, *_triplets(M), b)
        want = np.linalg.solve(M, b)
        assert not s
        assert np.abs(x - want).max() <= 1e-10 * (1 + np.abs(want).max())
        assert np.abs(M @ x - b).max() <= 1e-12 * (np.abs(M).max() * np.abs(x).max() + 1)


def test_mna_grid_with_source_rows_and_duplicates(oracle):
    # the 32x32 grid's Jacobian family: branch rows have zero diagonals, so
    # the factorisation must pivot rows; duplicates (each conductance stamps
    # four entries that overlap at the nodes) are summed
    n, rows, cols, vals, rhs = CK.grid_linear_batch(16, 3, seed=2)
    for i in range(3):
        M = np.zeros((n, n))
        np.add.at(M, (rows, cols), vals[i])
        assert np.any(np.diag(M) == 0)
        x, s = oracle.sparse_lu_solve(n, rows, cols, vals[i], rhs[i])
        want = np.linalg.solve(M, rhs[i])
        assert not s and np.abs(x - want).max() <= 1e-9 * np.abs(want).max()


def test_singular_systems(oracle):
    x, s = oracle.sparse_lu_solve(3, [0, 1, 0, 1], [0, 0, 1, 1], [1.0, 2.0, 2.0, 4.0],
                                  [1.0, 1.0, 1.0])
    assert s and np.all(x == 0)  # rank 1 on two rows, third row empty
    M = np.array([[1.0, 2.0], [2.0, 4.0]])
    x, s = oracle.sparse_lu_solve(2, *_triplets(M), [1.0, 2.0])
    assert s and np.all(x == 0)


def test_newton_size_rule_matches_dense_newton(oracle):
    # the 64-unknown inverter chain and a 64-unknown ladder cross the
    # threshold: the sparse-LU Newton equals the dense-LU Newton to rounding
    for c, B in ((CK.inverter_chain(), 64), (CK.ladder(64, seed=3), 16)):
        p = oracle.generate_params(c, B)
        cfg = A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM)
        try:
            oracle.set_linear_solver(1)
            dense = oracle.sweep_dcop(c, p, cfg, nthreads=4)
        finally:
            oracle.set_linear_solver(0)
        sparse = oracle.sweep_dcop(c, p, cfg, nthreads=4)
        np.testing.assert_array_equal(sparse["status"], dense["status"])
        assert np.mean(sparse["iterations"] == dense["iterations"]) >= 0.95
        same = sparse["iterations"] == dense["iterations"]
        d = np.abs(sparse["x"] - dense["x"])
        assert np.all(d[same] <= np.maximum(1e-12, 1e-9 * np.abs(dense["x"][same])))
        assert np.all(d <= 1e-3 * np.abs(dense["x"]) + 1e-6)
