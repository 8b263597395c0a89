"""The dense LU's pivot rule (nalgebra LU::new, spicier-batched-sweep
src/solver.rs:370-380: the largest |a| of the column, the first row on exact
ties) and its factors in LAPACK getrf format.

The oracle's factorisation is pinned against LAPACK's dgetrf (scipy's
lu_factor: idamax, the same first-maximum rule): pivots equal as integers,
factors to rounding.  The GPU's factorisation (spx_lu_factor_batch_f64, the
getrf half of the CUDA backend's solve) is compared with the oracle the same
way.  Two families of systems:
  * near-tie blocks: block-structured matrices, rows of the blocks
    interleaved, whose blocks open with a column of candidates equal or
    within 2^-52 .. 2^-20 relative of each other.  Those columns are
    untouched by earlier steps, so the pivot is decided on the inputs
    themselves, independent of rounding: a rule that compares fewer bits, or
    takes the last maximum, picks another row;
  * gmin-floored MNA systems: conductance graphs with 1e-12 S to ground at
    every node, weakly attached node clusters (down to 1e-9 S), voltage
    source rows (exact +-1 ties) and MOSFET-like one-way transconductances.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from spicier_amd import engine as E

REL = [0.0, 2.0 ** -52, 2.0 ** -40, 2.0 ** -30, 2.0 ** -20, -(2.0 ** -52), -(2.0 ** -20)]


def tie_blocks(n, rng):
    """n x n matrix of independent blocks (sizes 1..4) on random row sets and
    consecutive columns; each block's first column holds near-tied values."""
    M = np.zeros((n, n))
    rows = rng.permutation(n)
    c = 0
    while c < n:
        m = min(int(rng.integers(1, 5)), n - c)
        R = np.sort(rows[c:c + m])
        sub = rng.uniform(0.2, 1.0, (m, m)) * rng.choice([-1.0, 1.0], (m, m))
        base = 2.0 ** int(rng.integers(-30, 30))
        sub[:, 0] = base * (1.0 + rng.choice(REL, m)) * rng.choice([-1.0, 1.0], m)
        M[np.ix_(R, np.arange(c, c + m))] = sub
        c += m
    return M


def gmin_mna(n, rng):
    """A nodal MNA Jacobian: random conductances on a sparse graph, gmin 1e-12
    at every node, a weakly attached cluster, source rows/columns (+-1) and a
    few one-way transconductance entries."""
    nv = int(rng.integers(1, 4)) if n >= 8 else 1
    nn = n - nv
    G = np.zeros((n, n))
    weak = set(rng.choice(nn, size=max(1, nn // 8), replace=False).tolist())
    for i in range(nn):
        for j in rng.choice(nn, size=2, replace=False):
            if i == j:
                continue
            g = 10.0 ** rng.uniform(-6, -2)
            if i in weak or j in weak:
                g = 10.0 ** rng.uniform(-9, -7)
            G[i, i] += g; G[j, j] += g; G[i, j] -= g; G[j, i] -= g
    for i in range(nn):
        G[i, i] += 1e-12
    for k, a in enumerate(rng.choice(nn, size=nv, replace=False)):  # source k: node a to 0
        G[a, nn + k] = 1.0
        G[nn + k, a] = 1.0
    for _ in range(nn // 4):  # gm: current into d controlled by g - s
        d, g_, s_ = rng.choice(nn, size=3, replace=False)
        gm = 10.0 ** rng.uniform(-5, -3)
        G[d, g_] += gm; G[d, s_] -= gm; G[s_, g_] -= gm; G[s_, s_] += gm
    return G


def _cases(kind, n, count, seed):
    rng = np.random.default_rng(seed)
    mk = tie_blocks if kind == "ties" else gmin_mna
    return np.stack([mk(n, rng) for _ in range(count)])


def _lapack(M):
    lu, piv = sl.lu_factor(M, check_finite=True)
    return lu, piv.astype(np.int32)


def _backward_error(M, lu, piv):
    """max |P A - L U| / max |A| of getrf-format factors."""
    n = len(M)
    PA = M.copy()
    for k in range(n):
        PA[[k, piv[k]]] = PA[[piv[k], k]]
    L = np.tril(lu, -1) + np.eye(n)
    return np.abs(PA - L @ np.triu(lu)).max() / np.abs(M).max()


@pytest.mark.parametrize("n", [5, 16, 33, 64, 100])
def test_oracle_factor_matches_lapack(oracle, n):
    for kind in ("ties", "mna"):
        for M in _cases(kind, n, 40, seed=n):
            lu, piv, sing = oracle.lu_factor(M)
            assert not sing
            wlu, wpiv = _lapack(M)
            np.testing.assert_array_equal(piv, wpiv, err_msg=kind)
            assert _backward_error(M, lu, piv) <= 8 * n * 2.0 ** -53, kind
            if kind == "ties":  # no cancellation: the factors themselves agree
                assert np.abs(lu - wlu).max() <= 1e-12 * np.abs(wlu).max(), kind


def test_oracle_rule_is_strict_maximum_then_first_row(oracle):
    # by hand: candidates 1, 1 + 2^-20, 1 + 2^-20 in rows 0, 2, 3 of column 0;
    # the strict maximum is the tie of rows 2 and 3, the first of them wins
    M = np.array([[1.0, 2.0, 0.0, 0.0],
                  [0.0, 1.0, 3.0, 0.0],
                  [1.0 + 2.0 ** -20, 0.0, 1.0, 5.0],
                  [1.0 + 2.0 ** -20, 0.0, 0.0, 1.0]])
    _, piv, sing = oracle.lu_factor(M)
    assert not sing and piv[0] == 2
    _, piv, _ = oracle.lu_factor(M * np.array([[1.0], [1.0], [1.0 - 2.0 ** -52], [1.0]]))
    assert piv[0] == 3
    _, _, sing = oracle.lu_factor(np.array([[1.0, 2.0], [2.0, 4.0]]))
    assert sing  # exact zero pivot in the second column


@pytest.mark.gpu
@pytest.mark.parametrize("n", [3, 8, 13, 31, 32, 48, 64, 65, 100])
def test_gpu_pivots_equal_oracle_on_near_ties(oracle, n):
    M = _cases("ties", n, 400, seed=1000 + n)
    F, piv, sing = E.HipBatchedSolver().factor_batch(M.transpose(0, 2, 1).copy(), n, len(M))
    assert sing == []
    for i, m in enumerate(M):
        lu, p, s = oracle.lu_factor(m)
        np.testing.assert_array_equal(piv[i], p, err_msg=f"system {i}")
        assert np.abs(F[i] - lu).max() <= 1e-12 * np.abs(lu).max(), f"system {i}"


@pytest.mark.gpu
@pytest.mark.parametrize("n", [8, 24, 64, 90])
def test_gpu_gmin_floored_mna_matches_oracle(oracle, n):
    M = _cases("mna", n, 300, seed=2000 + n)
    rng = np.random.default_rng(n)
    b = rng.uniform(-1.0, 1.0, (len(M), n))
    F, piv, sing = E.HipBatchedSolver().factor_batch(M.transpose(0, 2, 1).copy(), n, len(M))
    r = E.HipBatchedSolver().solve_batch(M.transpose(0, 2, 1).copy(), b, n, len(M))
    x = r.solutions.reshape(len(M), n)
    assert sing == [] and r.singular_indices == []
    for i, m in enumerate(M):
        lu, p, _ = oracle.lu_factor(m)
        np.testing.assert_array_equal(piv[i], p, err_msg=f"system {i}")
        assert _backward_error(m, F[i], piv[i]) <= 8 * n * 2.0 ** -53, f"system {i}"
        want, s = oracle.lu_solve(m, b[i])
        assert not s
        # every unknown of every system: within 1e-9 relative of the oracle,
        # or as close to it as LAPACK's own solve (same pivots, other
        # rounding) comes, times 8 — these systems amplify rounding
        lap = sl.lu_solve(_lapack(m), b[i])
        tol = 1e-9 * np.abs(want) + 1e-12 * np.abs(want).max() + 8 * np.abs(lap - want)
        assert np.all(np.abs(x[i] - want) <= tol), f"system {i}"
