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


def test_grid24_dense_root_block_matches_oracle(oracle):
    # n = 580: the order's last 39 levels are single columns forming the dense
    # root block, which the kernel factors and solves in one wave's registers
    c = CK.grid_diodes(24)
    assert c.size == 580
    B = 8
    got = E.sweep_dcop(c, B, seed=5, config=_pnj())
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM),
                             nthreads=16)
    _check(got, want, "grid24")
    assert np.all(got.accepted_mask())


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


# ------------------------------ FP32 factors + FP64 refinement on the sparse path
def _ir(**kw):
    return E.nr_config(limiter=E.Limiter.Pnjlim, precision=E.Precision.F32_IR, **kw)


def _parity_ir(got, want, tag):
    """The refinement form against an FP64 result: same statuses; the same
    Newton iteration count on >= 99.9 % of instances, whose solutions agree
    to 1e-9 / 1e-12, the others at the SPICE tolerance; the same flags apart
    from FP64_ESCALATED, which at most 1 % carry."""
    np.testing.assert_array_equal(got.status, want["status"], err_msg=tag)
    same_it = got.iterations == want["iterations"]
    assert same_it.mean() >= 0.999, (tag, np.count_nonzero(~same_it))
    ok = close(got.solutions, want["x"])
    assert np.all(ok[same_it]), (tag, np.argwhere(~ok[same_it])[:5])
    if (~same_it).any():
        assert np.all(close(got.solutions[~same_it], want["x"][~same_it], rtol=1e-3, atol=1e-6))
    np.testing.assert_array_equal(got.flags & ~A.SPX_FLAG_FP64_ESCALATED, want["flags"],
                                  err_msg=tag)
    esc = np.count_nonzero(got.flags & A.SPX_FLAG_FP64_ESCALATED)
    assert esc <= 0.01 * len(got.flags), (tag, "escalated to FP64 on", esc)


@pytest.mark.parametrize("N,B", [(16, 96), (24, 16), (32, 8)])
def test_grid_fp32_ir_matches_oracle(oracle, N, B):
    # FP32 L+U values in LDS (grid32: 23,668 values, 95 KB), refined in FP64
    # against the FP64 assembled Jacobian
    c = CK.grid_diodes(N)
    got = E.sweep_dcop(c, B, seed=42, config=_ir())
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM),
                             nthreads=16)
    _parity_ir(got, want, f"grid{N} FP32+IR")
    assert np.all(got.accepted_mask())
    assert np.all(got.refinements > 0)
    assert np.all(got.factorizations <= got.iterations)


def test_grid32_fp32_ir_full_batch_matches_fp64():
    c = CK.grid_diodes(32)
    B = 4096
    f64 = E.sweep_dcop(c, B, seed=42, config=_pnj())
    ir = E.sweep_dcop(c, B, seed=42, config=_ir())
    assert np.all(f64.accepted_mask()) and np.all(ir.accepted_mask())
    want = {"status": f64.status, "iterations": f64.iterations, "flags": f64.flags,
            "x": f64.solutions}
    _parity_ir(ir, want, "grid32 FP32+IR vs FP64, 4096")
    assert np.all(ir.residual <= 1.0)


def test_grid40_fp32_ir_global_factors_match_fp64():
    # n = 1,604: the FP32 L+U values (41,436, 166 KB) exceed a workgroup's LDS
    # and live in the global slot; against the FP64 sweep (oracle-checked at
    # the smaller grids)
    c = CK.grid_diodes(40)
    B = 64
    f64 = E.sweep_dcop(c, B, seed=9, config=_pnj())
    ir = E.sweep_dcop(c, B, seed=9, config=_ir())
    assert np.all(f64.accepted_mask()) and np.all(ir.accepted_mask())
    want = {"status": f64.status, "iterations": f64.iterations, "flags": f64.flags,
            "x": f64.solutions}
    _parity_ir(ir, want, "grid40 FP32+IR vs FP64")


@pytest.mark.parametrize("prec", [E.Precision.F64, E.Precision.F32_IR])
def test_grid_singular_jacobian_takes_gmin_stepping(oracle, prec):
    # a node reached only through a capacitor (open at DC) leaves the plain
    # Newton Jacobian singular: a zero pivot, then Gmin stepping converges.
    # The refinement form hands every instance to the FP64 kernel at its
    # zero FP32 pivot, so its results are the FP64 path's.
    c = CK.grid_diodes(16)
    c.node("fl")
    c.capacitor("fl", "n3_3", c.const(1e-12, "Cf"))
    assert c.size == 261
    B = 16
    got = E.sweep_dcop(c, B, seed=5, config=E.nr_config(limiter=E.Limiter.Pnjlim,
                                                         precision=prec))
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM),
                             nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    np.testing.assert_array_equal(got.iterations, want["iterations"])
    np.testing.assert_array_equal(got.flags & ~A.SPX_FLAG_FP64_ESCALATED, want["flags"])
    assert np.all(got.flags & A.SPX_FLAG_GMIN_STEPPING)
    ok = close(got.solutions, want["x"])
    assert np.all(ok), np.abs(got.solutions - want["x"]).max()
    if prec == E.Precision.F32_IR:
        assert np.all(got.flags & A.SPX_FLAG_FP64_ESCALATED)


def test_grid_fp32_ir_handover_reruns_in_fp64():
    # one refinement step cannot meet the backward-error test, so every
    # instance is handed to the FP64 kernel: its results are the FP64 run's,
    # bit for bit, flagged FP64_ESCALATED
    c = CK.grid_diodes(16)
    B = 64
    f64 = E.sweep_dcop(c, B, seed=3, config=_pnj())
    ir = E.sweep_dcop(c, B, seed=3, config=_ir(ir_max_steps=1))
    np.testing.assert_array_equal(ir.status, f64.status)
    np.testing.assert_array_equal(ir.iterations, f64.iterations)
    np.testing.assert_array_equal(ir.solutions, f64.solutions)
    np.testing.assert_array_equal(ir.flags, f64.flags | A.SPX_FLAG_FP64_ESCALATED)


@pytest.mark.parametrize("prec", [E.Precision.F64, E.Precision.F32_IR])
def test_grid_stepping_fallbacks_match_oracle(oracle, prec):
    # a 4-iteration budget makes plain NR fail on the grid, so Gmin (then
    # source) stepping runs through the sparse kernels' saved-solution slot
    c = CK.grid_diodes(16)
    B = 16
    for kw in ({"gmin_stepping": True, "source_stepping": True},
               {"gmin_stepping": False, "source_stepping": True}):
        cfg = E.nr_config(E.ConvergenceCriteria(max_iterations=4), limiter=E.Limiter.Pnjlim,
                          precision=prec, **kw)
        got = E.sweep_dcop(c, B, seed=11, config=cfg)
        ocfg = A.default_nr_config(max_iterations=4, limiter=A.SPX_LIMIT_PNJLIM,
                                   gmin_stepping=int(kw["gmin_stepping"]))
        want = oracle.sweep_dcop(c, got.params, ocfg, nthreads=16)
        _parity_ir(got, want, f"grid16 stepping {prec.name} {kw}")
        assert np.all(got.flags & (A.SPX_FLAG_GMIN_STEPPING | A.SPX_FLAG_SOURCE_STEPPING))
    # every level fails: each failed level counts all its attempts
    # (newton.rs:300-315, 411-426), 1 + 5 + 5 x 5 = 31 iterations
    cfg = E.nr_config(E.ConvergenceCriteria(max_iterations=1), limiter=E.Limiter.Pnjlim,
                      precision=prec)
    got = E.sweep_dcop(c, B, seed=11, config=cfg)
    want = oracle.sweep_dcop(c, got.params,
                             A.default_nr_config(max_iterations=1, limiter=A.SPX_LIMIT_PNJLIM),
                             nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    np.testing.assert_array_equal(got.iterations, want["iterations"])
    np.testing.assert_array_equal(got.iterations, 31)


def test_linear_grid_fp32_ir_single_solve(oracle):
    # no diodes: one refined solve per instance
    c = CK.grid_diodes(20, every=10**9)
    B = 64
    got = E.sweep_dcop(c, B, seed=7, config=E.nr_config(precision=E.Precision.F32_IR))
    want = oracle.sweep_dcop(c, got.params, A.default_nr_config(), nthreads=16)
    _parity_ir(got, want, "linear grid FP32+IR")
    assert np.all(got.iterations == 1) and np.all(got.accepted_mask())


# ------------------------------------------------- batched sparse linear solve
def _grid_mna(N, rng, B):
    """B systems of an N x N conductance grid with four ideal corner sources
    (zero-diagonal branch rows), one shared triplet pattern with duplicates."""
    n = N * N + 4
    rows, cols, vals = [], [], []
    g = rng.uniform(0.9e-3, 1.1e-3, size=(B, 2 * N * (N - 1)))
    e = 0
    for r in range(N):
        for c in range(N):
            k = r * N + c
            for nb in ((k + 1) if c + 1 < N else None, (k + N) if r + 1 < N else None):
                if nb is None:
                    continue
                for (i, j, s) in ((k, k, 1), (nb, nb, 1), (k, nb, -1), (nb, k, -1)):
                    rows.append(i); cols.append(j); vals.append(s * g[:, e])
                e += 1
    for s, node in enumerate((0, N - 1, N * (N - 1), N * N - 1)):
        for (i, j) in ((node, N * N + s), (N * N + s, node)):
            rows.append(i); cols.append(j); vals.append(np.ones(B))
    return n, np.array(rows), np.array(cols), np.stack(vals, axis=1)


def _dense(n, rows, cols, vals):
    m = np.zeros((vals.shape[0], n, n))
    for t in range(rows.size):
        m[:, rows[t], cols[t]] += vals[:, t]
    return m


def test_sparse_solve_matches_lapack():
    rng = np.random.default_rng(3)
    B = 40
    n, rows, cols, vals = _grid_mna(14, rng, B)   # duplicates: every diagonal is summed
    b = rng.standard_normal((B, n))
    r = E.HipSparseBatchedSolver().solve_batch_triplets(rows, cols, vals, b, n)
    m = _dense(n, rows, cols, vals)
    want = np.linalg.solve(m, b[..., None])[..., 0]
    assert r.singular_indices == []
    x = r.solutions.reshape(B, n)
    # two backward-stable eliminations in different orders agree normwise to
    # about cond * eps; the residual pins the engine's own solution
    err = np.abs(x - want).max(axis=1) / np.abs(want).max(axis=1)
    assert np.all(err < 1e-10), err.max()
    resid = np.abs(b - np.einsum("bij,bj->bi", m, x)).max(axis=1)
    assert np.all(resid < 1e-12 * (np.abs(m).sum(axis=2).max(axis=1) * np.abs(x).max(axis=1))), resid.max()


@pytest.mark.parametrize("N", [24, 32, 40])
def test_sparse_solve_dense_root_block_matches_lapack(N):
    # grids whose order ends in a dense root block (39 and 48 columns; at
    # N = 40 the last 48 of a longer one-column chain): the block's
    # factorization and both substitutions run in registers
    rng = np.random.default_rng(N)
    B = 6
    n, rows, cols, vals = _grid_mna(N, rng, B)
    b = rng.standard_normal((B, n))
    r = E.HipSparseBatchedSolver().solve_batch_triplets(rows, cols, vals, b, n)
    m = _dense(n, rows, cols, vals)
    want = np.linalg.solve(m, b[..., None])[..., 0]
    assert r.singular_indices == []
    x = r.solutions.reshape(B, n)
    err = np.abs(x - want).max(axis=1) / np.abs(want).max(axis=1)
    assert np.all(err < 1e-9), err.max()
    resid = np.abs(b - np.einsum("bij,bj->bi", m, x)).max(axis=1)
    assert np.all(resid < 1e-12 * (np.abs(m).sum(axis=2).max(axis=1) * np.abs(x).max(axis=1))), resid.max()


def test_sparse_solve_singular_and_edge_cases():
    s = E.HipSparseBatchedSolver()
    rows, cols = np.array([0, 1, 2, 1]), np.array([0, 1, 2, 2])
    vals = np.array([[2.0, 4.0, 8.0, 1.0],     # regular
                     [2.0, 0.0, 8.0, 1.0],     # exact zero pivot
                     [1e300, 1e300, 1e-300, 1.0]])
    b = np.ones((3, 3))
    r = s.solve_batch_triplets(rows, cols, vals, b, 3)
    np.testing.assert_allclose(r.solution(0), [0.5, (1 - 1 / 8) / 4, 1 / 8], rtol=1e-15)
    assert r.singular_indices == [1, 2] or r.singular_indices == [1]
    assert np.all(r.solution(1) == 0)
    r = s.solve_batch_triplets(rows, cols, np.zeros((0, 4)), np.zeros((0, 3)), 3)
    assert r.batch_size == 0
    with pytest.raises(E.InvalidArgument):
        s.solve_batch_triplets(np.array([0, 5]), np.array([0, 0]), np.ones((1, 2)), np.ones(3), 3)


def test_sparse_dense_input_pattern_semantics():
    # FaerSparseCachedBatchedSolver: pattern from the first matrix (|a| > 1e-15);
    # a later matrix with a nonzero outside it is reported singular
    rng = np.random.default_rng(5)
    B, n = 6, 30
    m = np.zeros((B, n, n))
    for i in range(n):
        m[:, i, i] = 4.0 + rng.random(B)
        if i + 1 < n:
            m[:, i, i + 1] = m[:, i + 1, i] = -1.0
    m[3, 0, n - 1] = 0.5                       # outside the first matrix's pattern
    b = rng.standard_normal((B, n))
    cm = np.ascontiguousarray(np.transpose(m, (0, 2, 1))).reshape(-1)
    s = E.HipSparseBatchedSolver()
    r = s.solve_batch(cm, b.reshape(-1), n, B)
    assert s.has_cached_symbolic() and r.singular_indices == [3]
    want = np.linalg.solve(m, b[..., None])[..., 0]
    ok = [i for i in range(B) if i != 3]
    assert np.all(close(r.solutions.reshape(B, n)[ok], want[ok], rtol=1e-12, atol=1e-14))
    assert np.all(r.solution(3) == 0)
