"""GPU parity of the sweep bookkeeping and result reductions
(spicier_amd/postproc.py over include/spicier_hip.h) against the CPU
restatement in oracle/postproc_oracle.py: bit-exact for statuses, counts,
indices, histograms and packed layouts; sums of streaming statistics to
1e-12 relative (the device reduces in blocks, the reference sequentially)."""
import math

import numpy as np
import pytest

import postproc_oracle as P
from spicier_amd import _abi as A
from spicier_amd import engine as E
from spicier_amd import postproc as PP

pytestmark = pytest.mark.gpu


def _status(rng, B, p=(0.5, 0.3, 0.15, 0.05)):
    return rng.choice([P.ACTIVE, P.CONVERGED, P.FAILED, P.SINGULAR], size=B, p=p).astype(np.int32)


def _tracker(status, iterations=None, max_iterations=50):
    import torch
    t = PP.ConvergenceTracker(len(status), max_iterations)
    t.status_tensor.copy_(torch.as_tensor(status))
    if iterations is not None:
        t.iterations_tensor.copy_(torch.as_tensor(np.asarray(iterations, dtype=np.int32)))
    torch.cuda.synchronize()
    return t


def _host_status(t):
    t.device.synchronize()
    return t.status_tensor.cpu().numpy().tolist()


def test_tracker_reference_known_answers():
    t = PP.ConvergenceTracker(10)
    assert t.mark_converged(0) and t.mark_converged(5) and t.mark_converged(9)
    assert t.active_count() == 7 and t.converged_count() == 3
    assert t.status(0) == E.ConvergenceStatus.Converged and not t.mark_converged(0)
    t.mark_failed(3)
    t.mark_singular(7)
    assert t.failed_count() == 2 and t.status(7) == E.ConvergenceStatus.Singular
    t = PP.ConvergenceTracker.with_max_iterations(5, 3)
    for _ in range(3):
        t.increment_all_active()
    assert t.all_finished() and t.failed_count() == 5
    t = PP.ConvergenceTracker(5)
    t.mark_converged(1)
    t.mark_failed(3)
    assert t.active_indices() == [0, 2, 4]
    assert t.active_mask() == [True, False, True, False, True]
    assert t.active_mask_u32() == [1, 0, 1, 0, 1]
    t = PP.ConvergenceTracker(3)
    assert t.check_convergence([1e-8, 1e-8, 1.0, 1.0, 1e-9, 1e-9], 2, 1e-6, 1e-6,
                               [1.0, 2.0, 10.0, 20.0, 100.0, 200.0]) == 2
    assert [t.status(i) for i in range(3)] == [E.ConvergenceStatus.Converged,
                                               E.ConvergenceStatus.Active,
                                               E.ConvergenceStatus.Converged]
    t = PP.ConvergenceTracker(100)
    for i in range(50):
        for _ in range(3):
            t.increment_iteration(i)
        t.mark_converged(i)
    for i in range(50, 60):
        for _ in range(10):
            t.increment_iteration(i)
        t.mark_converged(i)
    for i in range(60, 65):
        t.mark_failed(i)
    s = t.summary()
    assert (s.total_points, s.converged_count, s.failed_count, s.active_count,
            s.max_iterations) == (100, 60, 5, 35, 10)
    assert "60/100" in str(s)


@pytest.mark.parametrize("B,n", [(1, 1), (5000, 37), (70001, 3), (3000, 130)])
def test_check_convergence_matches_oracle(B, n):
    rng = np.random.default_rng(B + n)
    st = _status(rng, B)
    x = rng.standard_normal(B * n) * 10.0 ** rng.integers(-3, 3, B * n)
    d = np.abs(x) * 1e-3 * rng.uniform(0.5, 1.5, B * n) + rng.uniform(0, 2e-6, B * n)
    d[rng.integers(0, B * n, 7)] = np.nan
    want, newly = P.check_convergence(st.tolist(), d, x, n, 1e-6, 1e-3)
    t = _tracker(st)
    assert t.check_convergence(d, n, 1e-6, 1e-3, x) == newly
    assert _host_status(t) == want


@pytest.mark.parametrize("B,n", [(4097, 5), (1000, 64), (300, 200)])
def test_check_residual_matches_oracle(B, n):
    rng = np.random.default_rng(7 * B + n)
    st = _status(rng, B)
    r = rng.standard_normal(B * n)
    # tolerance at the median norm: about half the instances sit on either side
    norms = np.sqrt((r.reshape(B, n) ** 2).sum(1))
    tol = float(np.median(norms))
    want, newly = P.check_residual(st.tolist(), r, n, tol)
    t = _tracker(st)
    assert t.check_residual_convergence(r, n, tol) == newly
    assert _host_status(t) == want


def test_increment_mark_and_summary_match_oracle():
    rng = np.random.default_rng(11)
    B = 20000
    st = _status(rng, B)
    it = rng.integers(0, 5, B).astype(np.int32)
    t = _tracker(st, it, max_iterations=5)
    wst, wit = st.tolist(), it.tolist()
    for _ in range(3):
        t.increment_all_active()
        wst, wit = P.increment_all_active(wst, wit, 5)
    t.device.synchronize()
    assert _host_status(t) == wst and t.iterations_tensor.cpu().numpy().tolist() == wit
    s = t.summary()
    w = P.summary(wst, wit)
    assert (s.converged_count, s.failed_count, s.active_count, s.max_iterations) == \
        (w["converged_count"], w["failed_count"], w["active_count"], w["max_iterations"])
    assert s.average_iterations == w["average_iterations"]
    # repeated indices: a mark succeeds once, increments count every time
    act = [i for i, v in enumerate(wst) if v == P.ACTIVE][:3]
    if act:
        assert t._mark([act[0], act[0], act[-1]], A.SPX_CONVERGED) == len(set([act[0], act[-1]]))


@pytest.mark.parametrize("B", [0, 1, 4095, 4096, 4097, 100003])
@pytest.mark.parametrize("which", [0, 1, 2])
def test_select_indices_match_oracle(B, which):
    rng = np.random.default_rng(B + 31 * which)
    st = _status(rng, B) if B else np.zeros(0, np.int32)
    if B == 0:
        t = PP.ConvergenceTracker(0)
        assert t._select(which).numel() == 0
        return
    t = _tracker(st)
    got = t._select(which).cpu().numpy().tolist()
    assert got == P.select(st.tolist(), which)


def test_compact_expand_match_oracle():
    rng = np.random.default_rng(5)
    B, item = 9000, 13
    data = rng.standard_normal(B * item)
    idx = np.sort(rng.choice(B, 3001, replace=False))
    comp = PP.compact_active(data, item, idx).cpu().numpy()
    want = P.compact_active(data.tolist(), item, idx.tolist())
    np.testing.assert_array_equal(comp, want)
    exp = PP.expand_active(comp, item, idx, B, -7.5).cpu().numpy()
    np.testing.assert_array_equal(exp, P.expand_active(want, item, idx.tolist(), B, -7.5))
    # the reference's own vectors (convergence.rs:580-607)
    np.testing.assert_array_equal(
        PP.compact_active([1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0], 2, [0, 2]).cpu().numpy(),
        [1.0, 2.0, 5.0, 6.0])
    np.testing.assert_array_equal(
        PP.expand_active([1.0, 2.0, 5.0, 6.0], 2, [0, 2], 4, 0.0).cpu().numpy(),
        [1.0, 2.0, 0.0, 0.0, 5.0, 6.0, 0.0, 0.0])
    with pytest.raises(E.InvalidArgument):
        PP.expand_active([1.0, 2.0], 1, [0, 0], 4, 0.0)


@pytest.mark.parametrize("bins", [1, 4, 7, 5000])
def test_histogram_matches_oracle(bins):
    rng = np.random.default_rng(bins)
    B, n, node = 30001, 5, 3
    x = rng.standard_normal(B * n)
    x[[3 * n + node, 10 * n + node, 99 * n + node]] = [np.nan, np.inf, -np.inf]
    h = PP.Histogram.from_sweep(x, n, B, node, bins)
    counts, lo, hi = P.histogram_from_sweep(x.tolist(), n, B, node, bins)
    assert (h.min, h.max) == (lo, hi) and h.bins == counts and h.total_count == B
    h = PP.Histogram.new(x[node::n], bins, -1.0, 2.0)
    assert h.bins == P.histogram(x[node::n].tolist(), bins, -1.0, 2.0)[0]


def test_histogram_reference_known_answers():
    h = PP.Histogram.new([1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 5.0], 4, 1.0, 5.0)
    assert len(h.bins) == 4 and h.total_count == 9 and abs(h.bin_width - 1.0) < 1e-10
    h = PP.Histogram.from_sweep([1.0, 0.0, 2.0, 0.0, 3.0, 0.0, 4.0, 0.0, 5.0, 0.0], 2, 5, 0, 4)
    assert h.total_count == 5 and len(h.bins) == 4
    h = PP.Histogram.new([1.0, 2.0, 3.0, 3.0, 3.0, 3.0, 4.0, 5.0], 5, 1.0, 5.0)
    assert abs(h.bin_center(h.mode_bin()) - 3.0) < h.bin_width
    h = PP.Histogram.from_sweep([2.0, 2.0, 2.0], 1, 3, 0, 2)
    assert (h.min, h.max, h.bins) == (1.5, 2.5, [0, 3])
    with pytest.raises(E.InvalidArgument):
        PP.Histogram.new([1.0], 0, 0.0, 1.0)


def test_yield_matches_oracle():
    rng = np.random.default_rng(9)
    B, n = 50000, 8
    x = rng.standard_normal(B * n)
    x[17] = np.nan
    specs = [(0, -1.0, 1.0), (3, -0.5, 2.0), (7, 0.0, 0.0), (7, -3.0, 3.0)]
    ya = PP.YieldAnalysis([PP.YieldSpec(*s) for s in specs])
    all_pass, each = P.yield_counts(x.tolist(), n, B, specs)
    assert ya._counts(x, n, B) == (all_pass, each)
    assert ya.compute_yield(x, n, B) == all_pass / B
    assert ya.individual_yields(x, n, B) == [c / B for c in each]
    # the reference's vectors (statistics.rs:862-904)
    sol = [1.0, 0.0, 2.0, 0.0, 3.0, 0.0, 4.0, 0.0, 5.0, 0.0]
    spec = PP.YieldSpec(0, 1.5, 3.5)
    assert abs(spec.compute_yield(sol, 2, 5) - 0.4) < 1e-10
    assert spec.count_pass_fail(sol, 2, 5) == (2, 3)
    yb = PP.YieldAnalysis([PP.YieldSpec(0, 1.5, 2.5), PP.YieldSpec(1, 15.0, 25.0)])
    assert abs(yb.compute_yield([1.0, 10.0, 2.0, 20.0, 3.0, 30.0], 2, 3) - 1 / 3) < 1e-10


def test_streaming_statistics_match_oracle():
    rng = np.random.default_rng(13)
    n = 11
    s1, s2 = PP.StreamingStatistics(n), PP.StreamingStatistics(n)
    acc = [[0, 0.0, 0.0, math.inf, -math.inf] for _ in range(n)]
    for k, B in enumerate((1, 700, 5003)):
        chunk = rng.standard_normal(B * n) * (k + 1) + k
        (s1 if k < 2 else s2).process_chunk(chunk, B)
        acc = P.stats_accumulate(acc, chunk.tolist(), n, B)
    s1.merge(s2)
    assert s1.total_samples() == 1 + 700 + 5003
    for j, st in enumerate(s1.finalize()):
        c, s, sq, mn, mx = acc[j]
        w = E.SweepStatistics.from_moments(c, s, sq, mn, mx)
        assert st.count == w.count and st.min == w.min and st.max == w.max
        assert abs(st.sum - w.sum) <= 1e-12 * abs(c * max(abs(mn), abs(mx)))
        assert abs(st.variance - w.variance) <= 1e-10 * w.variance
    # statistics.rs:988-996: two 2-point chunks of 2 nodes
    s = PP.StreamingStatistics(2)
    s.process_chunk([1.0, 10.0, 2.0, 20.0], 2)
    s.process_chunk([3.0, 30.0, 4.0, 40.0], 2)
    f = s.finalize()
    assert s.total_samples() == 4 and abs(f[0].mean - 2.5) < 1e-10 and abs(f[1].mean - 25.0) < 1e-10


@pytest.mark.parametrize("n", [1, 2, 10, 33, 64, 100])
def test_pack_matches_oracle(n):
    rng = np.random.default_rng(n)
    B = 5
    m = rng.standard_normal(B * n * n)
    lay = PP.BatchLayout.new(n, B)
    got = PP.pack_matrices_f32(m, n, B, lay).cpu().numpy()
    np.testing.assert_array_equal(got, P.pack_matrices_f32(m.tolist(), n, B))
    lay = PP.BatchLayout.unpadded(n, B)
    got = PP.pack_matrices_f32(m, n, B, lay).cpu().numpy()
    np.testing.assert_array_equal(got, P.pack_matrices_f32(m.tolist(), n, B, n))
    r = rng.standard_normal(B * n)
    pr = PP.pack_rhs_f32(r, n, B).cpu().numpy()
    np.testing.assert_array_equal(pr, r.astype(np.float32))
    np.testing.assert_array_equal(PP.unpack_solutions_f64(pr, n, B).cpu().numpy(),
                                  r.astype(np.float32).astype(np.float64))


def test_batch_layout_reference_known_answers():
    lay = PP.BatchLayout.new(10, 100)
    assert (lay.padded_row_stride(), lay.padded_matrix_size(), lay.has_padding(),
            lay.row_padding()) == (32, 320, True, 22)
    assert lay.matrix_offset(0, 1, 2) == 34 and PP.BatchLayout.new(10, 3).matrix_offset(1, 0, 0) == 320
    ps = lay.padding_stats()
    assert (ps.useful_elements, ps.total_elements, ps.wasted_bytes) == (100, 320, 220 * 100 * 4)
    p = PP.pack_matrices_f32([1.0, 0.0, 0.0, 1.0], 2, 1, PP.BatchLayout.new(2, 1)).cpu().numpy()
    assert p.size == 64 and (p[0], p[1], p[32], p[33]) == (1.0, 0.0, 0.0, 1.0)
    p = PP.pack_matrices_f32([1.0, 3.0, 2.0, 4.0, 5.0, 7.0, 6.0, 8.0], 2, 2,
                             PP.BatchLayout.unpadded(2, 2)).cpu().numpy()
    assert p.tolist() == [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0]


def _diag_factory(n):
    # the reference's TestStamper (pipeline.rs:415-450): A = diag(v), b_i = v (i + 1)
    def f(params):
        v = params[0] if len(params) else 1.0
        return np.eye(n) * v, v * (np.arange(n) + 1.0)
    return f


def test_pipelined_sweep_reference_cases():
    pts = [[1.0 + i * 0.1] for i in range(10)]
    r = PP.PipelinedSweep(PP.PipelineConfig.with_chunk_size(100)).execute(_diag_factory(3), pts, 3)
    assert r.total_count == 10 and r.converged_count() == 10 and r.chunks_processed == 1
    np.testing.assert_allclose(r.solution(0), [1.0, 2.0, 3.0], atol=1e-10)
    pts = [[1.0 + i * 0.01] for i in range(100)]
    r = PP.PipelinedSweep(PP.PipelineConfig.with_chunk_size(25)).execute(_diag_factory(3), pts, 3)
    assert r.chunks_processed == 4 and r.converged_count() == 100
    assert all(abs(s[0] - 1.0) < 1e-10 for s in r.solutions)
    r = PP.PipelinedSweep(PP.PipelineConfig.with_chunk_size(30)).execute(
        _diag_factory(2), [[2.0]] * 100, 2)
    assert r.total_count == 100 and r.converged_count() == 100 and r.chunks_processed == 4
    c = PP.PipelineConfig.auto(50, 10000)
    assert 100 <= c.chunk_size <= 2500 and c.num_buffers == 2


@pytest.mark.parametrize("n,B,chunk,nb", [(64, 20000, 1500, 3), (5, 70001, 8000, 2),
                                          (100, 900, 250, 2)])
def test_pipelined_solve_equals_one_shot(n, B, chunk, nb):
    # the same kernels on the same systems: bit-identical to spx_solve_batch_f64,
    # singular systems (zero rows planted) included
    rng = np.random.default_rng(n + B)
    m = rng.standard_normal((B, n, n)) + 0.5 * n * np.eye(n)
    m[[3, B // 2]] = 0.0
    b = rng.standard_normal((B, n))
    cm = np.ascontiguousarray(np.transpose(m, (0, 2, 1))).reshape(-1)
    want = E.HipBatchedSolver().solve_batch(cm, b.reshape(-1), n, B)
    r = PP.PipelinedSweep(PP.PipelineConfig(chunk, nb)).execute_arrays(cm, b.reshape(-1), n, B)
    assert r.chunks_processed == -(-B // chunk)
    assert r.singular_indices == want.singular_indices == [3, B // 2]
    np.testing.assert_array_equal(np.concatenate(r.solutions), want.solutions)
