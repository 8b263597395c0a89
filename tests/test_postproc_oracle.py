"""Pins the CPU restatement of the sweep bookkeeping and result reductions
(oracle/postproc_oracle.py) on the reference's own known answers: the unit
tests of convergence.rs:477-667, statistics.rs:843-996, batch_layout.rs:268-394
and pipeline.rs:452-470, plus closed forms where they leave gaps."""
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import postproc_oracle as P  # noqa: E402

A_, C_, F_, S_ = P.ACTIVE, P.CONVERGED, P.FAILED, P.SINGULAR


def test_check_convergence_known_answer():
    # convergence.rs:543-560
    st, newly = P.check_convergence([A_] * 3, [1e-8, 1e-8, 1.0, 1.0, 1e-9, 1e-9],
                                    [1.0, 2.0, 10.0, 20.0, 100.0, 200.0], 2, 1e-6, 1e-6)
    assert newly == 2 and st == [C_, A_, C_]
    # only active points are tested; a NaN change passes (`delta > tol` is false)
    st, newly = P.check_convergence([C_, A_], [1.0, 1.0, float("nan"), 0.0], [0.0] * 4, 2, 0, 0)
    assert newly == 1 and st == [C_, C_]


def test_check_residual_known_answer():
    # convergence.rs:562-578
    st, newly = P.check_residual([A_] * 3, [1e-8, 1e-8, 1.0, 1.0, 1e-9, 1e-9], 2, 1e-6)
    assert newly == 2 and st == [C_, A_, C_]
    # sqrt(3^2 + 4^2) = 5 is not < 5
    assert P.check_residual([A_], [3.0, 4.0], 2, 5.0)[1] == 0


def test_iteration_limit_and_summary():
    # convergence.rs:519-528: 3 increments at max 3 fail all 5
    st, it = [A_] * 5, [0] * 5
    for _ in range(3):
        st, it = P.increment_all_active(st, it, 3)
    assert st == [F_] * 5 and it == [3] * 5
    # convergence.rs:609-640
    st, it = [A_] * 100, [0] * 100
    for i in range(50):
        it[i], st[i] = 3, C_
    for i in range(50, 60):
        it[i], st[i] = 10, C_
    for i in range(60, 65):
        st[i] = F_
    s = P.summary(st, it)
    assert (s["total_points"], s["converged_count"], s["failed_count"], s["active_count"],
            s["max_iterations"]) == (100, 60, 5, 35, 10)
    assert s["average_iterations"] == (50 * 3 + 10 * 10) / 60


def test_selection_compact_expand_known_answers():
    # convergence.rs:530-541, 580-607
    st = [A_, C_, A_, C_, A_]
    assert P.select(st, 0) == [0, 2, 4]
    assert P.select([A_, C_, A_, F_, A_, S_], 2) == [3, 5]
    data = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0]
    assert P.compact_active(data, 2, [0, 2]) == [1.0, 2.0, 5.0, 6.0]
    assert P.expand_active([1.0, 2.0, 5.0, 6.0], 2, [0, 2], 4, 0.0) == \
        [1.0, 2.0, 0.0, 0.0, 5.0, 6.0, 0.0, 0.0]


def test_histogram_known_answers():
    # statistics.rs:843-851: 9 values, 4 bins over [1, 5], width 1
    counts, w = P.histogram([1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 5.0], 4, 1.0, 5.0)
    assert w == 1.0 and sum(counts) == 9
    assert counts == [2, 2, 2, 3]            # 5.0 lands past the last bin and is clamped
    # statistics.rs:853-860
    counts, lo, hi = P.histogram_from_sweep([1.0, 0.0, 2.0, 0.0, 3.0, 0.0, 4.0, 0.0, 5.0, 0.0],
                                            2, 5, 0, 4)
    assert (lo, hi, sum(counts)) == (1.0, 5.0, 5) and counts == [1, 1, 1, 2]
    # statistics.rs:932-941: the mode bin holds the four 3.0s
    counts, w = P.histogram([1.0, 2.0, 3.0, 3.0, 3.0, 3.0, 4.0, 5.0], 5, 1.0, 5.0)
    assert counts.index(max(counts)) == 2 and abs((1.0 + 2.5 * w) - 3.0) < w
    # equal values widen by 0.5 each way; NaN goes to bin 0, infinities clamp
    counts, lo, hi = P.histogram_from_sweep([2.0, 2.0, 2.0], 1, 3, 0, 2)
    assert (lo, hi, counts) == (1.5, 2.5, [0, 3])
    counts, _ = P.histogram([float("nan"), -math.inf, math.inf, 0.5], 4, 0.0, 1.0)
    assert counts == [2, 0, 1, 1]


def test_yield_known_answers():
    # statistics.rs:862-904
    sol = [1.0, 0.0, 2.0, 0.0, 3.0, 0.0, 4.0, 0.0, 5.0, 0.0]
    all_pass, each = P.yield_counts(sol, 2, 5, [(0, 1.5, 3.5)])
    assert all_pass == 2 and each == [2] and abs(all_pass / 5 - 0.4) < 1e-10
    all_pass, each = P.yield_counts([1.0, 10.0, 2.0, 20.0, 3.0, 30.0], 2, 3,
                                    [(0, 1.5, 2.5), (1, 15.0, 25.0)])
    assert all_pass == 1 and each == [1, 1]
    # limits are inclusive; NaN fails
    assert P.yield_counts([1.5, 3.5, float("nan")], 1, 3, [(0, 1.5, 3.5)]) == (2, [2])


def test_streaming_statistics_known_answers():
    # statistics.rs:947-996
    inf = math.inf
    acc = [[0, 0.0, 0.0, inf, -inf] for _ in range(2)]
    acc = P.stats_accumulate(acc, [1.0, 10.0, 2.0, 20.0], 2, 2)
    acc = P.stats_accumulate(acc, [3.0, 30.0, 4.0, 40.0], 2, 2)
    assert acc[0][0] == 4 and acc[0][1] / 4 == 2.5 and acc[1][1] / 4 == 25.0
    assert (acc[0][3], acc[0][4]) == (1.0, 4.0)
    # population variance of 1..5 from the moments is 2 (statistics.rs:760-776)
    a = P.stats_accumulate([[0, 0.0, 0.0, inf, -inf]], [1.0, 2.0, 3.0, 4.0, 5.0], 1, 5)[0]
    assert a[2] / a[0] - (a[1] / a[0]) ** 2 == 2.0


def test_batch_layout_known_answers():
    # batch_layout.rs:268-394
    assert [P.align_to_warp(v) for v in (1, 31, 32, 33, 64, 100)] == [32, 32, 32, 64, 64, 128]
    p = P.pack_matrices_f32([1.0, 0.0, 0.0, 1.0], 2, 1)
    assert p.size == 64 and (p[0], p[1], p[32], p[33]) == (1.0, 0.0, 0.0, 1.0)
    assert list(P.pack_matrices_f32([1.0, 3.0, 2.0, 4.0], 2, 1, 2)) == [1.0, 2.0, 3.0, 4.0]
    assert list(P.pack_matrices_f32([1.0, 3.0, 2.0, 4.0, 5.0, 7.0, 6.0, 8.0], 2, 2, 2)) == \
        [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0]
    # FP64 -> FP32 rounds to nearest even
    assert P.pack_matrices_f32([1.0 + 2 ** -24], 1, 1, 1)[0] == np.float32(1.0)
    assert P.pack_matrices_f32([1.0 + 3 * 2 ** -24], 1, 1, 1)[0] == np.float32(1.0 + 2 ** -22)


def test_pipeline_config_auto():
    # pipeline.rs:459-463, and the size classes of pipeline.rs:68-87
    c, nb = P.pipeline_config_auto(50, 10000)
    assert 100 <= c <= 2500 and nb == 2
    assert P.pipeline_config_auto(10, 100000)[0] == 2000
    assert P.pipeline_config_auto(64, 100000)[0] == 1000
    assert P.pipeline_config_auto(200, 100000)[0] == 500
    assert P.pipeline_config_auto(10, 200)[0] == 100
