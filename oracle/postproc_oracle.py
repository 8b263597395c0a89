"""CPU restatement of the reference's sweep bookkeeping and result
reductions: ConvergenceTracker and compact/expand
(crates/spicier-batched-sweep/src/convergence.rs), Histogram / YieldSpec /
YieldAnalysis / StreamingStatistics (statistics.rs), BatchLayout packing
(batch_layout.rs) and PipelineConfig::auto (pipeline.rs).

TEST INFRASTRUCTURE ONLY: imported by tests/ — never by the engine.  Plain
Python loops in the reference's order over small inputs; each function cites
the lines it restates.
"""
from __future__ import annotations

import math

import numpy as np

ACTIVE, CONVERGED, FAILED, SINGULAR = 0, 1, 2, 3


def check_convergence(status, changes, solutions, n, abstol, reltol):
    """ConvergenceTracker::check_convergence (convergence.rs:274-314).
    Returns (new status, newly converged)."""
    st = list(status)
    newly = 0
    for i in range(len(st)):
        if st[i] != ACTIVE:
            continue
        conv = True
        for j in range(n):
            delta = abs(changes[i * n + j])
            tol = abstol + reltol * abs(solutions[i * n + j])
            if delta > tol:
                conv = False
                break
        if conv:
            st[i] = CONVERGED
            newly += 1
    return st, newly


def check_residual(status, residuals, n, tolerance):
    """check_residual_convergence (convergence.rs:327-357): the squares summed
    in index order."""
    st = list(status)
    newly = 0
    for i in range(len(st)):
        if st[i] != ACTIVE:
            continue
        sq = 0.0
        for j in range(n):
            r = residuals[i * n + j]
            sq += r * r
        if math.sqrt(sq) < tolerance:
            st[i] = CONVERGED
            newly += 1
    return st, newly


def increment_all_active(status, iterations, max_iterations):
    """increment_all_active / increment_iteration (convergence.rs:184-205)."""
    st, it = list(status), list(iterations)
    for i in range(len(st)):
        if st[i] == ACTIVE:
            it[i] += 1
            if it[i] >= max_iterations:
                st[i] = FAILED
    return st, it


def summary(status, iterations):
    """ConvergenceTracker::summary (convergence.rs:360-387)."""
    conv = [i for i, s in enumerate(status) if s == CONVERGED]
    total = sum(iterations[i] for i in conv)
    return dict(total_points=len(status), converged_count=len(conv),
                failed_count=sum(1 for s in status if s in (FAILED, SINGULAR)),
                active_count=sum(1 for s in status if s == ACTIVE),
                iteration_sum=total,
                max_iterations=max((iterations[i] for i in conv), default=0),
                average_iterations=total / len(conv) if conv else 0.0)


def select(status, which):
    """active_indices / converged_indices / failed_indices (convergence.rs:209-238)."""
    if which == 0:
        return [i for i, s in enumerate(status) if s == ACTIVE]
    if which == 1:
        return [i for i, s in enumerate(status) if s == CONVERGED]
    return [i for i, s in enumerate(status) if s in (FAILED, SINGULAR)]


def compact_active(data, item_size, indices):
    """compact_active (convergence.rs:432-441)."""
    out = []
    for idx in indices:
        out.extend(data[idx * item_size:(idx + 1) * item_size])
    return out


def expand_active(compacted, item_size, indices, batch_size, default):
    """expand_active (convergence.rs:453-471)."""
    out = [default] * (batch_size * item_size)
    for k, idx in enumerate(indices):
        out[idx * item_size:(idx + 1) * item_size] = compacted[k * item_size:(k + 1) * item_size]
    return out


def _rust_as_isize(t: float) -> int:
    # `f64 as isize` saturates, NaN -> 0
    if t != t:
        return 0
    if t >= 9.2e18:
        return 2 ** 63 - 1
    if t <= -9.2e18:
        return -(2 ** 63)
    return int(t)


def histogram(values, num_bins, vmin, vmax):
    """Histogram::new (statistics.rs:244-270).  Returns (counts, width)."""
    assert num_bins > 0 and vmax > vmin
    width = (vmax - vmin) / float(num_bins)
    counts = [0] * num_bins
    for v in values:
        idx = _rust_as_isize(float(np.floor((v - vmin) / width)))
        idx = min(max(idx, 0), num_bins - 1)
        counts[idx] += 1
    return counts, width


def histogram_from_sweep(solutions, n, batch_size, node, num_bins):
    """Histogram::from_sweep (statistics.rs:273-293).  Returns (counts, min, max)."""
    lo, hi = math.inf, -math.inf
    for i in range(batch_size):
        v = solutions[i * n + node]
        lo = lo if v != v else min(lo, v)   # f64::min skips NaN
        hi = hi if v != v else max(hi, v)
    if abs(hi - lo) < 2.220446049250313e-16:
        lo -= 0.5
        hi += 0.5
    counts, _ = histogram([solutions[i * n + node] for i in range(batch_size)], num_bins, lo, hi)
    return counts, lo, hi


def yield_counts(solutions, n, batch_size, specs):
    """YieldSpec::passes / YieldAnalysis::compute_yield / individual_yields
    (statistics.rs:352-425): counts passing all specs and each spec.  specs =
    [(node, lower, upper)]."""
    all_pass, each = 0, [0] * len(specs)
    for i in range(batch_size):
        ok = True
        for s, (node, lo, hi) in enumerate(specs):
            v = solutions[i * n + node]
            p = lo <= v <= hi
            each[s] += p
            ok = ok and p
        all_pass += ok
    return all_pass, each


def stats_accumulate(acc, chunk, n, num_sweeps):
    """StreamingStatistics::process_chunk (statistics.rs:523-538) on per-node
    [count, sum, sumsq, min, max] accumulators (StatisticsAccumulator::add,
    :135-141)."""
    acc = [list(a) for a in acc]
    for i in range(num_sweeps):
        for j in range(n):
            v = chunk[i * n + j]
            a = acc[j]
            a[0] += 1
            a[1] += v
            a[2] += v * v
            a[3] = a[3] if v != v else min(a[3], v)
            a[4] = a[4] if v != v else max(a[4], v)
    return acc


def align_to_warp(size):
    """BatchLayout::align_to_warp (batch_layout.rs:70-73)."""
    return -(-size // 32) * 32


def pack_matrices_f32(matrices, n, batch_size, row_stride=None):
    """pack_matrices_f32 (batch_layout.rs:163-187): column-major FP64 to
    row-major FP32 with padded rows."""
    stride = align_to_warp(n) if row_stride is None else row_stride
    out = np.zeros(batch_size * n * stride, dtype=np.float32)
    for b in range(batch_size):
        for row in range(n):
            for col in range(n):
                out[b * n * stride + row * stride + col] = np.float32(matrices[b * n * n + col * n + row])
    return out


def pipeline_config_auto(system_size, total_points):
    """PipelineConfig::auto (pipeline.rs:68-87)."""
    chunk = 2000 if system_size < 50 else (1000 if system_size < 100 else 500)
    chunk = max(min(chunk, total_points // 4), 100)
    return chunk, 2
