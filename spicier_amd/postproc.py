"""Host mirror of the reference's sweep bookkeeping and result reductions,
over the C ABI (include/spicier_hip.h).  Every operation runs in
libspicier_hip.so on device buffers; torch tensors are only the device
memory the calls read and write.

Names, argument meaning and error behaviour follow:
  * ``ConvergenceTracker`` / ``ConvergenceSummary`` / ``compact_active`` /
    ``expand_active`` (crates/spicier-batched-sweep/src/convergence.rs:76-471),
  * ``Histogram`` / ``YieldSpec`` / ``YieldAnalysis`` / ``StreamingStatistics``
    (crates/spicier-batched-sweep/src/statistics.rs:215-587),
  * ``BatchLayout`` / ``pack_matrices_f32`` / ``pack_rhs_f32`` /
    ``unpack_solutions_f64`` / ``PaddingStats``
    (crates/spicier-batched-sweep/src/batch_layout.rs:30-242),
  * ``PipelineConfig`` / ``PipelinedSweep`` / ``PipelinedSweepResult``
    (crates/spicier-batched-sweep/src/pipeline.rs:7-393).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi as A
from .engine import (BackendType, ConvergenceStatus, Device, HipBatchedSolver,
                     InvalidArgument, InvalidDimension, SweepStatistics, _check)


def _torch():
    import torch
    return torch


def _dev(device_id):
    return _torch().device(f"cuda:{device_id}")


def _as_device(x, dtype, device_id=0):
    """Device tensor view of x (copied to the device if it is host data)."""
    torch = _torch()
    if isinstance(x, torch.Tensor) and x.device.type == "cuda":
        return x.to(dtype).contiguous()
    return torch.as_tensor(np.asarray(x), dtype=dtype).to(_dev(device_id)).contiguous()


# ------------------------------------------------------------ convergence
@dataclass
class ConvergenceSummary:
    """convergence.rs:390-400."""
    total_points: int
    converged_count: int
    failed_count: int
    active_count: int
    average_iterations: float
    max_iterations: int

    def __str__(self):  # convergence.rs:403-415
        pct = 100.0 * self.converged_count / self.total_points if self.total_points else 0.0
        return (f"Converged: {self.converged_count}/{self.total_points} ({pct:.1f}%), "
                f"Failed: {self.failed_count}, Avg iterations: {self.average_iterations:.1f}")


class ConvergenceTracker:
    """ConvergenceTracker (convergence.rs:76-387) with the status and
    iteration arrays in device memory (int32 tensors ``status`` and
    ``iterations``), so a device-resident Newton loop can update them in
    place.  Single-instance queries read one element back."""

    def __init__(self, batch_size: int, max_iterations: int = 50, device_id: int = 0):
        torch = _torch()
        self.device = Device.get(device_id)
        self.max_iterations = int(max_iterations)
        self.status_tensor = torch.full((batch_size,), A.SPX_ACTIVE, dtype=torch.int32,
                                        device=_dev(device_id))
        self.iterations_tensor = torch.zeros(batch_size, dtype=torch.int32, device=_dev(device_id))
        self._device_id = device_id

    @classmethod
    def with_max_iterations(cls, batch_size: int, max_iterations: int, device_id: int = 0):
        return cls(batch_size, max_iterations, device_id)

    def _sync_torch(self):
        # the library stream and torch's stream are different queues
        _torch().cuda.synchronize(self._device_id)

    # -- counts
    def batch_size(self) -> int:
        return int(self.status_tensor.numel())

    def summary(self) -> ConvergenceSummary:
        self._sync_torch()
        s = A.SpxTrackerSummary()
        _check(self.device.lib.spx_tracker_summary_device(
            self.device.handle, A.ptr(self.status_tensor), A.ptr(self.iterations_tensor),
            self.batch_size(), C.byref(s)))
        return ConvergenceSummary(int(s.total_points), int(s.converged_count),
                                  int(s.failed_count), int(s.active_count),
                                  float(s.average_iterations), int(s.max_iterations))

    def active_count(self) -> int:
        return self.summary().active_count

    def converged_count(self) -> int:
        return self.summary().converged_count

    def failed_count(self) -> int:
        return self.summary().failed_count

    def all_finished(self) -> bool:
        return self.active_count() == 0

    def status(self, index: int) -> ConvergenceStatus:
        self.device.synchronize()
        return ConvergenceStatus(int(self.status_tensor[index].item()))

    def iterations(self, index: int) -> int:
        self.device.synchronize()
        return int(self.iterations_tensor[index].item())

    # -- transitions
    def _mark(self, indices, to) -> int:
        self._sync_torch()
        idx = np.ascontiguousarray(np.atleast_1d(indices), dtype=np.int64)
        marked = C.c_int64(0)
        _check(self.device.lib.spx_tracker_mark_device(
            self.device.handle, A.ptr(self.status_tensor), self.batch_size(), A.ptr(idx),
            idx.size, to, C.byref(marked)))
        return int(marked.value)

    def mark_converged(self, index: int) -> bool:
        return self._mark(index, A.SPX_CONVERGED) == 1

    def mark_failed(self, index: int) -> bool:
        return self._mark(index, A.SPX_FAILED) == 1

    def mark_singular(self, index: int) -> bool:
        return self._mark(index, A.SPX_SINGULAR) == 1

    def increment_iteration(self, index: int) -> int:
        self._sync_torch()
        idx = np.array([index], dtype=np.int64)
        _check(self.device.lib.spx_tracker_increment_device(
            self.device.handle, A.ptr(self.status_tensor), A.ptr(self.iterations_tensor),
            self.batch_size(), self.max_iterations, A.ptr(idx), 1))
        return self.iterations(index)

    def increment_all_active(self):
        self._sync_torch()
        _check(self.device.lib.spx_tracker_increment_device(
            self.device.handle, A.ptr(self.status_tensor), A.ptr(self.iterations_tensor),
            self.batch_size(), self.max_iterations, None, 0))

    # -- selections
    def _select(self, which):
        torch = _torch()
        self._sync_torch()
        B = self.batch_size()
        out = torch.empty(max(B, 1), dtype=torch.int64, device=self.status_tensor.device)
        cnt = torch.zeros(1, dtype=torch.int64, device=self.status_tensor.device)
        _check(self.device.lib.spx_select_indices_device(
            self.device.handle, A.ptr(self.status_tensor), B, which, A.ptr(out), A.ptr(cnt)))
        self.device.synchronize()
        return out[:int(cnt.item())]

    def active_indices_device(self):
        """Ascending active indices as a device int64 tensor (feeds compact_active)."""
        return self._select(A.SPX_SELECT_ACTIVE)

    def active_indices(self) -> list:
        return [int(i) for i in self._select(A.SPX_SELECT_ACTIVE).cpu()]

    def converged_indices(self) -> list:
        return [int(i) for i in self._select(A.SPX_SELECT_CONVERGED).cpu()]

    def failed_indices(self) -> list:
        return [int(i) for i in self._select(A.SPX_SELECT_FAILED).cpu()]

    def active_mask(self) -> list:
        self.device.synchronize()
        return [bool(s == A.SPX_ACTIVE) for s in self.status_tensor.cpu().tolist()]

    def active_mask_u32(self) -> list:
        return [1 if m else 0 for m in self.active_mask()]

    # -- convergence checks
    def check_convergence(self, solution_changes, n: int, abstol: float, reltol: float,
                          solutions) -> int:
        torch = _torch()
        B = self.batch_size()
        d = _as_device(solution_changes, torch.float64, self._device_id)
        x = _as_device(solutions, torch.float64, self._device_id)
        if d.numel() != B * n or x.numel() != B * n:
            raise InvalidDimension(f"expected {B * n} values")
        newly = torch.zeros(1, dtype=torch.int64, device=d.device)
        self._sync_torch()
        _check(self.device.lib.spx_tracker_check_convergence_device(
            self.device.handle, A.ptr(self.status_tensor), B, n, A.ptr(d), A.ptr(x), abstol,
            reltol, A.ptr(newly)))
        self.device.synchronize()
        return int(newly.item())

    def check_residual_convergence(self, residuals, n: int, tolerance: float) -> int:
        torch = _torch()
        B = self.batch_size()
        r = _as_device(residuals, torch.float64, self._device_id)
        if r.numel() != B * n:
            raise InvalidDimension(f"expected {B * n} values")
        newly = torch.zeros(1, dtype=torch.int64, device=r.device)
        self._sync_torch()
        _check(self.device.lib.spx_tracker_check_residual_device(
            self.device.handle, A.ptr(self.status_tensor), B, n, A.ptr(r), tolerance,
            A.ptr(newly)))
        self.device.synchronize()
        return int(newly.item())


def compact_active(data, item_size: int, active_indices, device_id: int = 0):
    """compact_active (convergence.rs:432-441) on the device: returns a device
    tensor of the listed rows."""
    torch = _torch()
    dev = Device.get(device_id)
    d = _as_device(data, torch.float64, device_id)
    idx = _as_device(active_indices, torch.int64, device_id)
    out = torch.empty(idx.numel() * item_size, dtype=torch.float64, device=d.device)
    torch.cuda.synchronize(device_id)
    _check(dev.lib.spx_compact_rows_device(dev.handle, A.ptr(d), item_size, A.ptr(idx),
                                           idx.numel(), A.ptr(out)))
    dev.synchronize()
    return out


def expand_active(compacted, item_size: int, active_indices, batch_size: int, default: float,
                  device_id: int = 0):
    """expand_active (convergence.rs:453-471) on the device."""
    torch = _torch()
    dev = Device.get(device_id)
    c = _as_device(compacted, torch.float64, device_id)
    idx = _as_device(active_indices, torch.int64, device_id)
    if len(set(idx.cpu().tolist())) != idx.numel():
        raise InvalidArgument("expand_active needs distinct indices")
    out = torch.empty(batch_size * item_size, dtype=torch.float64, device=c.device)
    torch.cuda.synchronize(device_id)
    _check(dev.lib.spx_expand_rows_device(dev.handle, A.ptr(c), item_size, A.ptr(idx),
                                          idx.numel(), batch_size, float(default), A.ptr(out)))
    dev.synchronize()
    return out


# ------------------------------------------------------------ histograms
class Histogram:
    """Histogram (statistics.rs:221-327); binning on the device."""

    def __init__(self, counts, vmin: float, vmax: float):
        self.bins = [int(c) for c in counts]
        self.min = float(vmin)
        self.max = float(vmax)
        self.bin_width = (self.max - self.min) / len(self.bins)
        self.total_count = int(sum(self.bins))

    @classmethod
    def _run(cls, solutions, n, batch_size, node, num_bins, vmin, vmax, from_sweep, device_id):
        torch = _torch()
        if num_bins <= 0:
            raise InvalidArgument("num_bins must be > 0")
        dev = Device.get(device_id)
        x = _as_device(solutions, torch.float64, device_id)
        if x.numel() != n * batch_size:
            raise InvalidDimension(f"expected {n * batch_size} values")
        counts = torch.zeros(num_bins, dtype=torch.int64, device=x.device)
        rng = torch.zeros(2, dtype=torch.float64, device=x.device)
        torch.cuda.synchronize(device_id)
        _check(dev.lib.spx_histogram_device(dev.handle, A.ptr(x), batch_size, n, node, num_bins,
                                            float(vmin), float(vmax), int(from_sweep),
                                            A.ptr(counts), A.ptr(rng)))
        dev.synchronize()
        lo, hi = rng.cpu().tolist()
        return cls(counts.cpu().tolist(), lo, hi)

    @classmethod
    def new(cls, values, num_bins: int, vmin: float, vmax: float, device_id: int = 0):
        """Histogram::new (statistics.rs:244-270) over a value sequence."""
        if not vmax > vmin:
            raise InvalidArgument("max must be > min")
        v = np.asarray(list(values) if not hasattr(values, "__len__") else values,
                       dtype=np.float64)
        return cls._run(v, 1, v.size, 0, num_bins, vmin, vmax, False, device_id)

    @classmethod
    def from_sweep(cls, solutions, n: int, batch_size: int, node_idx: int, num_bins: int,
                   device_id: int = 0):
        """Histogram::from_sweep (statistics.rs:273-293)."""
        return cls._run(solutions, n, batch_size, node_idx, num_bins, 0.0, 0.0, True, device_id)

    def bin_index(self, value: float) -> int:  # :296-299
        t = np.floor((value - self.min) / self.bin_width)
        idx = 0 if t != t else int(max(min(t, 2 ** 62), -2 ** 62))
        return min(max(idx, 0), len(self.bins) - 1)

    def bin_center(self, bin_idx: int) -> float:  # :302-304
        return self.min + (bin_idx + 0.5) * self.bin_width

    def bin_percentage(self, bin_idx: int) -> float:  # :307-313
        return 0.0 if self.total_count == 0 else self.bins[bin_idx] / self.total_count

    def mode_bin(self) -> int:  # :316-323 (max_by_key: the last of equal maxima)
        best, idx = -1, 0
        for i, c in enumerate(self.bins):
            if c >= best:
                best, idx = c, i
        return idx


@dataclass
class YieldSpec:
    """YieldSpec (statistics.rs:331-376)."""
    node_idx: int
    lower_limit: float
    upper_limit: float

    def passes(self, value: float) -> bool:
        return self.lower_limit <= value <= self.upper_limit

    def _counts(self, solutions, n, batch_size, device_id=0):
        return YieldAnalysis([self])._counts(solutions, n, batch_size, device_id)

    def compute_yield(self, solutions, n: int, batch_size: int) -> float:
        _, each = self._counts(solutions, n, batch_size)
        return each[0] / batch_size

    def count_pass_fail(self, solutions, n: int, batch_size: int):
        _, each = self._counts(solutions, n, batch_size)
        return each[0], batch_size - each[0]


class YieldAnalysis:
    """YieldAnalysis (statistics.rs:379-425): all specs must pass."""

    def __init__(self, specs):
        self.specs = list(specs)

    def all_pass(self, solution) -> bool:
        return all(s.passes(solution[s.node_idx]) for s in self.specs)

    def _counts(self, solutions, n, batch_size, device_id=0):
        torch = _torch()
        dev = Device.get(device_id)
        x = _as_device(solutions, torch.float64, device_id)
        if x.numel() != n * batch_size:
            raise InvalidDimension(f"expected {n * batch_size} values")
        ns = len(self.specs)
        arr = (A.SpxYieldSpec * max(ns, 1))(*[A.SpxYieldSpec(s.node_idx, s.lower_limit,
                                                              s.upper_limit) for s in self.specs])
        counts = torch.zeros(ns + 1, dtype=torch.int64, device=x.device)
        torch.cuda.synchronize(device_id)
        _check(dev.lib.spx_yield_device(dev.handle, A.ptr(x), batch_size, n, arr, ns,
                                        A.ptr(counts)))
        dev.synchronize()
        c = counts.cpu().tolist()
        return c[0], c[1:]

    def compute_yield(self, solutions, n: int, batch_size: int) -> float:
        return self._counts(solutions, n, batch_size)[0] / batch_size

    def individual_yields(self, solutions, n: int, batch_size: int) -> list:
        return [c / batch_size for c in self._counts(solutions, n, batch_size)[1]]


class StreamingStatistics:
    """StreamingStatistics (statistics.rs:470-587): per-node accumulators in
    device memory, one reduction + merge per chunk."""

    def __init__(self, num_nodes: int, device_id: int = 0):
        self.device = Device.get(device_id)
        self._device_id = device_id
        self._n = int(num_nodes)
        self.reset()

    def reset(self):
        torch = _torch()
        acc = torch.zeros(self._n, 5, dtype=torch.float64, device=_dev(self._device_id))
        acc[:, 3] = float("inf")
        acc[:, 4] = float("-inf")
        self.acc = acc
        self._total = 0

    def num_nodes(self) -> int:
        return self._n

    def total_samples(self) -> int:
        return self._total

    def process_chunk(self, solutions, num_sweeps: int, mask=None):
        torch = _torch()
        x = _as_device(solutions, torch.float64, self._device_id)
        if x.numel() != num_sweeps * self._n:
            raise InvalidDimension(f"solutions.len() ({x.numel()}) != num_sweeps ({num_sweeps}) "
                                   f"* num_nodes ({self._n})")
        m = None if mask is None else _as_device(mask, torch.uint8, self._device_id)
        torch.cuda.synchronize(self._device_id)
        _check(self.device.lib.spx_stats_accumulate_device(
            self.device.handle, A.ptr(x), A.ptr(m), num_sweeps, self._n, A.ptr(self.acc)))
        self.device.synchronize()
        self._total += num_sweeps

    def process_chunk_f32(self, solutions, num_sweeps: int):
        self.process_chunk(np.asarray(solutions, dtype=np.float32).astype(np.float64), num_sweeps)

    def merge(self, other: "StreamingStatistics"):
        if other._n != self._n:
            raise InvalidDimension("Cannot merge statistics with different num_nodes")
        _torch().cuda.synchronize(self._device_id)
        _check(self.device.lib.spx_stats_merge_device(self.device.handle, A.ptr(self.acc),
                                                      A.ptr(other.acc), self._n))
        self.device.synchronize()
        self._total += other._total

    def node_statistics(self, node_idx: int) -> SweepStatistics:
        self.device.synchronize()
        c, s, sq, mn, mx = self.acc[node_idx].cpu().tolist()
        return SweepStatistics.from_moments(c, s, sq, mn, mx)

    def finalize(self) -> list:
        return [self.node_statistics(j) for j in range(self._n)]


# ------------------------------------------------------------ batch layout
WARP_SIZE = 32          # batch_layout.rs:21
BUFFER_ALIGNMENT = 256  # batch_layout.rs:24


@dataclass
class PaddingStats:
    """batch_layout.rs:203-213."""
    useful_elements: int
    total_elements: int
    overhead_ratio: float
    wasted_bytes: int


class BatchLayout:
    """BatchLayout (batch_layout.rs:30-133): row-major FP32 systems with rows
    padded to a multiple of 32."""

    def __init__(self, n: int, batch_size: int, padded: bool = True):
        self.n = n
        self.batch_size = batch_size
        self._stride = self.align_to_warp(n) if padded else n
        self._msize = self._stride * n

    @classmethod
    def new(cls, n, batch_size):
        return cls(n, batch_size, True)

    @classmethod
    def unpadded(cls, n, batch_size):
        return cls(n, batch_size, False)

    @staticmethod
    def align_to_warp(size: int) -> int:
        return -(-size // WARP_SIZE) * WARP_SIZE

    def padded_row_stride(self) -> int:
        return self._stride

    def padded_matrix_size(self) -> int:
        return self._msize

    def total_matrix_elements(self) -> int:
        return self._msize * self.batch_size

    def total_rhs_elements(self) -> int:
        return self.n * self.batch_size

    def matrix_offset(self, batch_idx: int, row: int, col: int) -> int:
        return batch_idx * self._msize + row * self._stride + col

    def rhs_offset(self, batch_idx: int, i: int) -> int:
        return batch_idx * self.n + i

    def has_padding(self) -> bool:
        return self._stride > self.n

    def row_padding(self) -> int:
        return self._stride - self.n

    def padding_stats(self) -> PaddingStats:
        useful = self.n * self.n
        total = self._msize
        pad = total - useful
        return PaddingStats(useful, total, pad / useful if useful else 0.0,
                            pad * self.batch_size * 4)


def pack_matrices_f32(matrices, n: int, batch_size: int, layout: BatchLayout, device_id: int = 0):
    """pack_matrices_f32 (batch_layout.rs:163-187) on the device; returns a
    device float32 tensor of layout.total_matrix_elements()."""
    torch = _torch()
    dev = Device.get(device_id)
    m = _as_device(matrices, torch.float64, device_id)
    if m.numel() != batch_size * n * n or layout.n != n or layout.batch_size != batch_size:
        raise InvalidDimension("matrices do not match the layout")
    out = torch.empty(max(layout.total_matrix_elements(), 1), dtype=torch.float32,
                      device=m.device)
    torch.cuda.synchronize(device_id)
    _check(dev.lib.spx_pack_matrices_f32_device(dev.handle, A.ptr(m), n, batch_size,
                                                layout.padded_row_stride(), A.ptr(out)))
    dev.synchronize()
    return out[:layout.total_matrix_elements()]


def pack_rhs_f32(rhs, n: int, batch_size: int, device_id: int = 0):
    """pack_rhs_f32 (batch_layout.rs:190-193)."""
    torch = _torch()
    dev = Device.get(device_id)
    r = _as_device(rhs, torch.float64, device_id)
    if r.numel() != n * batch_size:
        raise InvalidDimension("rhs length")
    out = torch.empty(max(r.numel(), 1), dtype=torch.float32, device=r.device)
    torch.cuda.synchronize(device_id)
    _check(dev.lib.spx_pack_rhs_f32_device(dev.handle, A.ptr(r), r.numel(), A.ptr(out)))
    dev.synchronize()
    return out[:r.numel()]


def unpack_solutions_f64(solutions_f32, n: int, batch_size: int, device_id: int = 0):
    """unpack_solutions_f64 (batch_layout.rs:196-199)."""
    torch = _torch()
    dev = Device.get(device_id)
    s = _as_device(solutions_f32, torch.float32, device_id)
    if s.numel() != n * batch_size:
        raise InvalidDimension("solutions length")
    out = torch.empty(max(s.numel(), 1), dtype=torch.float64, device=s.device)
    torch.cuda.synchronize(device_id)
    _check(dev.lib.spx_unpack_solutions_f64_device(dev.handle, A.ptr(s), s.numel(), A.ptr(out)))
    dev.synchronize()
    return out[:s.numel()]


# ------------------------------------------------------------ pipelined sweep
@dataclass
class PipelineConfig:
    """PipelineConfig (pipeline.rs:7-87)."""
    chunk_size: int = 1000
    num_buffers: int = 2

    @classmethod
    def with_chunk_size(cls, chunk_size: int) -> "PipelineConfig":
        return cls(chunk_size=chunk_size)

    @classmethod
    def auto(cls, system_size: int, total_points: int) -> "PipelineConfig":
        c = A.SpxPipelineConfig()
        A.load_library().spx_pipeline_config_auto(system_size, total_points, C.byref(c))
        return cls(int(c.chunk_size), int(c.num_buffers))


@dataclass
class PipelinedSweepResult:
    """pipeline.rs:365-393."""
    solutions: list
    singular_indices: list
    total_count: int
    backend_used: BackendType
    chunks_processed: int

    def solution(self, index: int):
        return self.solutions[index] if 0 <= index < len(self.solutions) else None

    def is_singular(self, index: int) -> bool:
        return index in self.singular_indices

    def converged_count(self) -> int:
        return self.total_count - len(self.singular_indices)


class PipelinedSweep:
    """PipelinedSweep (pipeline.rs:130-361).  The stamper builds each point's
    dense system on the host as in the reference; the chunks then go through
    spx_solve_batch_f64_pipelined, whose pinned staging slots and per-slot
    streams overlap host copies, transfers and solves."""

    def __init__(self, config: PipelineConfig, solver: HipBatchedSolver | None = None):
        self.config = config
        self.solver = solver or HipBatchedSolver()

    def execute(self, factory, points, system_size: int) -> PipelinedSweepResult:
        """factory(parameters) -> (matrix n x n, rhs n): the stamper of
        SweepStamperFactory::create_stamper + stamp_linear (sweep.rs:272-326)."""
        n, B = system_size, len(points)
        M = np.zeros((B, n, n))
        b = np.zeros((B, n))
        for i, p in enumerate(points):
            params = getattr(p, "parameters", p)
            m, r = factory(params)
            M[i], b[i] = m, r
        cm = np.ascontiguousarray(np.transpose(M, (0, 2, 1))).reshape(-1)  # column-major
        return self.execute_arrays(cm, b.reshape(-1), n, B)

    def execute_arrays(self, matrices, rhs, n: int, batch_size: int) -> PipelinedSweepResult:
        """The solve of pre-assembled column-major systems."""
        m = np.ascontiguousarray(matrices, dtype=np.float64).reshape(-1)
        r = np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1)
        if m.size != batch_size * n * n or r.size != batch_size * n:
            raise InvalidDimension("matrix / rhs length")
        chunk = min(self.config.chunk_size, batch_size) if batch_size else 1
        # small sweeps run as one batch (pipeline.rs:152-155)
        if batch_size <= chunk or batch_size < 100:
            chunk = max(batch_size, 1)
        cfg = A.SpxPipelineConfig(chunk, self.config.num_buffers)
        x = np.zeros(batch_size * n)
        sing = np.zeros(max(batch_size, 1), dtype=np.uint8)
        ns, nc = C.c_int64(0), C.c_int64(0)
        dev = self.solver.device
        _check(dev.lib.spx_solve_batch_f64_pipelined(dev.handle, A.ptr(m), A.ptr(r), n,
                                                     batch_size, C.byref(cfg), A.ptr(x),
                                                     A.ptr(sing), C.byref(ns), C.byref(nc)))
        sols = [x[i * n:(i + 1) * n] for i in range(batch_size)]
        idx = [int(i) for i in np.nonzero(sing[:batch_size])[0]]
        return PipelinedSweepResult(sols, idx, batch_size, BackendType.Hip,
                                    max(int(nc.value), 1 if batch_size else 0))
