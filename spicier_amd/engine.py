"""Host-side mirror of spicier-batched-sweep's backend interface, over the C ABI.

Names, argument meaning and error behaviour follow the reference:
  * ``BatchedLuSolver`` / ``BatchedSolveResult`` / ``GpuBatchConfig`` /
    ``BackendType`` / ``BackendSelector``
    (crates/spicier-batched-sweep/src/solver.rs:1-312),
  * the ``BatchedSweepError`` family (crates/spicier-batched-sweep/src/error.rs:7-46),
  * ``ConvergenceCriteria`` + Gmin / source stepping parameters
    (crates/spicier-solver/src/newton.rs:11-34,174-231),
  * ``ConvergenceStatus`` (crates/spicier-batched-sweep/src/convergence.rs:41-50),
  * ``SweepStatistics`` (crates/spicier-batched-sweep/src/statistics.rs:35-74).
Every computation runs in libspicier_hip.so; there is no CPU fallback: a
missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
import enum
import threading
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A

# ---------------------------------------------------------------- errors


class BatchedSweepError(RuntimeError):
    """Base of the reference's error enum (error.rs:7-46)."""


class NoBackendAvailable(BatchedSweepError):
    pass


class BackendInit(BatchedSweepError):
    pass


class MemoryAlloc(BatchedSweepError):
    pass


class Transfer(BatchedSweepError):
    pass


class InvalidDimension(BatchedSweepError):
    pass


class BatchTooLarge(BatchedSweepError):
    pass


class Backend(BatchedSweepError):
    pass


class InvalidArgument(BatchedSweepError):
    pass


_ERRORS = {
    A.SPX_ERR_INVALID_DIMENSION: InvalidDimension, A.SPX_ERR_BATCH_TOO_LARGE: BatchTooLarge,
    A.SPX_ERR_MEMORY_ALLOC: MemoryAlloc, A.SPX_ERR_TRANSFER: Transfer,
    A.SPX_ERR_BACKEND_INIT: BackendInit, A.SPX_ERR_BACKEND: Backend,
    A.SPX_ERR_NO_BACKEND: NoBackendAvailable, A.SPX_ERR_INVALID_ARGUMENT: InvalidArgument,
}


def _check(rc: int):
    if rc != A.SPX_OK:
        msg = A.load_library().spx_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, BatchedSweepError)(msg)


# ---------------------------------------------------------------- config / types
MAX_BATCH_SIZE = 65535     # solver.rs:6
MIN_BATCH_SIZE = 16        # solver.rs:9
MIN_MATRIX_SIZE = 32       # solver.rs:12


@dataclass
class GpuBatchConfig:
    """solver.rs:14-41."""
    min_batch_size: int = MIN_BATCH_SIZE
    min_matrix_size: int = MIN_MATRIX_SIZE
    max_batch_per_launch: int = MAX_BATCH_SIZE

    def should_use_gpu(self, matrix_size: int, batch_size: int) -> bool:
        return matrix_size >= self.min_matrix_size and batch_size >= self.min_batch_size


class BackendType(enum.Enum):
    """solver.rs:79-92 plus the HIP backend this engine adds."""
    Cuda = "CUDA"
    Metal = "Metal"
    Mps = "MPS"
    Accelerate = "Accelerate"
    Faer = "Faer"
    Cpu = "CPU"
    Hip = "HIP"

    def __str__(self):
        return self.value


class ConvergenceStatus(enum.IntEnum):
    Active = A.SPX_ACTIVE
    Converged = A.SPX_CONVERGED
    Failed = A.SPX_FAILED
    Singular = A.SPX_SINGULAR


@dataclass
class BatchedSolveResult:
    """solver.rs:43-75."""
    solutions: np.ndarray
    singular_indices: list
    n: int
    batch_size: int

    def solution(self, index: int):
        if index >= self.batch_size:
            return None
        return self.solutions[index * self.n:(index + 1) * self.n]

    def is_singular(self, index: int) -> bool:
        return index in self.singular_indices

    def num_solved(self) -> int:
        return self.batch_size - len(self.singular_indices)


# ---------------------------------------------------------------- device handles
class Device:
    """One spx handle (device + HIP stream).  Handles are cached per device so
    creating a solver per sweep point (parallel.rs:194) stays cheap."""

    _cache: dict = {}
    _lock = threading.Lock()

    def __init__(self, device_id: int = 0):
        lib = A.load_library()
        h = C.c_void_p()
        _check(lib.spx_create(device_id, C.byref(h)))
        self.lib = lib
        self.handle = h
        self.device_id = device_id

    @classmethod
    def get(cls, device_id: int = 0) -> "Device":
        with cls._lock:
            d = cls._cache.get(device_id)
            if d is None:
                d = cls(device_id)
                cls._cache[device_id] = d
            return d

    @staticmethod
    def count() -> int:
        lib = A.load_library()
        c = C.c_int32(0)
        rc = lib.spx_device_count(C.byref(c))
        return c.value if rc == A.SPX_OK else 0

    def synchronize(self):
        _check(self.lib.spx_synchronize(self.handle))

    @property
    def stream(self) -> int:
        return self.lib.spx_stream(self.handle)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                self.lib.spx_destroy(self.handle)
        except Exception:
            pass


# ---------------------------------------------------------------- the trait
class BatchedLuSolver:
    """The reference's `BatchedLuSolver` trait (solver.rs:108-135)."""

    def solve_batch(self, matrices, rhs, n: int, batch_size: int) -> BatchedSolveResult:
        raise NotImplementedError

    def should_use_gpu(self, matrix_size: int, batch_size: int) -> bool:
        raise NotImplementedError

    def backend_type(self) -> BackendType:
        raise NotImplementedError

    def config(self) -> GpuBatchConfig:
        raise NotImplementedError


class HipBatchedSolver(BatchedLuSolver):
    """The HIP drop-in for CudaBatchedSolver (spicier-batched-sweep/src/cuda.rs:12-87):
    matrices column-major per system, singular systems -> zeros + index."""

    def __init__(self, config: GpuBatchConfig | None = None, device_id: int = 0):
        self._config = config or GpuBatchConfig()
        self.device = Device.get(device_id)

    def solve_batch(self, matrices, rhs, n: int, batch_size: int) -> BatchedSolveResult:
        m = np.ascontiguousarray(matrices, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1)
        if m.size != batch_size * n * n:
            raise InvalidDimension(f"Expected {batch_size * n * n} matrix elements, got {m.size}")
        if b.size != batch_size * n:
            raise InvalidDimension(f"Expected {batch_size * n} RHS elements, got {b.size}")
        x = np.zeros(batch_size * n)
        sing = np.zeros(max(batch_size, 1), dtype=np.uint8)
        ns = C.c_int64(0)
        _check(self.device.lib.spx_solve_batch_f64(self.device.handle, A.ptr(m), A.ptr(b), n,
                                                   batch_size, A.ptr(x), A.ptr(sing),
                                                   C.byref(ns)))
        idx = [int(i) for i in np.nonzero(sing[:batch_size])[0]]
        return BatchedSolveResult(x, idx, n, batch_size)

    def solve_batch_device(self, d_matrices, d_rhs, n: int, batch_size: int, d_x, d_sing=None):
        """Device-pointer variant on torch tensors / raw pointers (stream-ordered)."""
        _check(self.device.lib.spx_solve_batch_f64_device(
            self.device.handle, A.ptr(d_matrices), A.ptr(d_rhs), n, batch_size, A.ptr(d_x),
            A.ptr(d_sing)))

    def factor_batch(self, matrices, n: int, batch_size: int):
        """The getrf half (cublasDgetrfBatched, batched_lu.rs:404-428) with the
        CPU path's pivot rule: (factors [B][n][n], factors[i] the n x n matrix
        of LAPACK's getrf format, 0-based pivots [B][n], singular indices)."""
        m = np.ascontiguousarray(matrices, dtype=np.float64).reshape(-1)
        if m.size != batch_size * n * n:
            raise InvalidDimension(f"Expected {batch_size * n * n} matrix elements, got {m.size}")
        f = np.zeros(batch_size * n * n)
        piv = np.zeros(max(batch_size * n, 1), dtype=np.int32)
        sing = np.zeros(max(batch_size, 1), dtype=np.uint8)
        ns = C.c_int64(0)
        _check(self.device.lib.spx_lu_factor_batch_f64(self.device.handle, A.ptr(m), n,
                                                       batch_size, A.ptr(f), A.ptr(piv),
                                                       A.ptr(sing), C.byref(ns)))
        idx = [int(i) for i in np.nonzero(sing[:batch_size])[0]]
        return (f.reshape(batch_size, n, n).transpose(0, 2, 1),
                piv[:batch_size * n].reshape(batch_size, n), idx)

    def should_use_gpu(self, matrix_size: int, batch_size: int) -> bool:
        return self._config.should_use_gpu(matrix_size, batch_size)

    def backend_type(self) -> BackendType:
        return BackendType.Hip

    def config(self) -> GpuBatchConfig:
        return self._config


class HipSparseBatchedSolver(BatchedLuSolver):
    """Shared-pattern sparse LU on the GPU: the drop-in for
    FaerSparseCachedBatchedSolver (dense input, pattern |a_ij| > 1e-15 of the
    first matrix) and FaerTripletBatchedSolver (triplets with one pattern),
    crates/spicier-batched-sweep/src/faer_sparse_solver.rs:29-362.  The
    symbolic analysis is done once per pattern and cached on the device
    handle; systems whose own nonzeros fall outside the cached pattern are
    reported singular, as the reference's try_new_with_symbolic rejects them."""

    def __init__(self, config: GpuBatchConfig | None = None, device_id: int = 0):
        self._config = config or GpuBatchConfig()
        self.device = Device.get(device_id)
        self._pattern = None  # (n, rows, cols) from the first dense batch

    def reset_cache(self):
        self._pattern = None

    def has_cached_symbolic(self) -> bool:
        return self._pattern is not None

    def solve_batch_triplets(self, rows, cols, values, rhs, n: int) -> BatchedSolveResult:
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        v = np.ascontiguousarray(values, dtype=np.float64)
        nnz = rows.size
        B = v.size // nnz if nnz else 0
        b = np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1)
        if cols.size != nnz or v.size != B * nnz or b.size != B * n:
            raise InvalidDimension(f"expected {B}x{nnz} values and {B}x{n} rhs")
        x = np.zeros(B * n)
        sing = np.zeros(max(B, 1), dtype=np.uint8)
        ns = C.c_int64(0)
        _check(self.device.lib.spx_sparse_solve_batch_f64(
            self.device.handle, n, nnz, A.ptr(rows), A.ptr(cols), A.ptr(v), A.ptr(b), B,
            A.ptr(x), A.ptr(sing), C.byref(ns)))
        return BatchedSolveResult(x, [int(i) for i in np.nonzero(sing[:B])[0]], n, B)

    def solve_batch(self, matrices, rhs, n: int, batch_size: int) -> BatchedSolveResult:
        m = np.ascontiguousarray(matrices, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1)
        if m.size != batch_size * n * n:
            raise InvalidDimension(f"Expected {batch_size * n * n} matrix elements, got {m.size}")
        if b.size != batch_size * n:
            raise InvalidDimension(f"Expected {batch_size * n} RHS elements, got {b.size}")
        if batch_size == 0:
            return BatchedSolveResult(np.zeros(0), [], n, 0)
        mats = m.reshape(batch_size, n, n)           # [i][col][row]
        nz = np.abs(mats) > 1e-15                    # dense_to_sparse_triplets threshold
        if self._pattern is None:
            cols, rows = np.nonzero(nz[0])
            self._pattern = (n, rows.astype(np.int32), cols.astype(np.int32))
        pn, rows, cols = self._pattern
        if pn != n:
            raise InvalidDimension(f"Matrix size {n} doesn't match cached size {pn}")
        inside = np.zeros((n, n), dtype=bool)
        inside[cols, rows] = True
        fits = ~np.any(nz & ~inside, axis=(1, 2))
        vals = mats[:, cols, rows]
        r = self.solve_batch_triplets(rows, cols, vals, b, n)
        x = r.solutions.reshape(batch_size, n)
        x[~fits] = 0.0
        sing = sorted(set(r.singular_indices) | {int(i) for i in np.nonzero(~fits)[0]})
        return BatchedSolveResult(x.reshape(-1), sing, n, batch_size)

    def should_use_gpu(self, matrix_size: int, batch_size: int) -> bool:
        return self._config.should_use_gpu(matrix_size, batch_size)

    def backend_type(self) -> BackendType:
        return BackendType.Hip

    def config(self) -> GpuBatchConfig:
        return self._config


@dataclass
class GmresConfig:
    """GmresConfig (crates/spicier-solver/src/gmres/mod.rs:34-48) plus the
    preconditioner choice (JacobiPreconditioner, preconditioner.rs:43-100)."""
    max_iter: int = 500
    tol: float = 1e-8
    restart: int = 30
    jacobi: bool = True
    # Precision.F32_IR: FP32 Krylov arithmetic inside FP64 iterative refinement
    precision: int = A.SPX_PREC_F64

    def to_c(self) -> A.SpxGmresConfig:
        return A.SpxGmresConfig(restart=self.restart, max_iter=self.max_iter, tol=self.tol,
                                jacobi=int(self.jacobi), precision=int(self.precision))


@dataclass
class GmresBatchResult:
    """Per system: RealGmresResult (gmres/real.rs:10-21)."""
    solutions: np.ndarray
    iterations: np.ndarray
    residual: np.ndarray
    converged: np.ndarray
    n: int
    batch_size: int

    def solution(self, i: int) -> np.ndarray:
        return self.solutions.reshape(self.batch_size, self.n)[i]


class HipBatchedGmres:
    """Batched restarted GMRES-Jacobi on the GPU, every system sharing one
    triplet pattern: solve_gmres_real_preconditioned per system
    (crates/spicier-solver/src/gmres/real.rs:208-400), one launch for the
    whole batch (the Metal sketch's GpuBatchedGmres::solve,
    crates/spicier-backend-metal/src/batched_gmres.rs:396-657)."""

    def __init__(self, config: GmresConfig | None = None, device_id: int = 0):
        self.config = config or GmresConfig()
        self.device = Device.get(device_id)

    def solve_batch_triplets(self, rows, cols, values, rhs, n: int, x0=None) -> GmresBatchResult:
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        v = np.ascontiguousarray(values, dtype=np.float64)
        nnz = rows.size
        B = v.size // nnz if nnz else 0
        b = np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1)
        if cols.size != nnz or v.size != B * nnz or b.size != B * n:
            raise InvalidDimension(f"expected {B}x{nnz} values and {B}x{n} rhs")
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1)
            if x0.size != B * n:
                raise InvalidDimension(f"expected {B}x{n} initial guesses")
        x = np.zeros(B * n)
        it = np.zeros(max(B, 1), dtype=np.int32)
        res = np.zeros(max(B, 1))
        conv = np.zeros(max(B, 1), dtype=np.uint8)
        cfg = self.config.to_c()
        _check(self.device.lib.spx_gmres_solve_batch_f64(
            self.device.handle, n, nnz, A.ptr(rows), A.ptr(cols), A.ptr(v), A.ptr(b), A.ptr(x0),
            B, C.byref(cfg), A.ptr(x), A.ptr(it), A.ptr(res), A.ptr(conv)))
        return GmresBatchResult(x, it[:B], res[:B], conv[:B].astype(bool), n, B)


@dataclass
class BackendSelector:
    """solver.rs:139-312, with HIP as the only backend this engine ships.  The
    reference falls through to its CPU solver when no GPU is present; this
    engine has no CPU path and raises NoBackendAvailable instead."""
    preferred: BackendType | None = None
    config: GpuBatchConfig = field(default_factory=GpuBatchConfig)

    @classmethod
    def auto(cls) -> "BackendSelector":
        return cls()

    @classmethod
    def prefer_hip(cls) -> "BackendSelector":
        return cls(BackendType.Hip)

    def create_solver(self) -> BatchedLuSolver:
        if self.preferred not in (None, BackendType.Hip):
            raise NoBackendAvailable(f"backend {self.preferred} is not part of this engine")
        if Device.count() == 0:
            raise NoBackendAvailable("No GPU backend available")
        return HipBatchedSolver(self.config)


# ---------------------------------------------------------------- Newton config
@dataclass
class ConvergenceCriteria:
    """newton.rs:11-34."""
    v_abstol: float = 1e-6
    v_reltol: float = 1e-3
    i_abstol: float = 1e-12
    max_iterations: int = 50
    gmin: float = 1e-12


@dataclass
class GminSteppingParams:
    """newton.rs:174-194."""
    initial_gmin: float = 1e-3
    final_gmin: float = 1e-12
    reduction_factor: float = 10.0
    max_attempts_per_level: int = 5


@dataclass
class SourceSteppingParams:
    """newton.rs:211-231."""
    initial_factor: float = 0.1
    factor_step: float = 0.1
    max_attempts_per_level: int = 5
    min_factor_step: float = 0.01


class Precision(enum.IntEnum):
    F64 = A.SPX_PREC_F64
    F32_IR = A.SPX_PREC_F32_IR


class Limiter(enum.IntEnum):
    Reference = A.SPX_LIMIT_REFERENCE
    NoLimit = A.SPX_LIMIT_NONE
    Pnjlim = A.SPX_LIMIT_PNJLIM


def nr_config(criteria: ConvergenceCriteria | None = None,
              gmin: GminSteppingParams | None = None,
              source: SourceSteppingParams | None = None,
              precision: Precision = Precision.F64, limiter: Limiter = Limiter.Reference,
              gmin_stepping: bool = True, source_stepping: bool = True,
              ir_max_steps: int = 8, ir_tol: float = 0.0) -> A.SpxNrConfig:
    c = criteria or ConvergenceCriteria()
    g = gmin or GminSteppingParams()
    s = source or SourceSteppingParams()
    return A.default_nr_config(
        v_abstol=c.v_abstol, v_reltol=c.v_reltol, i_abstol=c.i_abstol,
        max_iterations=c.max_iterations,
        gmin_stepping=int(gmin_stepping), gmin_initial=g.initial_gmin, gmin_final=g.final_gmin,
        gmin_factor=g.reduction_factor, gmin_max_attempts=g.max_attempts_per_level,
        source_stepping=int(source_stepping), src_initial=s.initial_factor,
        src_step=s.factor_step, src_min_step=s.min_factor_step,
        src_max_attempts=s.max_attempts_per_level,
        precision=int(precision), limiter=int(limiter), ir_max_steps=ir_max_steps,
        ir_tol=ir_tol)


# ---------------------------------------------------------------- statistics
@dataclass
class SweepStatistics:
    """statistics.rs:35-74 (from_moments: population variance E[x^2]-mean^2)."""
    count: int
    min: float
    max: float
    mean: float
    variance: float
    std_dev: float
    sum: float

    @classmethod
    def from_moments(cls, count, s, sq, mn, mx):
        count = int(count)
        mean = s / count if count > 0 else 0.0
        var = max((sq / count) - mean * mean, 0.0) if count > 0 else 0.0
        return cls(count, mn, mx, mean, var, var ** 0.5, s)


# ---------------------------------------------------------------- DC-OP sweep
@dataclass
class DcopSweepResult:
    """Nonlinear counterpart of GpuBatchedSweepResult (sweep.rs:17-55)."""
    solutions: np.ndarray      # [B, n]
    status: np.ndarray         # [B] ConvergenceStatus codes
    iterations: np.ndarray     # [B]
    flags: np.ndarray          # [B] SPX_FLAG_*
    residual: np.ndarray       # [B] max |KCL| / tol
    params: np.ndarray         # [B, P]
    node_stats: np.ndarray     # [n, 5] count,sum,sumsq,min,max over converged
    factorizations: np.ndarray = None  # [B] LU factorizations performed
    refinements: np.ndarray = None     # [B] FP64-residual refinement steps
    used_gpu: bool = True
    backend_used: BackendType = BackendType.Hip

    @property
    def total_count(self) -> int:
        return len(self.status)

    @property
    def converged_count(self) -> int:
        return int(np.count_nonzero(self.status == A.SPX_CONVERGED))

    @property
    def singular_indices(self):
        return [int(i) for i in np.nonzero(self.status == A.SPX_SINGULAR)[0]]

    def accepted_mask(self) -> np.ndarray:
        """Converged and passing the FP64 KCL residual test."""
        return (self.status == A.SPX_CONVERGED) & ((self.flags & A.SPX_FLAG_KCL_FAIL) == 0)

    def node_voltages(self, node_index: int) -> np.ndarray:
        return self.solutions[:, node_index]

    def statistics(self, node_index: int) -> SweepStatistics:
        return SweepStatistics.from_moments(*self.node_stats[node_index])


def sweep_dcop(circuit, batch: int, seed: int = 42, generator: int = A.SPX_GEN_HASH,
               sweep_offset: int = 0, config: A.SpxNrConfig | None = None, x0=None,
               explicit_params=None, device_id: int = 0) -> DcopSweepResult:
    """Monte Carlo / corner / linear DC-operating-point sweep on the GPU
    (host buffers in and out)."""
    dev = Device.get(device_id)
    cfg = config or nr_config()
    ckt, keep = circuit.to_c()
    if explicit_params is not None:
        explicit_params = np.ascontiguousarray(explicit_params, dtype=np.float64)
        generator = A.SPX_GEN_EXPLICIT
    spec, keep2 = circuit.sweep_spec(batch, seed, generator, sweep_offset, explicit_params)
    n, P = circuit.size, circuit.num_params
    out = DcopSweepResult(np.zeros((batch, n)), np.zeros(batch, np.int32),
                          np.zeros(batch, np.int32), np.zeros(batch, np.int32),
                          np.zeros(batch), np.zeros((batch, P)), np.zeros((n, 5)),
                          np.zeros(batch, np.int32), np.zeros(batch, np.int32))
    res = A.SpxSweepResult(A.ptr(out.solutions), A.ptr(out.status), A.ptr(out.iterations),
                           A.ptr(out.flags), A.ptr(out.residual), A.ptr(out.params),
                           A.ptr(out.node_stats), A.ptr(out.factorizations),
                           A.ptr(out.refinements))
    if x0 is not None:
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        cfg.warm_start = A.SPX_WARM_GIVEN
    _check(dev.lib.spx_sweep_dcop(dev.handle, C.byref(ckt), C.byref(spec), C.byref(cfg),
                                  A.ptr(x0), C.byref(res)))
    return out


class IntegrationMethod(enum.IntEnum):
    """IntegrationMethod (crates/spicier-solver/src/transient/types.rs:50-63)."""
    BackwardEuler = A.SPX_TRAN_BE
    Trapezoidal = A.SPX_TRAN_TRAP
    TrBdf2 = A.SPX_TRAN_TRBDF2


@dataclass
class AdaptiveTransientParams:
    """AdaptiveTransientParams (transient/types.rs:76-117), with the defaults
    of its Default impl, plus the per-instance capacity of time points."""
    tstop: float = 1e-3
    h_init: float = 1e-9
    h_min: float = 1e-15
    h_max: float = 1e-6
    reltol: float = 1e-3
    abstol: float = 1e-6
    max_points: int = 4096

    @classmethod
    def for_tstop(cls, tstop: float, max_points: int = 4096) -> "AdaptiveTransientParams":
        return cls(tstop=tstop, h_max=tstop / 100.0, max_points=max_points)

    def to_c(self) -> A.SpxTranAdaptive:
        return A.SpxTranAdaptive(self.tstop, self.h_init, self.h_min, self.h_max, self.reltol,
                                 self.abstol, int(self.max_points))


@dataclass
class TransientSweepResult:
    """Transient of every instance (TransientResult / AdaptiveTransientResult,
    crates/spicier-solver/src/transient/result.rs, restricted to the probed
    unknowns): waveforms[i, k, p] = x[probes[p]] at time point k.  Fixed
    steps: t = k * tstep.  Adaptive: point_times[i, k] for k < num_points[i],
    with total_steps, rejected_steps, min_step and max_step per instance."""
    solutions: np.ndarray
    status: np.ndarray
    iterations: np.ndarray
    flags: np.ndarray
    residual: np.ndarray
    params: np.ndarray
    waveforms: np.ndarray
    fail_step: np.ndarray
    tstep: float
    point_times: np.ndarray | None = None
    num_points: np.ndarray | None = None
    total_steps: np.ndarray | None = None
    rejected_steps: np.ndarray | None = None
    min_step: np.ndarray | None = None
    max_step: np.ndarray | None = None

    @property
    def times(self) -> np.ndarray:
        """Fixed steps: the shared grid [points]; adaptive: [batch][max_points]."""
        if self.point_times is not None:
            return self.point_times
        return np.arange(self.waveforms.shape[1]) * self.tstep

    def accepted_mask(self) -> np.ndarray:
        return (self.status == A.SPX_CONVERGED) & ((self.flags & A.SPX_FLAG_KCL_FAIL) == 0)


def _tran_config(tstep, steps, pr, method, adaptive):
    ad = adaptive.to_c() if adaptive is not None else None
    tr = A.SpxTranConfig(tstep=tstep, steps=steps, num_probes=pr.size, probes=A.ptr(pr),
                         method=int(method), adaptive=C.pointer(ad) if ad is not None else None)
    return tr, ad


def transient_sweep(circuit, batch: int, tstep: float, steps: int, probes=None, seed: int = 42,
                    generator: int = A.SPX_GEN_HASH, sweep_offset: int = 0,
                    config: A.SpxNrConfig | None = None, explicit_params=None,
                    method: IntegrationMethod = IntegrationMethod.BackwardEuler,
                    adaptive: AdaptiveTransientParams | None = None,
                    device_id: int = 0) -> TransientSweepResult:
    """Monte Carlo transient on the GPU (host buffers): `steps` fixed steps
    of `tstep` with `method` (solve_transient, transient/solver.rs:41-249),
    or, with `adaptive`, trapezoidal steps under LTE control
    (solve_transient_adaptive, solver.rs:461-676; tstep and steps unused)."""
    dev = Device.get(device_id)
    cfg = config or nr_config()
    ckt, keep = circuit.to_c()
    if explicit_params is not None:
        explicit_params = np.ascontiguousarray(explicit_params, dtype=np.float64)
        generator = A.SPX_GEN_EXPLICIT
    spec, keep2 = circuit.sweep_spec(batch, seed, generator, sweep_offset, explicit_params)
    n, P = circuit.size, circuit.num_params
    pr = np.ascontiguousarray(np.arange(n) if probes is None else probes, dtype=np.int32)
    tr, ad = _tran_config(tstep, steps, pr, method, adaptive)
    pts = adaptive.max_points if adaptive is not None else steps + 1
    out = TransientSweepResult(np.zeros((batch, n)), np.zeros(batch, np.int32),
                               np.zeros(batch, np.int32), np.zeros(batch, np.int32),
                               np.zeros(batch), np.zeros((batch, P)),
                               np.zeros((batch, pts, pr.size)), np.zeros(batch, np.int32),
                               tstep)
    if adaptive is not None:
        out.point_times = np.zeros((batch, pts))
        out.num_points = np.zeros(batch, np.int32)
        out.total_steps = np.zeros(batch, np.int32)
        out.rejected_steps = np.zeros(batch, np.int32)
        out.min_step = np.zeros(batch)
        out.max_step = np.zeros(batch)
    res = A.SpxTranResult(A.ptr(out.solutions), A.ptr(out.status), A.ptr(out.iterations),
                          A.ptr(out.flags), A.ptr(out.residual), A.ptr(out.params),
                          A.ptr(out.waveforms), A.ptr(out.fail_step), A.ptr(out.point_times),
                          A.ptr(out.num_points), A.ptr(out.total_steps),
                          A.ptr(out.rejected_steps), A.ptr(out.min_step), A.ptr(out.max_step))
    _check(dev.lib.spx_transient_sweep(dev.handle, C.byref(ckt), C.byref(spec), C.byref(cfg),
                                       C.byref(tr), C.byref(res)))
    return out


@dataclass
class TranClass:
    """One size class of a mixed transient batch: a circuit, its instance
    count and sampling, its transient settings, and (filled by
    transient_mixed_device) torch device buffers of the results."""
    circuit: object
    batch: int
    tstep: float
    steps: int
    probes: object = None
    seed: int = 42
    sweep_offset: int = 0
    buffers: dict = field(default_factory=dict)
    method: IntegrationMethod = IntegrationMethod.BackwardEuler
    adaptive: AdaptiveTransientParams | None = None

    def allocate(self, device_id: int = 0):
        import torch
        dev = torch.device(f"cuda:{device_id}")
        n, P, B = self.circuit.size, self.circuit.num_params, self.batch
        pr = np.arange(n) if self.probes is None else np.asarray(self.probes)
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        pts = self.adaptive.max_points if self.adaptive is not None else self.steps + 1
        self.buffers = dict(x=torch.empty(B, n, **f64), status=torch.empty(B, **i32),
                            iterations=torch.empty(B, **i32), flags=torch.empty(B, **i32),
                            residual=torch.empty(B, **f64), params=torch.empty(B, P, **f64),
                            waveforms=torch.empty(B, pts, pr.size, **f64),
                            fail_step=torch.empty(B, **i32))
        if self.adaptive is not None:
            self.buffers.update(times=torch.empty(B, pts, **f64),
                                num_points=torch.empty(B, **i32),
                                total_steps=torch.empty(B, **i32),
                                rejected_steps=torch.empty(B, **i32),
                                min_step=torch.empty(B, **f64), max_step=torch.empty(B, **f64))
        return self

    def host_result(self) -> "TransientSweepResult":
        b = {k: v.cpu().numpy() for k, v in self.buffers.items()}
        return TransientSweepResult(b["x"], b["status"], b["iterations"], b["flags"],
                                    b["residual"], b["params"], b["waveforms"], b["fail_step"],
                                    self.tstep, b.get("times"), b.get("num_points"),
                                    b.get("total_steps"), b.get("rejected_steps"),
                                    b.get("min_step"), b.get("max_step"))


class KernelKind(enum.IntEnum):
    Auto = A.SPX_KERNEL_AUTO
    Dense = A.SPX_KERNEL_DENSE
    Sparse64 = A.SPX_KERNEL_SPARSE64
    Sparse256 = A.SPX_KERNEL_SPARSE256
    Lanes = A.SPX_KERNEL_LANES


def transient_mixed_device(classes, config: A.SpxNrConfig | None = None, graph: bool = True,
                           autotune: bool = False, kernels=None, device_id: int = 0):
    """One launch over a mixed-size transient batch (include/spicier_hip.h,
    spx_transient_mixed_device), enqueued on the library stream.  Classes
    without buffers get them allocated.  Returns the kernel kind each class
    ran with."""
    dev = Device.get(device_id)
    cfg = config or nr_config()
    keep = []
    arr = (A.SpxTranClass * max(1, len(classes)))()
    for i, c in enumerate(classes):
        if not c.buffers:
            c.allocate(device_id)
        ckt, k1 = c.circuit.to_c()
        spec, k2 = c.circuit.sweep_spec(c.batch, c.seed, A.SPX_GEN_HASH, c.sweep_offset)
        pr = np.ascontiguousarray(np.arange(c.circuit.size) if c.probes is None else c.probes,
                                  dtype=np.int32)
        keep += [ckt, k1, k2, pr]
        b = c.buffers
        arr[i].circuit = C.pointer(ckt)
        arr[i].spec = spec
        tr, ad = _tran_config(c.tstep, c.steps, pr, c.method, c.adaptive)
        keep.append(ad)
        arr[i].tran = tr
        arr[i].result = A.SpxTranResult(A.ptr(b["x"]), A.ptr(b["status"]),
                                        A.ptr(b["iterations"]), A.ptr(b["flags"]),
                                        A.ptr(b["residual"]), A.ptr(b["params"]),
                                        A.ptr(b["waveforms"]), A.ptr(b["fail_step"]),
                                        *[A.ptr(b.get(k)) for k in
                                          ("times", "num_points", "total_steps",
                                           "rejected_steps", "min_step", "max_step")])
    kinds = np.array([int(k) for k in (kernels or [KernelKind.Auto] * len(classes))],
                     dtype=np.int32)
    opts = (A.SPX_MIXED_GRAPH if graph else 0) | (A.SPX_MIXED_AUTOTUNE if autotune else 0)
    _check(dev.lib.spx_transient_mixed_device(dev.handle, arr, len(classes), C.byref(cfg), opts,
                                              A.ptr(kinds)))
    return [KernelKind(int(k)) for k in kinds]


def sample_params(circuit, batch: int, seed: int = 42, generator: int = A.SPX_GEN_HASH,
                  sweep_offset: int = 0, device_id: int = 0) -> np.ndarray:
    dev = Device.get(device_id)
    spec, keep = circuit.sweep_spec(batch, seed, generator, sweep_offset)
    out = np.zeros((batch, circuit.num_params))
    _check(dev.lib.spx_sample_params(dev.handle, C.byref(spec), circuit.num_params, A.ptr(out)))
    return out


# ---------------------------------------------------------------- learned warm start
class Surrogate:
    """MLP warm start (include/spicier_hip.h spx_surrogate_*): n_in sampled
    parameters -> H -> H (SiLU) -> n_out node voltages, trained and evaluated
    by the library's BF16-MFMA kernels.  Device buffers are torch tensors."""

    def __init__(self, n_in: int = 32, hidden: int = 256, n_out: int = 64, seed: int = 1,
                 device_id: int = 0):
        self.device = Device.get(device_id)
        self.lib = self.device.lib
        self._bind()
        s = C.c_void_p()
        _check(self.lib.spx_surrogate_create(self.device.handle, n_in, hidden, n_out, seed,
                                             C.byref(s)))
        self.ptr = s
        self.n_in, self.hidden, self.n_out = n_in, hidden, n_out
        self.loss_history = []

    def _bind(self):
        L, vp, i32, i64 = self.lib, C.c_void_p, C.c_int32, C.c_int64
        sigs = {
            "spx_surrogate_create": (C.c_int, [vp, i32, i32, i32, C.c_uint64, C.POINTER(vp)]),
            "spx_surrogate_destroy": (C.c_int, [vp]),
            "spx_surrogate_set_inputs": (C.c_int, [vp, vp, i32]),
            "spx_surrogate_num_weights": (C.c_int64, [vp]),
            "spx_surrogate_get_weights": (C.c_int, [vp, vp, i64]),
            "spx_surrogate_set_weights": (C.c_int, [vp, vp, i64]),
            "spx_surrogate_set_normalization": (C.c_int, [vp, vp, vp, vp, vp]),
            "spx_surrogate_train_device": (C.c_int, [vp, vp, vp, i32, vp, i64, i32, i32,
                                                     C.c_float, C.c_float, C.c_float,
                                                     C.c_uint64, vp]),
            "spx_surrogate_predict_device": (C.c_int, [vp, vp, i64, i32, vp]),
            "spx_surrogate_gradient_device": (C.c_int, [vp, vp, vp, i64, vp,
                                                        C.POINTER(C.c_float)]),
            "spx_set_surrogate": (C.c_int, [vp, vp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args

    @property
    def num_weights(self) -> int:
        return int(self.lib.spx_surrogate_num_weights(self.ptr))

    def set_inputs(self, slots, num_params: int):
        arr = np.ascontiguousarray(slots, dtype=np.int32)
        assert arr.size == self.n_in
        _check(self.lib.spx_surrogate_set_inputs(self.ptr, A.ptr(arr), num_params))

    def get_weights(self) -> np.ndarray:
        w = np.zeros(self.num_weights, dtype=np.float32)
        _check(self.lib.spx_surrogate_get_weights(self.ptr, A.ptr(w), w.size))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(self.lib.spx_surrogate_set_weights(self.ptr, A.ptr(w), w.size))

    def set_normalization(self, in_mean, in_std, out_mean, out_std):
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (in_mean, in_std, out_mean,
                                                               out_std)]
        _check(self.lib.spx_surrogate_set_normalization(self.ptr, *[A.ptr(v) for v in a]))

    def gradient(self, X, T):
        """Loss and gradient for standardised inputs/targets (host arrays)."""
        import torch
        dev = torch.device(f"cuda:{self.device.device_id}")
        x = torch.as_tensor(np.ascontiguousarray(X, dtype=np.float32), device=dev)
        t = torch.as_tensor(np.ascontiguousarray(T, dtype=np.float32), device=dev)
        g = torch.empty(self.num_weights, dtype=torch.float32, device=dev)
        loss = C.c_float(0)
        torch.cuda.synchronize()
        _check(self.lib.spx_surrogate_gradient_device(self.ptr, A.ptr(x), A.ptr(t), x.shape[0],
                                                      A.ptr(g), C.byref(loss)))
        return float(loss.value), g.cpu().numpy()

    def train(self, d_params, d_x, n_x: int, d_mask=None, epochs: int = 30, batch: int = 1024,
              lr: float = 2e-3, tol_rel: float = 1e-3, tol_abs: float = 1e-6,
              shuffle_seed: int = 7):
        """Train on device rows; errors are weighted in units of the Newton
        tolerance tol_rel*|x| + tol_abs (0, 0 = plain standardised MSE)."""
        B = int(d_params.shape[0])
        hist = np.zeros(max(epochs, 1), dtype=np.float32)
        _check(self.lib.spx_surrogate_train_device(self.ptr, A.ptr(d_params), A.ptr(d_x), n_x,
                                                   A.ptr(d_mask), B, epochs, batch, lr, tol_rel,
                                                   tol_abs, shuffle_seed, A.ptr(hist)))
        self.loss_history = hist[:epochs].tolist()
        return self.loss_history

    def predict(self, d_params, n_x: int, d_x0):
        _check(self.lib.spx_surrogate_predict_device(self.ptr, A.ptr(d_params),
                                                     int(d_params.shape[0]), n_x, A.ptr(d_x0)))

    def attach(self):
        """Make this the surrogate used by SPX_WARM_SURROGATE sweeps."""
        _check(self.lib.spx_set_surrogate(self.device.handle, self.ptr))

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                self.lib.spx_surrogate_destroy(self.ptr)
        except Exception:
            pass


def sweep_dcop_device(circuit, batch: int, buffers: dict, seed: int = 42,
                      generator: int = A.SPX_GEN_HASH, sweep_offset: int = 0,
                      config: A.SpxNrConfig | None = None, device_id: int = 0):
    """Sweep into caller-owned device buffers (torch tensors keyed x, status,
    iterations, flags, residual, params, node_stats, factorizations,
    refinements; any but x may be absent).  Enqueued on the library stream."""
    dev = Device.get(device_id)
    cfg = config or nr_config()
    ckt, keep = circuit.to_c()
    spec, keep2 = circuit.sweep_spec(batch, seed, generator, sweep_offset)
    g = buffers.get
    res = A.SpxSweepResult(A.ptr(g("x")), A.ptr(g("status")), A.ptr(g("iterations")),
                           A.ptr(g("flags")), A.ptr(g("residual")), A.ptr(g("params")),
                           A.ptr(g("node_stats")), A.ptr(g("factorizations")),
                           A.ptr(g("refinements")))
    _check(dev.lib.spx_sweep_dcop_device(dev.handle, C.byref(ckt), C.byref(spec), C.byref(cfg),
                                         None, C.byref(res)))


def train_surrogate(circuit, input_slots, n_train: int = 262144, train_offset: int = 1 << 20,
                    seed: int = 42, epochs: int = 30, batch: int = 1024, lr: float = 2e-3,
                    hidden: int = 256, device_id: int = 0):
    """Produce FP64 Newton operating points of `circuit` for parameters drawn at
    sweep indices [train_offset, train_offset + n_train), then train a
    surrogate on them.  Returns the surrogate (not yet attached)."""
    import torch
    dev = torch.device(f"cuda:{device_id}")
    n, P = circuit.size, circuit.num_params
    bufs = dict(x=torch.empty(n_train, n, dtype=torch.float64, device=dev),
                status=torch.empty(n_train, dtype=torch.int32, device=dev),
                flags=torch.empty(n_train, dtype=torch.int32, device=dev),
                params=torch.empty(n_train, P, dtype=torch.float64, device=dev))
    torch.cuda.synchronize()
    sweep_dcop_device(circuit, n_train, bufs, seed=seed, sweep_offset=train_offset,
                      config=nr_config(), device_id=device_id)
    Device.get(device_id).synchronize()
    mask = ((bufs["status"] == A.SPX_CONVERGED) &
            ((bufs["flags"] & A.SPX_FLAG_KCL_FAIL) == 0)).to(torch.uint8)
    s = Surrogate(len(input_slots), hidden, n, seed=seed, device_id=device_id)
    s.set_inputs(input_slots, P)
    s.train(bufs["params"], bufs["x"], n, mask, epochs=epochs, batch=batch, lr=lr)
    return s
