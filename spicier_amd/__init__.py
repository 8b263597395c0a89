"""MI355X batched-sweep engine: hand-written HIP kernels behind a C ABI
(include/spicier_hip.h, libspicier_hip.so), with a Python mirror of
spicier-batched-sweep's backend interface."""
from . import _abi  # noqa: F401
from .circuit import Circuit, Param  # noqa: F401
from .engine import (  # noqa: F401
    Backend, BackendInit, BackendSelector, BackendType, BatchedLuSolver, BatchedSolveResult,
    BatchedSweepError, BatchTooLarge, ConvergenceCriteria, ConvergenceStatus, DcopSweepResult,
    Device, GminSteppingParams, GpuBatchConfig, HipBatchedSolver, InvalidArgument,
    InvalidDimension, Limiter, MemoryAlloc, NoBackendAvailable, Precision, SourceSteppingParams,
    SweepStatistics, Transfer, nr_config, sample_params, sweep_dcop,
)
