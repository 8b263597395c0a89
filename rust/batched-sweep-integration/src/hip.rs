//! HIP backend implementation for batched LU solving: the file a maintainer
//! adds to crates/spicier-batched-sweep/src/ (beside cuda.rs, whose shape it
//! follows, cuda.rs:12-87), behind `#[cfg(feature = "hip")] mod hip;`.
//!
//! It needs three one-line additions in that crate:
//!   * `BackendType::Hip` in the closed enum (solver.rs:79-92) and its
//!     `Display` arm (`"HIP"`);
//!   * the `BackendType::Hip` arm of `try_create_backend` (solver.rs:258-311)
//!     calling `HipBatchedSolver::new`, and `Hip` first in
//!     `create_solver`'s auto order (solver.rs:212-256);
//!   * the `From<HipError>` impl below, next to the CUDA one (error.rs:48-71).

use crate::error::{BatchedSweepError, Result};
use crate::solver::{BackendType, BatchedLuSolver, BatchedSolveResult, GpuBatchConfig};
use spicier_backend_hip::{GpuBatchedSweepConfig, HipBatchedLuSolver, HipContext, HipError};

/// HIP-accelerated batched LU solver (AMD Instinct MI355X).
pub struct HipBatchedSolver {
    solver: HipBatchedLuSolver,
    config: GpuBatchConfig,
}

impl HipBatchedSolver {
    /// Creates the solver on device 0.  Cheap after the first call: the
    /// device context is process-wide (parallel.rs:194 creates a solver per
    /// point).
    ///
    /// # Errors
    /// `BackendInit` if no HIP device or runtime is available.
    pub fn new(config: GpuBatchConfig) -> Result<Self> {
        let ctx = HipContext::shared(0)
            .map_err(|e| BatchedSweepError::BackendInit(format!("HIP context: {e}")))?;
        let hip_config = GpuBatchedSweepConfig {
            min_batch_size: config.min_batch_size,
            min_matrix_size: config.min_matrix_size,
            max_batch_per_launch: config.max_batch_per_launch,
        };
        Ok(Self { solver: HipBatchedLuSolver::with_config(ctx, hip_config), config })
    }

    /// Try to create a HIP solver, returning None if unavailable.
    pub fn try_new(config: GpuBatchConfig) -> Option<Self> {
        Self::new(config).ok()
    }
}

impl BatchedLuSolver for HipBatchedSolver {
    fn solve_batch(
        &self,
        matrices: &[f64],
        rhs: &[f64],
        n: usize,
        batch_size: usize,
    ) -> Result<BatchedSolveResult> {
        let r = self.solver.solve_batch(matrices, rhs, n, batch_size)?;
        Ok(BatchedSolveResult {
            solutions: r.solutions,
            singular_indices: r.singular_indices,
            n: r.n,
            batch_size: r.batch_size,
        })
    }

    fn should_use_gpu(&self, matrix_size: usize, batch_size: usize) -> bool {
        self.solver.should_use_gpu(matrix_size, batch_size)
    }

    fn backend_type(&self) -> BackendType {
        BackendType::Hip
    }

    fn config(&self) -> &GpuBatchConfig {
        &self.config
    }
}

/// For error.rs, beside `From<spicier_backend_cuda::CudaError>`: one arm per
/// library status code.
impl From<HipError> for BatchedSweepError {
    fn from(e: HipError) -> Self {
        match e {
            HipError::NoDevice => BatchedSweepError::NoBackendAvailable,
            HipError::DeviceInit(m) => BatchedSweepError::BackendInit(m),
            HipError::MemoryAlloc(m) => BatchedSweepError::MemoryAlloc(m),
            HipError::Transfer(m) => BatchedSweepError::Transfer(m),
            HipError::InvalidDimension(m) => BatchedSweepError::InvalidDimension(m),
            HipError::BatchTooLarge { size, max } => BatchedSweepError::BatchTooLarge { size, max },
            other => BatchedSweepError::Backend(other.to_string()),
        }
    }
}
