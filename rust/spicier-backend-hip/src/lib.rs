//! HIP backend for spicier's batched sweeps on AMD Instinct MI355X (gfx950).
//!
//! This crate wraps `libspicier_hip.so` (the C ABI of `include/spicier_hip.h`)
//! the way `spicier-backend-cuda` wraps cuBLAS: a device context, a batched
//! LU solver with the `GpuBatchedSweepConfig` thresholds, and an error enum
//! that `spicier-batched-sweep` maps onto `BatchedSweepError`
//! (crates/spicier-batched-sweep/src/error.rs:7-46).  The adapter that
//! implements `BatchedLuSolver` for it lives in spicier-batched-sweep
//! (`src/hip.rs`, the counterpart of `src/cuda.rs:12-87`).
//!
//! Beyond the linear drop-in the crate exposes the sweep-level entry point:
//! sampling, Newton with Gmin/source stepping, the FP64 KCL acceptance test
//! and per-node statistics of a whole Monte Carlo / corner sweep on the
//! device, which the trait cannot express because its stamper is a Rust
//! closure.
//!
//! The library never falls back to a CPU path: a missing device is
//! `HipError::NoDevice`, and falling through to Faer/CPU is the selector's
//! decision (solver.rs:258-311).

pub mod ffi;

use std::collections::HashMap;
use std::ffi::CStr;
use std::fmt;
use std::os::raw::c_int;
use std::ptr;
use std::sync::{Arc, Mutex, OnceLock};

/// Errors of the HIP backend, one per library status code.
#[derive(Debug, Clone, PartialEq)]
pub enum HipError {
    /// No HIP device (SPX_ERR_NO_BACKEND).
    NoDevice,
    /// Device or runtime initialisation failed (SPX_ERR_BACKEND_INIT).
    DeviceInit(String),
    /// Device memory allocation failed (SPX_ERR_MEMORY_ALLOC).
    MemoryAlloc(String),
    /// Host <-> device copy failed (SPX_ERR_TRANSFER).
    Transfer(String),
    /// Lengths or sizes do not match (SPX_ERR_INVALID_DIMENSION).
    InvalidDimension(String),
    /// Batch larger than one launch accepts (SPX_ERR_BATCH_TOO_LARGE).
    BatchTooLarge { size: usize, max: usize },
    /// An argument outside its documented domain (SPX_ERR_INVALID_ARGUMENT).
    InvalidArgument(String),
    /// Any other runtime failure (SPX_ERR_BACKEND).
    Backend(String),
}

impl fmt::Display for HipError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        match self {
            HipError::NoDevice => write!(f, "no HIP device available"),
            HipError::DeviceInit(m) => write!(f, "HIP initialisation failed: {m}"),
            HipError::MemoryAlloc(m) => write!(f, "device allocation failed: {m}"),
            HipError::Transfer(m) => write!(f, "transfer failed: {m}"),
            HipError::InvalidDimension(m) => write!(f, "invalid dimensions: {m}"),
            HipError::BatchTooLarge { size, max } => {
                write!(f, "batch size {size} exceeds maximum {max}")
            }
            HipError::InvalidArgument(m) => write!(f, "invalid argument: {m}"),
            HipError::Backend(m) => write!(f, "HIP backend error: {m}"),
        }
    }
}

impl std::error::Error for HipError {}

pub type Result<T> = std::result::Result<T, HipError>;

fn last_error() -> String {
    // spx_last_error is per thread and never null
    unsafe { CStr::from_ptr(ffi::spx_last_error()) }.to_string_lossy().into_owned()
}

/// Maps a library status code to Ok or the matching error.
fn check(rc: c_int) -> Result<()> {
    match rc {
        ffi::SPX_OK => Ok(()),
        ffi::SPX_ERR_NO_BACKEND => Err(HipError::NoDevice),
        ffi::SPX_ERR_BACKEND_INIT => Err(HipError::DeviceInit(last_error())),
        ffi::SPX_ERR_MEMORY_ALLOC => Err(HipError::MemoryAlloc(last_error())),
        ffi::SPX_ERR_TRANSFER => Err(HipError::Transfer(last_error())),
        ffi::SPX_ERR_INVALID_DIMENSION => Err(HipError::InvalidDimension(last_error())),
        ffi::SPX_ERR_BATCH_TOO_LARGE => Err(HipError::BatchTooLarge {
            size: 0,
            max: unsafe { ffi::spx_max_batch_per_launch() } as usize,
        }),
        ffi::SPX_ERR_INVALID_ARGUMENT => Err(HipError::InvalidArgument(last_error())),
        ffi::SPX_ERR_BACKEND => Err(HipError::Backend(last_error())),
        _ => Err(HipError::Backend(format!("status {rc}: {}", last_error()))),
    }
}

/// Number of HIP devices the library sees.
pub fn device_count() -> Result<usize> {
    let mut n = 0i32;
    check(unsafe { ffi::spx_device_count(&mut n) })?;
    Ok(n as usize)
}

/// One device and one HIP stream.  Calls on a context are serialised by the
/// library (one mutex per handle), so a context is shared freely between
/// threads, as the trait's `Send + Sync` requires.
pub struct HipContext {
    handle: *mut ffi::SpxHandle,
    device: i32,
}

// The handle is an opaque pointer owned by this context; the library locks
// it per call.
unsafe impl Send for HipContext {}
unsafe impl Sync for HipContext {}

impl HipContext {
    /// A new context (and stream) on `device`.
    pub fn new(device: i32) -> Result<Self> {
        if device_count()? == 0 {
            return Err(HipError::NoDevice);
        }
        let mut h = ptr::null_mut();
        check(unsafe { ffi::spx_create(device, &mut h) })?;
        Ok(Self { handle: h, device })
    }

    /// The process-wide context of `device`, created on first use.
    /// `solve_batched_sweep_parallel` creates a solver per sweep point
    /// (parallel.rs:194), so solver creation must not create a stream.
    pub fn shared(device: i32) -> Result<Arc<Self>> {
        static CONTEXTS: OnceLock<Mutex<HashMap<i32, Arc<HipContext>>>> = OnceLock::new();
        let map = CONTEXTS.get_or_init(|| Mutex::new(HashMap::new()));
        let mut m = map.lock().unwrap_or_else(|e| e.into_inner());
        if let Some(c) = m.get(&device) {
            return Ok(Arc::clone(c));
        }
        let c = Arc::new(Self::new(device)?);
        m.insert(device, Arc::clone(&c));
        Ok(c)
    }

    pub fn device(&self) -> i32 {
        self.device
    }

    /// The raw handle, for the device-pointer entry points of `ffi`.
    pub fn raw(&self) -> *mut ffi::SpxHandle {
        self.handle
    }

    /// Waits for every call enqueued on this context's stream.
    pub fn synchronize(&self) -> Result<()> {
        check(unsafe { ffi::spx_synchronize(self.handle) })
    }
}

impl Drop for HipContext {
    fn drop(&mut self) {
        unsafe {
            ffi::spx_destroy(self.handle);
        }
    }
}

/// The thresholds of spicier-batched-sweep's `GpuBatchConfig`
/// (solver.rs:6-41); the same fields as spicier-backend-cuda's config.
#[derive(Debug, Clone)]
pub struct GpuBatchedSweepConfig {
    pub min_batch_size: usize,
    pub min_matrix_size: usize,
    pub max_batch_per_launch: usize,
}

impl Default for GpuBatchedSweepConfig {
    fn default() -> Self {
        Self {
            min_batch_size: 16,
            min_matrix_size: 32,
            max_batch_per_launch: unsafe { ffi::spx_max_batch_per_launch() } as usize,
        }
    }
}

/// Solutions of a batch, with the indices of its singular systems.
#[derive(Debug, Clone)]
pub struct HipBatchedSolveResult {
    pub solutions: Vec<f64>,
    pub singular_indices: Vec<usize>,
    pub n: usize,
    pub batch_size: usize,
}

/// LU factors of a batch in LAPACK getrf format.
#[derive(Debug, Clone)]
pub struct HipBatchedFactors {
    /// batch * n * n, column-major per system: unit-lower L strictly below
    /// the diagonal, U on and above.
    pub factors: Vec<f64>,
    /// batch * n: at step k row k was swapped with row pivots[k] (0-based).
    pub pivots: Vec<i32>,
    pub singular_indices: Vec<usize>,
}

/// Batched dense LU with partial pivoting (first maximum on ties, as the CPU
/// path's nalgebra LU), singular systems zeroed and listed.
pub struct HipBatchedLuSolver {
    ctx: Arc<HipContext>,
    config: GpuBatchedSweepConfig,
}

impl HipBatchedLuSolver {
    pub fn new(ctx: Arc<HipContext>) -> Self {
        Self::with_config(ctx, GpuBatchedSweepConfig::default())
    }

    pub fn with_config(ctx: Arc<HipContext>, config: GpuBatchedSweepConfig) -> Self {
        Self { ctx, config }
    }

    pub fn config(&self) -> &GpuBatchedSweepConfig {
        &self.config
    }

    pub fn should_use_gpu(&self, matrix_size: usize, batch_size: usize) -> bool {
        matrix_size >= self.config.min_matrix_size && batch_size >= self.config.min_batch_size
    }

    fn check_lengths(m: usize, r: usize, n: usize, batch: usize) -> Result<()> {
        // the checks every backend makes (solver.rs:336-352)
        if m != batch * n * n {
            return Err(HipError::InvalidDimension(format!(
                "Expected {} matrix elements, got {}",
                batch * n * n,
                m
            )));
        }
        if r != batch * n {
            return Err(HipError::InvalidDimension(format!(
                "Expected {} RHS elements, got {}",
                batch * n,
                r
            )));
        }
        Ok(())
    }

    /// Solves `batch_size` systems; matrices column-major per system,
    /// batch-major (solver.rs:112-113).  Batches larger than
    /// `max_batch_per_launch` are split.
    pub fn solve_batch(
        &self,
        matrices: &[f64],
        rhs: &[f64],
        n: usize,
        batch_size: usize,
    ) -> Result<HipBatchedSolveResult> {
        Self::check_lengths(matrices.len(), rhs.len(), n, batch_size)?;
        let mut solutions = vec![0.0f64; batch_size * n];
        let mut singular = vec![0u8; batch_size];
        let step = self.config.max_batch_per_launch.max(1);
        let mut b = 0;
        while b < batch_size {
            let cnt = step.min(batch_size - b);
            check(unsafe {
                ffi::spx_solve_batch_f64(
                    self.ctx.raw(),
                    matrices[b * n * n..].as_ptr(),
                    rhs[b * n..].as_ptr(),
                    n as i64,
                    cnt as i64,
                    solutions[b * n..].as_mut_ptr(),
                    singular[b..].as_mut_ptr(),
                    ptr::null_mut(),
                )
            })?;
            b += cnt;
        }
        let singular_indices =
            singular.iter().enumerate().filter(|(_, &s)| s != 0).map(|(i, _)| i).collect();
        Ok(HipBatchedSolveResult { solutions, singular_indices, n, batch_size })
    }

    /// The getrf half alone (cublasDgetrfBatched's place in
    /// spicier-backend-cuda/src/batched_lu.rs:404-428).
    pub fn factor_batch(
        &self,
        matrices: &[f64],
        n: usize,
        batch_size: usize,
    ) -> Result<HipBatchedFactors> {
        Self::check_lengths(matrices.len(), batch_size * n, n, batch_size)?;
        let mut factors = vec![0.0f64; batch_size * n * n];
        let mut pivots = vec![0i32; batch_size * n];
        let mut singular = vec![0u8; batch_size];
        if batch_size > 0 {
            check(unsafe {
                ffi::spx_lu_factor_batch_f64(
                    self.ctx.raw(),
                    matrices.as_ptr(),
                    n as i64,
                    batch_size as i64,
                    factors.as_mut_ptr(),
                    pivots.as_mut_ptr(),
                    singular.as_mut_ptr(),
                    ptr::null_mut(),
                )
            })?;
        }
        let singular_indices =
            singular.iter().enumerate().filter(|(_, &s)| s != 0).map(|(i, _)| i).collect();
        Ok(HipBatchedFactors { factors, pivots, singular_indices })
    }
}

/// Shared-pattern sparse batched solve (FaerTripletBatchedSolver,
/// faer_sparse_solver.rs:230-362): one symbolic analysis per pattern, cached
/// on the context.
pub struct HipSparseBatchedSolver {
    ctx: Arc<HipContext>,
}

impl HipSparseBatchedSolver {
    pub fn new(ctx: Arc<HipContext>) -> Self {
        Self { ctx }
    }

    /// `values` is batch * rows.len(): system i's triplet values; duplicates
    /// are summed in triplet order.
    pub fn solve_batch_triplets(
        &self,
        n: usize,
        rows: &[i32],
        cols: &[i32],
        values: &[f64],
        rhs: &[f64],
        batch_size: usize,
    ) -> Result<HipBatchedSolveResult> {
        let nnz = rows.len();
        if cols.len() != nnz || values.len() != nnz * batch_size || rhs.len() != n * batch_size {
            return Err(HipError::InvalidDimension(format!(
                "triplets {} / {} / {} values for batch {} and n {}",
                rows.len(),
                cols.len(),
                values.len(),
                batch_size,
                n
            )));
        }
        let mut solutions = vec![0.0f64; batch_size * n];
        let mut singular = vec![0u8; batch_size];
        if batch_size > 0 {
            check(unsafe {
                ffi::spx_sparse_solve_batch_f64(
                    self.ctx.raw(),
                    n as i64,
                    nnz as i64,
                    rows.as_ptr(),
                    cols.as_ptr(),
                    values.as_ptr(),
                    rhs.as_ptr(),
                    batch_size as i64,
                    solutions.as_mut_ptr(),
                    singular.as_mut_ptr(),
                    ptr::null_mut(),
                )
            })?;
        }
        let singular_indices =
            singular.iter().enumerate().filter(|(_, &s)| s != 0).map(|(i, _)| i).collect();
        Ok(HipBatchedSolveResult { solutions, singular_indices, n, batch_size })
    }
}

/// Per-instance results of a DC-operating-point sweep.
#[derive(Debug, Clone)]
pub struct SweepResult {
    pub n: usize,
    pub batch_size: usize,
    /// batch * n: node voltages then branch currents.
    pub x: Vec<f64>,
    /// `ConvergenceStatus` codes (SPX_ACTIVE .. SPX_SINGULAR).
    pub status: Vec<i32>,
    pub iterations: Vec<i32>,
    /// SPX_FLAG_* bits.
    pub flags: Vec<i32>,
    /// max_i |KCL_i| / tol_i of the final solution (<= 1 passes).
    pub residual: Vec<f64>,
    /// batch * num_params sampled parameter values.
    pub params: Vec<f64>,
    /// n * 5: count, sum, sum of squares, min, max of each unknown over the
    /// accepted instances.
    pub node_stats: Vec<f64>,
}

impl SweepResult {
    /// Instances that converged and passed the FP64 KCL test.
    pub fn accepted(&self) -> usize {
        self.status
            .iter()
            .zip(&self.flags)
            .filter(|(&s, &f)| s == ffi::SPX_CONVERGED && f & ffi::SPX_FLAG_KCL_FAIL == 0)
            .count()
    }
}

/// The Newton, stepping and acceptance defaults of ConvergenceCriteria,
/// GminSteppingParams and SourceSteppingParams (newton.rs:24-34,185-194,222-231).
pub fn nr_config_default() -> ffi::SpxNrConfig {
    let mut c = std::mem::MaybeUninit::<ffi::SpxNrConfig>::zeroed();
    unsafe {
        ffi::spx_nr_config_default(c.as_mut_ptr());
        c.assume_init()
    }
}

/// Monte Carlo / corner / parameter sweep of a circuit's DC operating point
/// on the device: the circuit crosses as its element table (the stamper of
/// solve_batched_sweep_gpu, sweep.rs:101-217, cannot cross an ABI), the
/// parameters are drawn on the device from `specs` and `seed`.
pub fn sweep_dcop(
    ctx: &HipContext,
    num_nodes: i32,
    elements: &[ffi::SpxElement],
    specs: &[ffi::SpxParamSpec],
    generator: i32,
    seed: u64,
    batch_size: usize,
    config: &ffi::SpxNrConfig,
) -> Result<SweepResult> {
    let ckt = ffi::SpxCircuit {
        num_nodes,
        num_elements: elements.len() as i32,
        elements: elements.as_ptr(),
        num_params: specs.len() as i32,
    };
    let n = unsafe { ffi::spx_circuit_size(&ckt) };
    if n <= 0 {
        return Err(HipError::InvalidArgument(last_error()));
    }
    let n = n as usize;
    let spec = ffi::SpxSweepSpec {
        generator,
        seed,
        batch: batch_size as i64,
        sweep_offset: 0,
        params: specs.as_ptr(),
        explicit_params: ptr::null(),
    };
    let mut r = SweepResult {
        n,
        batch_size,
        x: vec![0.0; batch_size * n],
        status: vec![0; batch_size],
        iterations: vec![0; batch_size],
        flags: vec![0; batch_size],
        residual: vec![0.0; batch_size],
        params: vec![0.0; batch_size * specs.len()],
        node_stats: vec![0.0; n * 5],
    };
    let mut out = ffi::SpxSweepResult {
        x: r.x.as_mut_ptr(),
        status: r.status.as_mut_ptr(),
        iterations: r.iterations.as_mut_ptr(),
        flags: r.flags.as_mut_ptr(),
        residual: r.residual.as_mut_ptr(),
        params: r.params.as_mut_ptr(),
        node_stats: r.node_stats.as_mut_ptr(),
        factorizations: ptr::null_mut(),
        refinements: ptr::null_mut(),
    };
    check(unsafe { ffi::spx_sweep_dcop(ctx.raw(), &ckt, &spec, config, ptr::null(), &mut out) })?;
    Ok(r)
}

#[cfg(test)]
mod tests {
    use super::*;

    fn ctx() -> Option<Arc<HipContext>> {
        match device_count() {
            Ok(n) if n > 0 => HipContext::shared(0).ok(),
            _ => None,
        }
    }

    #[test]
    fn known_answers() {
        // solver.rs:408-460: two 2x2 identities, [[2,1],[1,3]] x = [5,10],
        // and a singular second system
        let Some(c) = ctx() else { return };
        let s = HipBatchedLuSolver::new(c);
        let r = s.solve_batch(&[1., 0., 0., 1., 1., 0., 0., 1.], &[1., 2., 3., 4.], 2, 2).unwrap();
        assert!(r.singular_indices.is_empty());
        for (a, b) in r.solutions.iter().zip([1., 2., 3., 4.]) {
            assert!((a - b).abs() < 1e-15);
        }
        let r = s.solve_batch(&[2., 1., 1., 3.], &[5., 10.], 2, 1).unwrap();
        assert!((r.solutions[0] - 1.0).abs() < 1e-12 && (r.solutions[1] - 3.0).abs() < 1e-12);
        let r = s.solve_batch(&[1., 0., 0., 1., 1., 1., 2., 2.], &[1., 2., 1., 2.], 2, 2).unwrap();
        assert_eq!(r.singular_indices, vec![1]);
        assert!(r.solutions[2..].iter().all(|&v| v == 0.0));
        assert!(matches!(
            s.solve_batch(&[1.0, 2.0], &[1.0], 2, 1),
            Err(HipError::InvalidDimension(_))
        ));
        let r = s.solve_batch(&[], &[], 3, 0).unwrap();
        assert_eq!(r.batch_size, 0);
    }

    #[test]
    fn divider_sweep() {
        // V1 = 10 V across R1 = R2 = 1 kOhm nominal, +-10 % uniform
        let Some(c) = ctx() else { return };
        let el = |kind, a, b, p| ffi::SpxElement {
            kind,
            node: [a, b, 0, 0],
            param: [p, -1, -1, -1, -1, -1, -1],
            ctrl: -1,
        };
        let elements = [
            el(ffi::SPX_EL_VSOURCE, 1, 0, 0),
            el(ffi::SPX_EL_RESISTOR, 1, 2, 1),
            el(ffi::SPX_EL_RESISTOR, 2, 0, 2),
        ];
        let spec = |kind, slot, a, b| ffi::SpxParamSpec { kind, rng_index: slot, a, b, c: 0.0, d: 0.0 };
        let specs = [
            spec(ffi::SPX_P_CONST, 0, 10.0, 0.0),
            spec(ffi::SPX_P_UNIFORM, 1, 900.0, 1100.0),
            spec(ffi::SPX_P_UNIFORM, 2, 900.0, 1100.0),
        ];
        let r = sweep_dcop(&c, 2, &elements, &specs, ffi::SPX_GEN_HASH, 42, 100,
                           &nr_config_default()).unwrap();
        assert_eq!(r.accepted(), 100);
        for i in 0..100 {
            let (r1, r2) = (r.params[i * 3 + 1], r.params[i * 3 + 2]);
            assert!((r.x[i * 3 + 1] - 10.0 * r2 / (r1 + r2)).abs() < 1e-12);
        }
    }
}
