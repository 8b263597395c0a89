//! Raw bindings to `libspicier_hip.so`, one to one with
//! `include/spicier_hip.h` (structs `#[repr(C)]` in the header's field order,
//! enumerators as `c_int` constants).  `tests/test_rust_ffi.py` checks this
//! file against the header: every function with its argument types, every
//! struct field in order, every constant's value.
#![allow(non_camel_case_types, non_snake_case, dead_code)]

use std::os::raw::{c_char, c_int, c_void};

// ------------------------------------------------------------- constants
pub const SPX_OK: c_int = 0;
pub const SPX_ERR_INVALID_DIMENSION: c_int = 1;
pub const SPX_ERR_BATCH_TOO_LARGE: c_int = 2;
pub const SPX_ERR_MEMORY_ALLOC: c_int = 3;
pub const SPX_ERR_TRANSFER: c_int = 4;
pub const SPX_ERR_BACKEND_INIT: c_int = 5;
pub const SPX_ERR_BACKEND: c_int = 6;
pub const SPX_ERR_NO_BACKEND: c_int = 7;
pub const SPX_ERR_INVALID_ARGUMENT: c_int = 8;
pub const SPX_ACTIVE: c_int = 0;
pub const SPX_CONVERGED: c_int = 1;
pub const SPX_FAILED: c_int = 2;
pub const SPX_SINGULAR: c_int = 3;
pub const SPX_FLAG_KCL_FAIL: c_int = 1;
pub const SPX_FLAG_GMIN_STEPPING: c_int = 2;
pub const SPX_FLAG_SOURCE_STEPPING: c_int = 4;
pub const SPX_FLAG_FP64_ESCALATED: c_int = 8;
pub const SPX_FLAG_WARM_START: c_int = 16;
pub const SPX_EL_RESISTOR: c_int = 0;
pub const SPX_EL_CAPACITOR: c_int = 1;
pub const SPX_EL_VSOURCE: c_int = 2;
pub const SPX_EL_ISOURCE: c_int = 3;
pub const SPX_EL_DIODE: c_int = 4;
pub const SPX_EL_NMOS: c_int = 5;
pub const SPX_EL_PMOS: c_int = 6;
pub const SPX_EL_VPULSE: c_int = 7;
pub const SPX_EL_VSIN: c_int = 8;
pub const SPX_EL_INDUCTOR: c_int = 9;
pub const SPX_EL_VCVS: c_int = 10;
pub const SPX_EL_VCCS: c_int = 11;
pub const SPX_EL_CCCS: c_int = 12;
pub const SPX_EL_CCVS: c_int = 13;
pub const SPX_EL_VPWL: c_int = 14;
pub const SPX_P_CONST: c_int = 0;
pub const SPX_P_GAUSS_REL: c_int = 1;
pub const SPX_P_GAUSS: c_int = 2;
pub const SPX_P_UNIFORM: c_int = 3;
pub const SPX_P_LOGUNIFORM: c_int = 4;
pub const SPX_GEN_HASH: c_int = 0;
pub const SPX_GEN_LCG: c_int = 1;
pub const SPX_GEN_CORNER: c_int = 2;
pub const SPX_GEN_LINEAR: c_int = 3;
pub const SPX_GEN_EXPLICIT: c_int = 4;
pub const SPX_PREC_F64: c_int = 0;
pub const SPX_PREC_F32_IR: c_int = 1;
pub const SPX_LIMIT_REFERENCE: c_int = 0;
pub const SPX_LIMIT_NONE: c_int = 1;
pub const SPX_LIMIT_PNJLIM: c_int = 2;
pub const SPX_WARM_ZERO: c_int = 0;
pub const SPX_WARM_GIVEN: c_int = 1;
pub const SPX_WARM_SURROGATE: c_int = 2;
pub const SPX_TRAN_BE: c_int = 0;
pub const SPX_TRAN_TRAP: c_int = 1;
pub const SPX_TRAN_TRBDF2: c_int = 2;
pub const SPX_MIXED_GRAPH: c_int = 1;
pub const SPX_MIXED_AUTOTUNE: c_int = 2;
pub const SPX_KERNEL_AUTO: c_int = 0;
pub const SPX_KERNEL_DENSE: c_int = 1;
pub const SPX_KERNEL_SPARSE64: c_int = 2;
pub const SPX_KERNEL_SPARSE256: c_int = 3;
pub const SPX_KERNEL_LANES: c_int = 4;
pub const SPX_SELECT_ACTIVE: c_int = 0;
pub const SPX_SELECT_CONVERGED: c_int = 1;
pub const SPX_SELECT_FAILED: c_int = 2;
pub const SPX_WINDOW_RECTANGULAR: c_int = 0;
pub const SPX_WINDOW_HANNING: c_int = 1;
pub const SPX_WINDOW_HAMMING: c_int = 2;
pub const SPX_WINDOW_BLACKMAN: c_int = 3;
pub const SPX_MAX_ELEMENT_PARAMS: usize = 7;

// ------------------------------------------------------------- types
#[repr(C)]
pub struct SpxHandle {
    _opaque: [u8; 0],
}

#[repr(C)]
pub struct SpxSurrogate {
    _opaque: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxGmresConfig {
    pub restart: i32,
    pub max_iter: i32,
    pub tol: f64,
    pub jacobi: i32,
    pub precision: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxElement {
    pub kind: i32,
    pub node: [i32; 4],
    pub param: [i32; SPX_MAX_ELEMENT_PARAMS],
    pub ctrl: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxCircuit {
    pub num_nodes: i32,
    pub num_elements: i32,
    pub elements: *const SpxElement,
    pub num_params: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxParamSpec {
    pub kind: i32,
    pub rng_index: i32,
    pub a: f64,
    pub b: f64,
    pub c: f64,
    pub d: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxSweepSpec {
    pub generator: i32,
    pub seed: u64,
    pub batch: i64,
    pub sweep_offset: i64,
    pub params: *const SpxParamSpec,
    pub explicit_params: *const f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxNrConfig {
    pub v_abstol: f64,
    pub v_reltol: f64,
    pub i_abstol: f64,
    pub max_iterations: i32,
    pub gmin_stepping: i32,
    pub gmin_initial: f64,
    pub gmin_final: f64,
    pub gmin_factor: f64,
    pub gmin_max_attempts: i32,
    pub source_stepping: i32,
    pub src_initial: f64,
    pub src_step: f64,
    pub src_min_step: f64,
    pub src_max_attempts: i32,
    pub precision: i32,
    pub limiter: i32,
    pub warm_start: i32,
    pub ir_max_steps: i32,
    pub ir_tol: f64,
    pub kcl_reltol: f64,
    pub kcl_abstol: f64,
    pub kcl_vntol: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxSweepResult {
    pub x: *mut f64,
    pub status: *mut i32,
    pub iterations: *mut i32,
    pub flags: *mut i32,
    pub residual: *mut f64,
    pub params: *mut f64,
    pub node_stats: *mut f64,
    pub factorizations: *mut i32,
    pub refinements: *mut i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxTranAdaptive {
    pub tstop: f64,
    pub h_init: f64,
    pub h_min: f64,
    pub h_max: f64,
    pub reltol: f64,
    pub abstol: f64,
    pub max_points: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxTranConfig {
    pub tstep: f64,
    pub steps: i32,
    pub num_probes: i32,
    pub probes: *const i32,
    pub method: i32,
    pub adaptive: *const SpxTranAdaptive,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxTranResult {
    pub x: *mut f64,
    pub status: *mut i32,
    pub iterations: *mut i32,
    pub flags: *mut i32,
    pub residual: *mut f64,
    pub params: *mut f64,
    pub waveforms: *mut f64,
    pub fail_step: *mut i32,
    pub times: *mut f64,
    pub num_points: *mut i32,
    pub total_steps: *mut i32,
    pub rejected_steps: *mut i32,
    pub min_step: *mut f64,
    pub max_step: *mut f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxTranClass {
    pub circuit: *const SpxCircuit,
    pub spec: SpxSweepSpec,
    pub tran: SpxTranConfig,
    pub result: SpxTranResult,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxTrackerSummary {
    pub total_points: i64,
    pub converged_count: i64,
    pub failed_count: i64,
    pub active_count: i64,
    pub iteration_sum: i64,
    pub max_iterations: i32,
    pub average_iterations: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxYieldSpec {
    pub node: i32,
    pub lower_limit: f64,
    pub upper_limit: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxPipelineConfig {
    pub chunk_size: i64,
    pub num_buffers: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxSpectralConfig {
    pub window: i32,
    pub zero_padding: i32,
    pub fft_size: i64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxWaveformBatch {
    pub data: *const f64,
    pub batch: i64,
    pub batch_stride: i64,
    pub num_points: i32,
    pub point_stride: i32,
    pub times: *const f64,
    pub sample_rate: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxSpectrumInfo {
    pub num_samples: i64,
    pub fft_size: i64,
    pub num_bins: i64,
    pub sample_rate: f64,
    pub freq_resolution: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxSpectrumOut {
    pub magnitude: *mut f64,
    pub magnitude_db: *mut f64,
    pub phase: *mut f64,
    pub psd: *mut f64,
    pub mean_magnitude: *mut f64,
    pub std_magnitude: *mut f64,
    pub min_magnitude: *mut f64,
    pub max_magnitude: *mut f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxThdOut {
    pub thd_percent: *mut f64,
    pub thd_db: *mut f64,
    pub fundamental_freq: *mut f64,
    pub fundamental_magnitude: *mut f64,
    pub harmonic_rms: *mut f64,
    pub harmonics: *mut f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct SpxThdStatistics {
    pub mean_thd: f64,
    pub std_thd: f64,
    pub min_thd: f64,
    pub max_thd: f64,
    pub median_thd: f64,
    pub yield_percent: f64,
    pub num_samples: i64,
    pub thd_spec: f64,
}

// ------------------------------------------------------------- functions
extern "C" {
    pub fn spx_device_count(count: *mut i32) -> c_int;
    pub fn spx_create(device_id: i32, out: *mut *mut SpxHandle) -> c_int;
    pub fn spx_destroy(h: *mut SpxHandle) -> c_int;
    pub fn spx_last_error() -> *const c_char;
    pub fn spx_synchronize(h: *mut SpxHandle) -> c_int;
    pub fn spx_stream(h: *mut SpxHandle) -> *mut c_void;
    pub fn spx_solve_batch_f64(h: *mut SpxHandle, matrices: *const f64, rhs: *const f64, n: i64, batch: i64, solutions: *mut f64, singular: *mut u8, num_singular: *mut i64) -> c_int;
    pub fn spx_solve_batch_f64_device(h: *mut SpxHandle, d_matrices: *const f64, d_rhs: *const f64, n: i64, batch: i64, d_solutions: *mut f64, d_singular: *mut u8) -> c_int;
    pub fn spx_solve_batch_f64_device_noalloc(h: *mut SpxHandle, d_matrices: *const f64, d_rhs: *const f64, n: i64, batch: i64, d_solutions: *mut f64, d_singular: *mut u8) -> c_int;
    pub fn spx_lu_factor_batch_f64(h: *mut SpxHandle, matrices: *const f64, n: i64, batch: i64, factors: *mut f64, pivots: *mut i32, singular: *mut u8, num_singular: *mut i64) -> c_int;
    pub fn spx_lu_factor_batch_f64_device(h: *mut SpxHandle, d_matrices: *const f64, n: i64, batch: i64, d_factors: *mut f64, d_pivots: *mut i32, d_singular: *mut u8) -> c_int;
    pub fn spx_max_batch_per_launch() -> i64;
    pub fn spx_sparse_solve_batch_f64(h: *mut SpxHandle, n: i64, nnz: i64, rows: *const i32, cols: *const i32, values: *const f64, rhs: *const f64, batch: i64, solutions: *mut f64, singular: *mut u8, num_singular: *mut i64) -> c_int;
    pub fn spx_sparse_solve_batch_f64_device(h: *mut SpxHandle, n: i64, nnz: i64, rows: *const i32, cols: *const i32, d_values: *const f64, d_rhs: *const f64, batch: i64, d_solutions: *mut f64, d_singular: *mut u8) -> c_int;
    pub fn spx_gmres_config_default(cfg: *mut SpxGmresConfig);
    pub fn spx_gmres_solve_batch_f64(h: *mut SpxHandle, n: i64, nnz: i64, rows: *const i32, cols: *const i32, values: *const f64, rhs: *const f64, x0: *const f64, batch: i64, cfg: *const SpxGmresConfig, solutions: *mut f64, iterations: *mut i32, residuals: *mut f64, converged: *mut u8) -> c_int;
    pub fn spx_gmres_solve_batch_f64_device(h: *mut SpxHandle, n: i64, nnz: i64, rows: *const i32, cols: *const i32, d_values: *const f64, d_rhs: *const f64, d_x0: *const f64, batch: i64, cfg: *const SpxGmresConfig, d_solutions: *mut f64, d_iterations: *mut i32, d_residuals: *mut f64, d_converged: *mut u8) -> c_int;
    pub fn spx_sample_params(h: *mut SpxHandle, spec: *const SpxSweepSpec, num_params: i32, params: *mut f64) -> c_int;
    pub fn spx_sample_params_device(h: *mut SpxHandle, spec: *const SpxSweepSpec, num_params: i32, d_params: *mut f64) -> c_int;
    pub fn spx_nr_config_default(cfg: *mut SpxNrConfig);
    pub fn spx_sweep_dcop(h: *mut SpxHandle, ckt: *const SpxCircuit, spec: *const SpxSweepSpec, cfg: *const SpxNrConfig, x0: *const f64, out: *mut SpxSweepResult) -> c_int;
    pub fn spx_sweep_dcop_device(h: *mut SpxHandle, ckt: *const SpxCircuit, spec: *const SpxSweepSpec, cfg: *const SpxNrConfig, d_x0: *const f64, d_out: *mut SpxSweepResult) -> c_int;
    pub fn spx_transient_sweep(h: *mut SpxHandle, ckt: *const SpxCircuit, spec: *const SpxSweepSpec, cfg: *const SpxNrConfig, tran: *const SpxTranConfig, out: *mut SpxTranResult) -> c_int;
    pub fn spx_transient_sweep_device(h: *mut SpxHandle, ckt: *const SpxCircuit, spec: *const SpxSweepSpec, cfg: *const SpxNrConfig, tran: *const SpxTranConfig, d_out: *mut SpxTranResult) -> c_int;
    pub fn spx_transient_mixed_device(h: *mut SpxHandle, classes: *const SpxTranClass, num_classes: i32, cfg: *const SpxNrConfig, options: u32, kernels: *mut i32) -> c_int;
    pub fn spx_circuit_size(ckt: *const SpxCircuit) -> i32;
    pub fn spx_surrogate_create(h: *mut SpxHandle, n_in: i32, n_hidden: i32, n_out: i32, seed: u64, out: *mut *mut SpxSurrogate) -> c_int;
    pub fn spx_surrogate_destroy(s: *mut SpxSurrogate) -> c_int;
    pub fn spx_surrogate_set_inputs(s: *mut SpxSurrogate, slots: *const i32, P: i32) -> c_int;
    pub fn spx_surrogate_num_weights(s: *const SpxSurrogate) -> i64;
    pub fn spx_surrogate_get_weights(s: *mut SpxSurrogate, host: *mut f32, count: i64) -> c_int;
    pub fn spx_surrogate_set_weights(s: *mut SpxSurrogate, host: *const f32, count: i64) -> c_int;
    pub fn spx_surrogate_set_normalization(s: *mut SpxSurrogate, in_mean: *const f32, in_std: *const f32, out_mean: *const f32, out_std: *const f32) -> c_int;
    pub fn spx_surrogate_train_device(s: *mut SpxSurrogate, d_params: *const f64, d_x: *const f64, n_x: i32, d_mask: *const u8, B: i64, epochs: i32, batch: i32, lr: f32, tol_rel: f32, tol_abs: f32, shuffle_seed: u64, loss_history: *mut f32) -> c_int;
    pub fn spx_surrogate_predict_device(s: *mut SpxSurrogate, d_params: *const f64, B: i64, n_x: i32, d_x0: *mut f64) -> c_int;
    pub fn spx_surrogate_gradient_device(s: *mut SpxSurrogate, d_in: *const f32, d_target: *const f32, B: i64, d_grad: *mut f32, loss: *mut f32) -> c_int;
    pub fn spx_set_surrogate(h: *mut SpxHandle, s: *mut SpxSurrogate) -> c_int;
    pub fn spx_stats_reduce_device(h: *mut SpxHandle, d_values: *const f64, d_mask: *const u8, batch: i64, n: i32, d_out: *mut f64) -> c_int;
    pub fn spx_tracker_check_convergence_device(h: *mut SpxHandle, d_status: *mut i32, batch: i64, n: i32, d_changes: *const f64, d_solutions: *const f64, abstol: f64, reltol: f64, d_newly_converged: *mut i64) -> c_int;
    pub fn spx_tracker_check_residual_device(h: *mut SpxHandle, d_status: *mut i32, batch: i64, n: i32, d_residuals: *const f64, tolerance: f64, d_newly_converged: *mut i64) -> c_int;
    pub fn spx_tracker_increment_device(h: *mut SpxHandle, d_status: *mut i32, d_iterations: *mut i32, batch: i64, max_iterations: i32, indices: *const i64, count: i64) -> c_int;
    pub fn spx_tracker_mark_device(h: *mut SpxHandle, d_status: *mut i32, batch: i64, indices: *const i64, count: i64, new_status: i32, marked: *mut i64) -> c_int;
    pub fn spx_tracker_summary_device(h: *mut SpxHandle, d_status: *const i32, d_iterations: *const i32, batch: i64, out: *mut SpxTrackerSummary) -> c_int;
    pub fn spx_select_indices_device(h: *mut SpxHandle, d_status: *const i32, batch: i64, which: i32, d_indices: *mut i64, d_count: *mut i64) -> c_int;
    pub fn spx_compact_rows_device(h: *mut SpxHandle, d_data: *const f64, item_size: i64, d_indices: *const i64, count: i64, d_out: *mut f64) -> c_int;
    pub fn spx_expand_rows_device(h: *mut SpxHandle, d_compacted: *const f64, item_size: i64, d_indices: *const i64, count: i64, batch: i64, fill_value: f64, d_out: *mut f64) -> c_int;
    pub fn spx_histogram_device(h: *mut SpxHandle, d_solutions: *const f64, batch: i64, n: i32, node: i32, num_bins: i32, min: f64, max: f64, from_sweep: i32, d_counts: *mut i64, d_range: *mut f64) -> c_int;
    pub fn spx_yield_device(h: *mut SpxHandle, d_solutions: *const f64, batch: i64, n: i32, specs: *const SpxYieldSpec, num_specs: i32, d_counts: *mut i64) -> c_int;
    pub fn spx_stats_accumulate_device(h: *mut SpxHandle, d_values: *const f64, d_mask: *const u8, batch: i64, n: i32, d_acc: *mut f64) -> c_int;
    pub fn spx_stats_merge_device(h: *mut SpxHandle, d_acc: *mut f64, d_other: *const f64, n: i32) -> c_int;
    pub fn spx_pack_matrices_f32_device(h: *mut SpxHandle, d_matrices: *const f64, n: i64, batch: i64, row_stride: i64, d_packed: *mut f32) -> c_int;
    pub fn spx_pack_rhs_f32_device(h: *mut SpxHandle, d_rhs: *const f64, count: i64, d_packed: *mut f32) -> c_int;
    pub fn spx_unpack_solutions_f64_device(h: *mut SpxHandle, d_packed: *const f32, count: i64, d_solutions: *mut f64) -> c_int;
    pub fn spx_pipeline_config_auto(n: i64, total_points: i64, cfg: *mut SpxPipelineConfig);
    pub fn spx_solve_batch_f64_pipelined(h: *mut SpxHandle, matrices: *const f64, rhs: *const f64, n: i64, batch: i64, cfg: *const SpxPipelineConfig, solutions: *mut f64, singular: *mut u8, num_singular: *mut i64, chunks_processed: *mut i64) -> c_int;
    pub fn spx_spectral_config_default(cfg: *mut SpxSpectralConfig);
    pub fn spx_spectrum_info_query(h: *mut SpxHandle, wf: *const SpxWaveformBatch, cfg: *const SpxSpectralConfig, info: *mut SpxSpectrumInfo) -> c_int;
    pub fn spx_batch_fft_device(h: *mut SpxHandle, wf: *const SpxWaveformBatch, cfg: *const SpxSpectralConfig, out: *const SpxSpectrumOut, info: *mut SpxSpectrumInfo) -> c_int;
    pub fn spx_batch_thd_device(h: *mut SpxHandle, wf: *const SpxWaveformBatch, fundamental_freq: f64, num_harmonics: i32, out: *const SpxThdOut, found_harmonics: *mut i32) -> c_int;
    pub fn spx_thd_statistics_device(h: *mut SpxHandle, d_thd_percent: *const f64, count: i64, thd_spec: f64, stats: *mut SpxThdStatistics) -> c_int;
    pub fn spx_window_coefficients(window: i32, n: i64, out: *mut f64) -> c_int;
    pub fn spx_window_coherent_gain(window: i32, n: i64) -> f64;
}

