"""ctypes mirror of include/spicier_hip.h and the loader of libspicier_hip.so.

The structures here are the boundary format only; all computation happens in
the HIP library.  ``load_library()`` raises if the library is missing — there
is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspicier_hip.so")

# spx_status (error.rs:7-46)
SPX_OK = 0
SPX_ERR_INVALID_DIMENSION = 1
SPX_ERR_BATCH_TOO_LARGE = 2
SPX_ERR_MEMORY_ALLOC = 3
SPX_ERR_TRANSFER = 4
SPX_ERR_BACKEND_INIT = 5
SPX_ERR_BACKEND = 6
SPX_ERR_NO_BACKEND = 7
SPX_ERR_INVALID_ARGUMENT = 8

# spx_conv_status (convergence.rs:41-50)
SPX_ACTIVE, SPX_CONVERGED, SPX_FAILED, SPX_SINGULAR = 0, 1, 2, 3

SPX_FLAG_KCL_FAIL = 1
SPX_FLAG_GMIN_STEPPING = 2
SPX_FLAG_SOURCE_STEPPING = 4
SPX_FLAG_FP64_ESCALATED = 8
SPX_FLAG_WARM_START = 16

# element kinds
SPX_EL_RESISTOR, SPX_EL_CAPACITOR, SPX_EL_VSOURCE, SPX_EL_ISOURCE = 0, 1, 2, 3
SPX_EL_DIODE, SPX_EL_NMOS, SPX_EL_PMOS, SPX_EL_VPULSE = 4, 5, 6, 7
SPX_EL_VSIN = 8
SPX_EL_INDUCTOR = 9
SPX_EL_VCVS, SPX_EL_VCCS, SPX_EL_CCCS, SPX_EL_CCVS = 10, 11, 12, 13
SPX_EL_VPWL = 14
SPX_MAX_ELEMENT_PARAMS = 7

# parameter kinds
SPX_P_CONST, SPX_P_GAUSS_REL, SPX_P_GAUSS, SPX_P_UNIFORM, SPX_P_LOGUNIFORM = 0, 1, 2, 3, 4

# generators
SPX_GEN_HASH, SPX_GEN_LCG, SPX_GEN_CORNER, SPX_GEN_LINEAR, SPX_GEN_EXPLICIT = 0, 1, 2, 3, 4

SPX_PREC_F64, SPX_PREC_F32_IR = 0, 1
SPX_LIMIT_REFERENCE, SPX_LIMIT_NONE, SPX_LIMIT_PNJLIM = 0, 1, 2
SPX_WARM_ZERO, SPX_WARM_GIVEN, SPX_WARM_SURROGATE = 0, 1, 2


class SpxElement(C.Structure):
    _fields_ = [("kind", C.c_int32), ("node", C.c_int32 * 4),
                ("param", C.c_int32 * SPX_MAX_ELEMENT_PARAMS), ("ctrl", C.c_int32)]


class SpxCircuit(C.Structure):
    _fields_ = [("num_nodes", C.c_int32), ("num_elements", C.c_int32),
                ("elements", C.POINTER(SpxElement)), ("num_params", C.c_int32)]


class SpxParamSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("rng_index", C.c_int32),
                ("a", C.c_double), ("b", C.c_double), ("c", C.c_double), ("d", C.c_double)]


class SpxSweepSpec(C.Structure):
    _fields_ = [("generator", C.c_int32), ("seed", C.c_uint64), ("batch", C.c_int64),
                ("sweep_offset", C.c_int64), ("params", C.POINTER(SpxParamSpec)),
                ("explicit_params", C.c_void_p)]


class SpxNrConfig(C.Structure):
    _fields_ = [
        ("v_abstol", C.c_double), ("v_reltol", C.c_double), ("i_abstol", C.c_double),
        ("max_iterations", C.c_int32),
        ("gmin_stepping", C.c_int32), ("gmin_initial", C.c_double), ("gmin_final", C.c_double),
        ("gmin_factor", C.c_double), ("gmin_max_attempts", C.c_int32),
        ("source_stepping", C.c_int32), ("src_initial", C.c_double), ("src_step", C.c_double),
        ("src_min_step", C.c_double), ("src_max_attempts", C.c_int32),
        ("precision", C.c_int32), ("limiter", C.c_int32), ("warm_start", C.c_int32),
        ("ir_max_steps", C.c_int32), ("ir_tol", C.c_double),
        ("kcl_reltol", C.c_double), ("kcl_abstol", C.c_double), ("kcl_vntol", C.c_double),
    ]


class SpxGmresConfig(C.Structure):
    _fields_ = [("restart", C.c_int32), ("max_iter", C.c_int32), ("tol", C.c_double),
                ("jacobi", C.c_int32), ("precision", C.c_int32)]


class SpxSweepResult(C.Structure):
    _fields_ = [("x", C.c_void_p), ("status", C.c_void_p), ("iterations", C.c_void_p),
                ("flags", C.c_void_p), ("residual", C.c_void_p), ("params", C.c_void_p),
                ("node_stats", C.c_void_p), ("factorizations", C.c_void_p),
                ("refinements", C.c_void_p)]


SPX_TRAN_BE, SPX_TRAN_TRAP, SPX_TRAN_TRBDF2 = 0, 1, 2


class SpxTranAdaptive(C.Structure):
    _fields_ = [("tstop", C.c_double), ("h_init", C.c_double), ("h_min", C.c_double),
                ("h_max", C.c_double), ("reltol", C.c_double), ("abstol", C.c_double),
                ("max_points", C.c_int32)]


class SpxTranConfig(C.Structure):
    _fields_ = [("tstep", C.c_double), ("steps", C.c_int32), ("num_probes", C.c_int32),
                ("probes", C.c_void_p), ("method", C.c_int32),
                ("adaptive", C.POINTER(SpxTranAdaptive))]


class SpxTranResult(C.Structure):
    _fields_ = [("x", C.c_void_p), ("status", C.c_void_p), ("iterations", C.c_void_p),
                ("flags", C.c_void_p), ("residual", C.c_void_p), ("params", C.c_void_p),
                ("waveforms", C.c_void_p), ("fail_step", C.c_void_p), ("times", C.c_void_p),
                ("num_points", C.c_void_p), ("total_steps", C.c_void_p),
                ("rejected_steps", C.c_void_p), ("min_step", C.c_void_p),
                ("max_step", C.c_void_p)]


class SpxTranClass(C.Structure):
    _fields_ = [("circuit", C.POINTER(SpxCircuit)), ("spec", SpxSweepSpec),
                ("tran", SpxTranConfig), ("result", SpxTranResult)]


class SpxTrackerSummary(C.Structure):
    _fields_ = [("total_points", C.c_int64), ("converged_count", C.c_int64),
                ("failed_count", C.c_int64), ("active_count", C.c_int64),
                ("iteration_sum", C.c_int64), ("max_iterations", C.c_int32),
                ("average_iterations", C.c_double)]


class SpxYieldSpec(C.Structure):
    _fields_ = [("node", C.c_int32), ("lower_limit", C.c_double), ("upper_limit", C.c_double)]


class SpxPipelineConfig(C.Structure):
    _fields_ = [("chunk_size", C.c_int64), ("num_buffers", C.c_int32)]


SPX_SELECT_ACTIVE, SPX_SELECT_CONVERGED, SPX_SELECT_FAILED = 0, 1, 2

SPX_WINDOW_RECTANGULAR, SPX_WINDOW_HANNING, SPX_WINDOW_HAMMING, SPX_WINDOW_BLACKMAN = 0, 1, 2, 3


class SpxSpectralConfig(C.Structure):
    _fields_ = [("window", C.c_int32), ("zero_padding", C.c_int32), ("fft_size", C.c_int64)]


class SpxWaveformBatch(C.Structure):
    _fields_ = [("data", C.c_void_p), ("batch", C.c_int64), ("batch_stride", C.c_int64),
                ("num_points", C.c_int32), ("point_stride", C.c_int32),
                ("times", C.POINTER(C.c_double)), ("sample_rate", C.c_double)]


class SpxSpectrumInfo(C.Structure):
    _fields_ = [("num_samples", C.c_int64), ("fft_size", C.c_int64), ("num_bins", C.c_int64),
                ("sample_rate", C.c_double), ("freq_resolution", C.c_double)]


class SpxSpectrumOut(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in ("magnitude", "magnitude_db", "phase", "psd",
                                          "mean_magnitude", "std_magnitude", "min_magnitude",
                                          "max_magnitude")]


class SpxThdOut(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in ("thd_percent", "thd_db", "fundamental_freq",
                                          "fundamental_magnitude", "harmonic_rms", "harmonics")]


class SpxThdStatistics(C.Structure):
    _fields_ = [("mean_thd", C.c_double), ("std_thd", C.c_double), ("min_thd", C.c_double),
                ("max_thd", C.c_double), ("median_thd", C.c_double),
                ("yield_percent", C.c_double), ("num_samples", C.c_int64),
                ("thd_spec", C.c_double)]

SPX_MIXED_GRAPH, SPX_MIXED_AUTOTUNE = 1, 2
SPX_KERNEL_AUTO, SPX_KERNEL_DENSE, SPX_KERNEL_SPARSE64, SPX_KERNEL_SPARSE256 = 0, 1, 2, 3
SPX_KERNEL_LANES = 4


def default_nr_config(**overrides) -> SpxNrConfig:
    """Defaults of ConvergenceCriteria / GminSteppingParams / SourceSteppingParams
    (crates/spicier-solver/src/newton.rs:24-34,185-194,222-231)."""
    cfg = SpxNrConfig(
        v_abstol=1e-6, v_reltol=1e-3, i_abstol=1e-12, max_iterations=50,
        gmin_stepping=1, gmin_initial=1e-3, gmin_final=1e-12, gmin_factor=10.0,
        gmin_max_attempts=5,
        source_stepping=1, src_initial=0.1, src_step=0.1, src_min_step=0.01, src_max_attempts=5,
        precision=SPX_PREC_F64, limiter=SPX_LIMIT_REFERENCE, warm_start=SPX_WARM_ZERO,
        ir_max_steps=8, ir_tol=0.0,
        kcl_reltol=1e-3, kcl_abstol=1e-12, kcl_vntol=1e-6)
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


_lib = None


def load_library():
    """Load libspicier_hip.so (built in-tree by ``__graft_entry__.build()``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP engine first (python -c "
            "'import __graft_entry__; __graft_entry__.build()')")
    # torch bundles its own libamdhip64 (same soname, different file name).  If
    # ours were loaded first, torch would later bring in a second HIP runtime
    # and see no GPU; loading torch first makes both resolve to one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    vp, i32, i64, u8p = C.c_void_p, C.c_int32, C.c_int64, C.c_void_p
    sigs = {
        "spx_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
        "spx_create": (C.c_int, [i32, C.POINTER(C.c_void_p)]),
        "spx_destroy": (C.c_int, [vp]),
        "spx_last_error": (C.c_char_p, []),
        "spx_synchronize": (C.c_int, [vp]),
        "spx_stream": (C.c_void_p, [vp]),
        "spx_solve_batch_f64": (C.c_int, [vp, vp, vp, i64, i64, vp, u8p, C.POINTER(C.c_int64)]),
        "spx_solve_batch_f64_device": (C.c_int, [vp, vp, vp, i64, i64, vp, u8p]),
        "spx_solve_batch_f64_device_noalloc": (C.c_int, [vp, vp, vp, i64, i64, vp, u8p]),
        "spx_lu_factor_batch_f64": (C.c_int, [vp, vp, i64, i64, vp, vp, u8p,
                                              C.POINTER(C.c_int64)]),
        "spx_lu_factor_batch_f64_device": (C.c_int, [vp, vp, i64, i64, vp, vp, u8p]),
        "spx_max_batch_per_launch": (C.c_int64, []),
        "spx_sparse_solve_batch_f64": (C.c_int, [vp, i64, i64, vp, vp, vp, vp, i64, vp, u8p,
                                                 C.POINTER(C.c_int64)]),
        "spx_sparse_solve_batch_f64_device": (C.c_int, [vp, i64, i64, vp, vp, vp, vp, i64, vp,
                                                        u8p]),
        "spx_gmres_config_default": (None, [C.POINTER(SpxGmresConfig)]),
        "spx_gmres_solve_batch_f64": (C.c_int, [vp, i64, i64, vp, vp, vp, vp, vp, i64,
                                                C.POINTER(SpxGmresConfig), vp, vp, vp, u8p]),
        "spx_gmres_solve_batch_f64_device": (C.c_int, [vp, i64, i64, vp, vp, vp, vp, vp, i64,
                                                       C.POINTER(SpxGmresConfig), vp, vp, vp,
                                                       u8p]),
        "spx_sample_params": (C.c_int, [vp, C.POINTER(SpxSweepSpec), i32, vp]),
        "spx_sample_params_device": (C.c_int, [vp, C.POINTER(SpxSweepSpec), i32, vp]),
        "spx_nr_config_default": (None, [C.POINTER(SpxNrConfig)]),
        "spx_sweep_dcop": (C.c_int, [vp, C.POINTER(SpxCircuit), C.POINTER(SpxSweepSpec),
                                     C.POINTER(SpxNrConfig), vp, C.POINTER(SpxSweepResult)]),
        "spx_sweep_dcop_device": (C.c_int, [vp, C.POINTER(SpxCircuit), C.POINTER(SpxSweepSpec),
                                            C.POINTER(SpxNrConfig), vp,
                                            C.POINTER(SpxSweepResult)]),
        "spx_circuit_size": (C.c_int32, [C.POINTER(SpxCircuit)]),
        "spx_transient_mixed_device": (C.c_int, [vp, C.POINTER(SpxTranClass), i32,
                                                 C.POINTER(SpxNrConfig), C.c_uint32, vp]),
        "spx_transient_sweep": (C.c_int, [vp, C.POINTER(SpxCircuit), C.POINTER(SpxSweepSpec),
                                          C.POINTER(SpxNrConfig), C.POINTER(SpxTranConfig),
                                          C.POINTER(SpxTranResult)]),
        "spx_transient_sweep_device": (C.c_int, [vp, C.POINTER(SpxCircuit),
                                                 C.POINTER(SpxSweepSpec), C.POINTER(SpxNrConfig),
                                                 C.POINTER(SpxTranConfig),
                                                 C.POINTER(SpxTranResult)]),
        "spx_stats_reduce_device": (C.c_int, [vp, vp, vp, i64, i32, vp]),
        "spx_tracker_check_convergence_device": (C.c_int, [vp, vp, i64, i32, vp, vp, C.c_double,
                                                           C.c_double, vp]),
        "spx_tracker_check_residual_device": (C.c_int, [vp, vp, i64, i32, vp, C.c_double, vp]),
        "spx_tracker_increment_device": (C.c_int, [vp, vp, vp, i64, i32, vp, i64]),
        "spx_tracker_mark_device": (C.c_int, [vp, vp, i64, vp, i64, i32, C.POINTER(C.c_int64)]),
        "spx_tracker_summary_device": (C.c_int, [vp, vp, vp, i64, C.POINTER(SpxTrackerSummary)]),
        "spx_select_indices_device": (C.c_int, [vp, vp, i64, i32, vp, vp]),
        "spx_compact_rows_device": (C.c_int, [vp, vp, i64, vp, i64, vp]),
        "spx_expand_rows_device": (C.c_int, [vp, vp, i64, vp, i64, i64, C.c_double, vp]),
        "spx_histogram_device": (C.c_int, [vp, vp, i64, i32, i32, i32, C.c_double, C.c_double,
                                           i32, vp, vp]),
        "spx_yield_device": (C.c_int, [vp, vp, i64, i32, C.POINTER(SpxYieldSpec), i32, vp]),
        "spx_stats_accumulate_device": (C.c_int, [vp, vp, vp, i64, i32, vp]),
        "spx_stats_merge_device": (C.c_int, [vp, vp, vp, i32]),
        "spx_pack_matrices_f32_device": (C.c_int, [vp, vp, i64, i64, i64, vp]),
        "spx_pack_rhs_f32_device": (C.c_int, [vp, vp, i64, vp]),
        "spx_unpack_solutions_f64_device": (C.c_int, [vp, vp, i64, vp]),
        "spx_pipeline_config_auto": (None, [i64, i64, C.POINTER(SpxPipelineConfig)]),
        "spx_solve_batch_f64_pipelined": (C.c_int, [vp, vp, vp, i64, i64,
                                                    C.POINTER(SpxPipelineConfig), vp, u8p,
                                                    C.POINTER(C.c_int64),
                                                    C.POINTER(C.c_int64)]),
        "spx_spectral_config_default": (None, [C.POINTER(SpxSpectralConfig)]),
        "spx_spectrum_info_query": (C.c_int, [vp, C.POINTER(SpxWaveformBatch),
                                              C.POINTER(SpxSpectralConfig),
                                              C.POINTER(SpxSpectrumInfo)]),
        "spx_batch_fft_device": (C.c_int, [vp, C.POINTER(SpxWaveformBatch),
                                           C.POINTER(SpxSpectralConfig),
                                           C.POINTER(SpxSpectrumOut),
                                           C.POINTER(SpxSpectrumInfo)]),
        "spx_batch_thd_device": (C.c_int, [vp, C.POINTER(SpxWaveformBatch), C.c_double, i32,
                                           C.POINTER(SpxThdOut), C.POINTER(C.c_int32)]),
        "spx_thd_statistics_device": (C.c_int, [vp, vp, i64, C.c_double,
                                                C.POINTER(SpxThdStatistics)]),
        "spx_window_coefficients": (C.c_int, [i32, i64, vp]),
        "spx_window_coherent_gain": (C.c_double, [i32, i64]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a) -> C.c_void_p:
    """Raw data pointer of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return C.c_void_p(a.ctypes.data)
