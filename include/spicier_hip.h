/*
 * spicier_hip.h — C ABI of the MI355X batched-sweep engine (libspicier_hip.so).
 *
 * The engine replaces the GPU backend behind spicier-batched-sweep's
 * `BatchedLuSolver` trait and adds the nonlinear DC-operating-point sweep that
 * the trait's callers assemble on the CPU today.  Every entry point takes plain
 * pointers and sizes; no torch or HIP types cross this boundary.
 *
 * Conventions shared by every entry point:
 *   - return value: spx_status (0 = success); a message is kept per thread and
 *     read with spx_last_error().
 *   - matrices are column-major per system and batch-major across systems,
 *     exactly as `BatchedLuSolver::solve_batch` receives them
 *     (crates/spicier-batched-sweep/src/solver.rs:108-125).
 *   - "host" entry points take host pointers, copy in, run on the handle's
 *     stream and copy out before returning; "_device" entry points take device
 *     pointers (memory on the handle's device) and only enqueue work on the
 *     handle's stream unless stated otherwise.
 *   - calls on one handle are serialized by an internal mutex; a handle may be
 *     shared between threads (the trait is `Send + Sync`).
 */
#ifndef SPICIER_HIP_H
#define SPICIER_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Status codes: one per `BatchedSweepError` variant
 * (crates/spicier-batched-sweep/src/error.rs:7-46). */
typedef enum spx_status {
    SPX_OK = 0,
    SPX_ERR_INVALID_DIMENSION = 1, /* InvalidDimension */
    SPX_ERR_BATCH_TOO_LARGE = 2,   /* BatchTooLarge{size,max} */
    SPX_ERR_MEMORY_ALLOC = 3,      /* MemoryAlloc */
    SPX_ERR_TRANSFER = 4,          /* Transfer */
    SPX_ERR_BACKEND_INIT = 5,      /* BackendInit */
    SPX_ERR_BACKEND = 6,           /* Backend */
    SPX_ERR_NO_BACKEND = 7,        /* NoBackendAvailable */
    SPX_ERR_INVALID_ARGUMENT = 8   /* argument outside the documented domain */
} spx_status;

/* Per-instance convergence status: `ConvergenceStatus`
 * (crates/spicier-batched-sweep/src/convergence.rs:41-50). */
typedef enum spx_conv_status {
    SPX_ACTIVE = 0,
    SPX_CONVERGED = 1,
    SPX_FAILED = 2,
    SPX_SINGULAR = 3
} spx_conv_status;

/* Bit flags OR-ed into spx_sweep_result.flags[i]. */
enum {
    SPX_FLAG_KCL_FAIL = 1,      /* Δx test passed but the FP64 KCL residual test did not */
    SPX_FLAG_GMIN_STEPPING = 2, /* plain Newton failed; Gmin stepping was run */
    SPX_FLAG_SOURCE_STEPPING = 4,/* Gmin stepping failed; source stepping was run */
    SPX_FLAG_FP64_ESCALATED = 8,/* FP32+IR could not reach FP64 accuracy; FP64 LU used */
    SPX_FLAG_WARM_START = 16    /* Newton started from a caller or surrogate guess */
};

/* ------------------------------------------------------------------------- */
/* Handle: one device, one HIP stream, device memory owned by the library. */
typedef struct spx_handle spx_handle;

int spx_device_count(int32_t* count);
/* Creates a handle on `device_id`.  Cheap after the first call per process:
 * `parallel.rs:194` creates a solver per sweep point, so handles share a
 * process-wide device context. */
int spx_create(int32_t device_id, spx_handle** out);
int spx_destroy(spx_handle* h);
/* Message of the last failing call on this thread ("" if none). */
const char* spx_last_error(void);
/* Blocks until every call enqueued on the handle's stream has finished. */
int spx_synchronize(spx_handle* h);
/* The handle's stream as an opaque pointer (hipStream_t), for callers that
 * order their own device work against the library's. */
void* spx_stream(spx_handle* h);

/* ------------------------------------------------------------------------- */
/* Batched dense linear solve — the `BatchedLuSolver::solve_batch` drop-in
 * (replaces CudaBatchedLuSolver::solve_batch,
 *  crates/spicier-backend-cuda/src/batched_lu.rs:373-494, and matches
 *  CpuBatchedSolver, crates/spicier-batched-sweep/src/solver.rs:327-385).
 *
 * matrices: batch*n*n doubles, column-major per system; rhs: batch*n;
 * solutions: batch*n out.  Partial-pivot LU, first maximum on ties, a system
 * is singular iff an exact zero pivot column occurs or the solution is not
 * finite; singular systems get an all-zero solution and singular[i] = 1
 * (singular may be NULL).  *num_singular receives the count (may be NULL).
 * batch == 0 succeeds and touches nothing.  n must be 1..4096. */
int spx_solve_batch_f64(spx_handle* h, const double* matrices, const double* rhs,
                        int64_t n, int64_t batch, double* solutions,
                        uint8_t* singular, int64_t* num_singular);
/* Same contract on device pointers; enqueued on the handle's stream. */
int spx_solve_batch_f64_device(spx_handle* h, const double* d_matrices,
                               const double* d_rhs, int64_t n, int64_t batch,
                               double* d_solutions, uint8_t* d_singular);
/* No-allocation variant of spx_solve_batch_f64_device.  It has no
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same inputs, outputs, semantics and kernels; the
 * LU workspace of n > 64 is a 64 MiB array that the library holds on each
 * device from its load until the process ends, so the call makes no
 * allocation.  One system must fit it: n <= 2895 (n > 2895 is
 * SPX_ERR_INVALID_DIMENSION).  Every call on a device shares that device's
 * array, so calls on one device that overlap in time (handles on different
 * streams) race; calls on one handle are stream-ordered and do not. */
int spx_solve_batch_f64_device_noalloc(spx_handle* h, const double* d_matrices,
                                       const double* d_rhs, int64_t n, int64_t batch,
                                       double* d_solutions, uint8_t* d_singular);
/* Batched LU factorisation alone — the getrf half of the CUDA backend's
 * solve (cublasDgetrfBatched, crates/spicier-backend-cuda/src/batched_lu.rs:
 * 404-428), with the CPU path's pivot rule (nalgebra LU::new,
 * crates/spicier-batched-sweep/src/solver.rs:370): the largest |a| of the
 * column, the first row on exact ties, every bit compared.
 * matrices: batch*n*n column-major; factors out in LAPACK getrf's format
 * (column-major P A = L U, unit-lower L strictly below the diagonal, U on and
 * above); pivots: batch*n int32, pivots[i*n + k] = the 0-based row swapped
 * with row k at step k (>= k).  A system with a zero pivot column or a
 * non-finite factor is singular: singular[i] = 1 (may be NULL) and its
 * factors and pivots are zero; *num_singular receives the count (may be
 * NULL).  batch == 0 succeeds and touches nothing.  n must be 1..4096. */
int spx_lu_factor_batch_f64(spx_handle* h, const double* matrices, int64_t n, int64_t batch,
                            double* factors, int32_t* pivots, uint8_t* singular,
                            int64_t* num_singular);
/* Same contract on device pointers; enqueued on the handle's stream. */
int spx_lu_factor_batch_f64_device(spx_handle* h, const double* d_matrices, int64_t n,
                                   int64_t batch, double* d_factors, int32_t* d_pivots,
                                   uint8_t* d_singular);
/* Largest batch one launch of the solve accepts (GpuBatchConfig
 * max_batch_per_launch, solver.rs:6-41). Larger batches are split. */
int64_t spx_max_batch_per_launch(void);

/* Batched sparse solve with one pattern shared by every system — the GPU
 * counterpart of FaerTripletBatchedSolver::solve_batch_triplets and of the
 * numeric half of FaerSparseCachedBatchedSolver
 * (crates/spicier-batched-sweep/src/faer_sparse_solver.rs:196-225,230-362).
 * The pattern is nnz triplets (rows[t], cols[t]), 0-based, duplicates summed
 * in triplet order; system i has values[i*nnz + t] and rhs[i*n ..].  The
 * symbolic analysis (minimum-degree order, fill, elimination levels) runs
 * once per pattern and is cached on the handle, like the reference's
 * RwLock-cached SymbolicLu.  The numeric LU makes no row interchanges; an
 * unknown without a diagonal entry (an MNA voltage-source branch) is
 * eliminated right after its first neighbour.  A system is singular when a
 * pivot is exactly zero or the solution is not finite: its solution is all
 * zero and singular[i] = 1 (singular may be NULL).  n must be 1..4096. */
int spx_sparse_solve_batch_f64(spx_handle* h, int64_t n, int64_t nnz, const int32_t* rows,
                               const int32_t* cols, const double* values, const double* rhs,
                               int64_t batch, double* solutions, uint8_t* singular,
                               int64_t* num_singular);
/* Same on device pointers (values, rhs, solutions, singular); rows/cols are
 * host arrays.  Enqueued on the handle's stream. */
int spx_sparse_solve_batch_f64_device(spx_handle* h, int64_t n, int64_t nnz, const int32_t* rows,
                                      const int32_t* cols, const double* d_values,
                                      const double* d_rhs, int64_t batch, double* d_solutions,
                                      uint8_t* d_singular);

/* Batched restarted GMRES with right Jacobi preconditioning, one shared
 * triplet pattern (duplicates summed) — solve_gmres_real_preconditioned with
 * JacobiPreconditioner (crates/spicier-solver/src/gmres/real.rs:208-400,
 * crates/spicier-solver/src/preconditioner.rs:43-100; GmresConfig defaults
 * gmres/mod.rs:34-48), run for every system on the device in one launch in
 * place of the Metal sketch's host-driven loop
 * (crates/spicier-backend-metal/src/batched_gmres.rs:396-657).  Each system
 * stops when ||b - A x|| / ||b|| < tol or after max_iter Arnoldi steps (the
 * count that trips the budget is included, as in the reference).  x0 (may be
 * NULL = zero start) is [B][n]; iterations (int32), residuals (final relative
 * residual) and converged (uint8) may be NULL.  n must be 1..4096. */
typedef struct spx_gmres_config {
    int32_t restart;  /* Krylov dimension before restart (30) */
    int32_t max_iter; /* total Arnoldi steps (500) */
    double tol;       /* relative residual (1e-8) */
    int32_t jacobi;   /* 1: right Jacobi preconditioning (M = diag(A), |d| < 1e-30 -> 1) */
    int32_t precision; /* SPX_PREC_F64, or SPX_PREC_F32_IR: FP32 Krylov vectors, inner
                          products and rotations inside FP64 iterative refinement (the
                          residual and x in FP64, the same tol on the FP64 residual;
                          iterations count the FP32 Arnoldi steps) */
} spx_gmres_config;

void spx_gmres_config_default(spx_gmres_config* cfg);
int spx_gmres_solve_batch_f64(spx_handle* h, int64_t n, int64_t nnz, const int32_t* rows,
                              const int32_t* cols, const double* values, const double* rhs,
                              const double* x0, int64_t batch, const spx_gmres_config* cfg,
                              double* solutions, int32_t* iterations, double* residuals,
                              uint8_t* converged);
/* Same on device pointers (rows/cols stay host arrays); enqueued on the
 * handle's stream. */
int spx_gmres_solve_batch_f64_device(spx_handle* h, int64_t n, int64_t nnz, const int32_t* rows,
                                     const int32_t* cols, const double* d_values,
                                     const double* d_rhs, const double* d_x0, int64_t batch,
                                     const spx_gmres_config* cfg, double* d_solutions,
                                     int32_t* d_iterations, double* d_residuals,
                                     uint8_t* d_converged);

/* ------------------------------------------------------------------------- */
/* Circuit description: the device-resident "circuit program".  Element kinds
 * and stamp conventions follow spicier-core MnaSystem
 * (crates/spicier-core/src/mna.rs:87-169) and the device files cited below.
 * Nodes are numbered as spicier NodeId: 0 is ground, node k>0 is unknown k-1;
 * voltage sources, inductors, VCVS and CCVS get branch unknowns num_nodes +
 * (order of appearance). */
typedef enum spx_element_kind {
    SPX_EL_RESISTOR = 0,  /* node[0..1]; param[0] = resistance Ω (passive.rs:53-58) */
    SPX_EL_CAPACITOR = 1, /* node[0..1]; param[0] = C (F); open at DC, BE companion in transient */
    SPX_EL_VSOURCE = 2,   /* node[0]=pos,node[1]=neg; param[0] = V (sources.rs:96-101) */
    SPX_EL_ISOURCE = 3,   /* current param[0] flows node[0] -> node[1] (sources.rs:230-237) */
    SPX_EL_DIODE = 4,     /* node[0]=anode,node[1]=cathode; param[0]=IS, param[1]=N (diode.rs:91-131) */
    SPX_EL_NMOS = 5,      /* node[0..2]=d,g,s; param[0]=VTO,[1]=KP,[2]=LAMBDA,[3]=W,[4]=L (level1.rs:154-248) */
    SPX_EL_PMOS = 6,      /* as NMOS, PMOS sign convention */
    SPX_EL_VPULSE = 7,    /* voltage source whose transient value is PULSE(param[0]=V1, [1]=V2, [2]=TD,
                             [3]=TR, [4]=TF, [5]=PW, [6]=PER) (waveforms.rs:160-180); DC value V1 */
    SPX_EL_VSIN = 8,      /* voltage source whose transient value is SIN(param[0]=VO, [1]=VA,
                             [2]=FREQ, [3]=TD, [4]=THETA, [5]=PHASE in degrees): VO before TD,
                             then VO + VA e^(-THETA (t-TD)) sin(2 pi FREQ (t-TD) + PHASE)
                             (eval_sin, waveforms.rs:185-201); DC value VO */
    SPX_EL_INDUCTOR = 9,  /* node[0..1]; param[0] = L (H); a branch unknown like a voltage source
                             (passive.rs:283-375): 0 V source at DC; in a transient step its
                             row is i - G (v0 - v1) = I with the method's companion G, I
                             (InductorState, transient/companion.rs:158-302) */
    SPX_EL_VCVS = 10,     /* E: node[0..1] = out+/out-, node[2..3] = ctrl+/ctrl-; param[0] = gain;
                             a branch unknown, row v(out+) - v(out-) - gain (v(ctrl+) - v(ctrl-))
                             = 0 (controlled.rs:58-90) */
    SPX_EL_VCCS = 11,     /* G: nodes as VCVS; param[0] = gm; gm v(ctrl+, ctrl-) enters out+ and
                             leaves out- (controlled.rs:172-201) */
    SPX_EL_CCCS = 12,     /* F: node[0..1] = out+/out-; `ctrl` = the controlling element; param[0]
                             = gain; KCL +gain i(ctrl) at out+, -gain i(ctrl) at out-
                             (controlled.rs:267-285) */
    SPX_EL_CCVS = 13,     /* H: nodes and `ctrl` as CCCS; param[0] = gain; a branch unknown, row
                             v(out+) - v(out-) - gain i(ctrl) = 0 (controlled.rs:351-375) */
    SPX_EL_VPWL = 14      /* voltage source PWL(t1 v1 ... tk vk): `ctrl` = k points (1..4096),
                             param[0] = the slot of t1, the 2k values in consecutive slots;
                             eval_pwl (waveforms.rs:204-232); DC value v1 */
} spx_element_kind;

#define SPX_MAX_ELEMENT_PARAMS 7

typedef struct spx_element {
    int32_t kind;                           /* spx_element_kind */
    int32_t node[4];                        /* NodeId numbers, 0 = ground */
    int32_t param[SPX_MAX_ELEMENT_PARAMS];  /* parameter slot per quantity, -1 = unused */
    int32_t ctrl;                           /* CCCS / CCVS: index of the controlling element, an
                                               earlier voltage source, inductor, VCVS or CCVS
                                               (find_vsource_branch_index, netlist.rs:412-420);
                                               VPWL: its number of points; -1 otherwise */
} spx_element;

typedef struct spx_circuit {
    int32_t num_nodes;            /* nodes excluding ground */
    int32_t num_elements;
    const spx_element* elements;  /* stamped in this order (netlist.rs:441-445) */
    int32_t num_params;           /* P: per-instance parameter slots */
} spx_circuit;

/* Distribution of one parameter slot. */
typedef enum spx_param_kind {
    SPX_P_CONST = 0,      /* a */
    SPX_P_GAUSS_REL = 1,  /* (a + z*b*a).clamp(c, d): ParameterVariation{nominal=a, sigma=b, min=c, max=d}
                             (spicier-solver/src/sweep.rs:17-54,193) */
    SPX_P_GAUSS = 2,      /* a + z*b (rng.rs:148-150 gaussian_scaled) */
    SPX_P_UNIFORM = 3,    /* a + (b-a)*u */
    SPX_P_LOGUNIFORM = 4  /* exp(ln a + (ln b - ln a)*u) */
} spx_param_kind;

typedef struct spx_param_spec {
    int32_t kind;       /* spx_param_kind */
    int32_t rng_index;  /* hash-RNG param_idx of this slot (rng.rs:51-56) */
    double a, b, c, d;
} spx_param_spec;

/* How instance parameters are produced. */
typedef enum spx_generator {
    SPX_GEN_HASH = 0,     /* stateless hash RNG, (seed, sweep_offset+i, rng_index) (rng.rs:39-150) */
    SPX_GEN_LCG = 1,      /* MonteCarloGenerator's sequential LCG + Box-Muller, reproduced with
                             jump-ahead; every varied slot must be SPX_P_GAUSS_REL
                             (spicier-solver/src/sweep.rs:170-200) */
    SPX_GEN_CORNER = 2,   /* CornerGenerator: bit j of i -> varied slot j at c (min) or d (max)
                             (sweep.rs:204-227) */
    SPX_GEN_LINEAR = 3,   /* LinearSweepGenerator: c + t*(d-c), t=i/(B-1) (sweep.rs:231-266) */
    SPX_GEN_EXPLICIT = 4  /* caller-supplied parameters [batch][P] */
} spx_generator;

typedef struct spx_sweep_spec {
    int32_t generator;             /* spx_generator */
    uint64_t seed;
    int64_t batch;                 /* B instances */
    int64_t sweep_offset;          /* instance i uses sweep index sweep_offset + i */
    const spx_param_spec* params;  /* num_params entries */
    const double* explicit_params; /* [batch][P] for SPX_GEN_EXPLICIT (host or device pointer
                                      matching the entry point), else NULL */
} spx_sweep_spec;

/* Fills params[B][P] (sweep-major, as generate_gaussian_parameters,
 * rng.rs:208-227).  Host and device variants. */
int spx_sample_params(spx_handle* h, const spx_sweep_spec* spec, int32_t num_params,
                      double* params);
int spx_sample_params_device(spx_handle* h, const spx_sweep_spec* spec, int32_t num_params,
                             double* d_params);

/* ------------------------------------------------------------------------- */
/* Newton configuration: ConvergenceCriteria, GminSteppingParams and
 * SourceSteppingParams (crates/spicier-solver/src/newton.rs:11-34,174-231)
 * plus the engine's precision / limiter / warm-start choices. */
typedef enum spx_precision {
    SPX_PREC_F64 = 0,    /* FP64 LU every Newton iteration */
    SPX_PREC_F32_IR = 1  /* FP32 LU + FP64 iterative refinement; n <= 64 reuses its FP32 factors
                            across iterations, n > 64 (sparse LU, FP32 L+U values in LDS)
                            refactors each iteration; instances FP32 cannot resolve rerun in
                            FP64, flagged SPX_FLAG_FP64_ESCALATED */
} spx_precision;

typedef enum spx_limiter {
    SPX_LIMIT_REFERENCE = 0, /* diode.rs:139-148 memoryless log limiter, as the reference stamps
                                (its Newton fixed point above vcrit is not the Shockley solution) */
    SPX_LIMIT_NONE = 1,      /* plain Shockley, no limiting */
    SPX_LIMIT_PNJLIM = 2     /* SPICE pnjlim on each diode's junction voltage, vcrit from the
                                device's own IS; an iteration that limited cannot converge */
} spx_limiter;

typedef enum spx_warm_start {
    SPX_WARM_ZERO = 0,      /* zero initial guess (newton.rs:80-83) */
    SPX_WARM_GIVEN = 1,     /* x0[B][n] supplied by the caller */
    SPX_WARM_SURROGATE = 2  /* predicted by the surrogate attached with spx_set_surrogate */
} spx_warm_start;

typedef struct spx_nr_config {
    double v_abstol;            /* 1e-6 */
    double v_reltol;            /* 1e-3 */
    double i_abstol;            /* 1e-12 */
    int32_t max_iterations;     /* 50 */
    int32_t gmin_stepping;      /* 1: fall back to Gmin stepping if plain NR fails */
    double gmin_initial;        /* 1e-3 */
    double gmin_final;          /* 1e-12 */
    double gmin_factor;         /* 10 */
    int32_t gmin_max_attempts;  /* 5 */
    int32_t source_stepping;    /* 1: fall back to source stepping if Gmin stepping fails */
    double src_initial;         /* 0.1 */
    double src_step;            /* 0.1 */
    double src_min_step;        /* 0.01 */
    int32_t src_max_attempts;   /* 5 */
    int32_t precision;          /* spx_precision */
    int32_t limiter;            /* spx_limiter */
    int32_t warm_start;         /* spx_warm_start */
    int32_t ir_max_steps;       /* FP32+IR: refinement steps per Newton solve before refactoring */
    double ir_tol;              /* FP32+IR: stop when ||b - J x||inf <= ir_tol ||J||inf ||x||inf
                                   (LAPACK dsgesv's test); 0 = sqrt(n) * 2^-53 */
    /* KCL residual test of the final x, FP64: devices evaluated unlimited
     * (Shockley / square law) at x; node row i passes if
     * |sum_j J_ij x_j - b_i| <= kcl_reltol * max(|J_ij x_j|, |b_i|) + kcl_abstol. */
    double kcl_reltol;          /* SPICE reltol of the KCL residual test (1e-3) */
    double kcl_abstol;          /* abstol (A) of node rows (1e-12) */
    double kcl_vntol;           /* vntol (V) of branch rows (1e-6) */
} spx_nr_config;

/* Fills *cfg with the defaults quoted above (FP64, reference limiter, zero start). */
void spx_nr_config_default(spx_nr_config* cfg);

/* Caller-owned result buffers (host for spx_sweep_dcop, device for
 * spx_sweep_dcop_device).  Any pointer except x may be NULL. */
typedef struct spx_sweep_result {
    double* x;            /* [B][n] converged solution (node voltages, then branch currents) */
    int32_t* status;      /* [B] spx_conv_status */
    int32_t* iterations;  /* [B] Newton iterations performed (all stages) */
    int32_t* flags;       /* [B] SPX_FLAG_* */
    double* residual;     /* [B] max_i |KCL_i| / tol_i of the final solution (<= 1 passes) */
    double* params;       /* [B][P] sampled parameters */
    double* node_stats;   /* [n][5] count,sum,sumsq,min,max over converged instances
                             (StatisticsAccumulator, statistics.rs:135-154) */
    int32_t* factorizations; /* [B] LU factorizations performed (FP64 + FP32) */
    int32_t* refinements;    /* [B] FP64-residual refinement steps performed (FP32+IR) */
} spx_sweep_result;

/* Monte Carlo / corner / linear DC-operating-point sweep (the nonlinear
 * counterpart of solve_batched_sweep_gpu, crates/spicier-batched-sweep/src/sweep.rs:101-217,
 * with the per-instance semantics of solve_newton_raphson + Gmin + source
 * stepping, crates/spicier-solver/src/newton.rs:70-459).  x0 is [B][n] when
 * cfg->warm_start == SPX_WARM_GIVEN, else NULL. */
int spx_sweep_dcop(spx_handle* h, const spx_circuit* ckt, const spx_sweep_spec* spec,
                   const spx_nr_config* cfg, const double* x0, spx_sweep_result* out);
/* Device-pointer variant: spec->explicit_params, x0 and every result buffer
 * are device pointers; returns once the work is enqueued. */
int spx_sweep_dcop_device(spx_handle* h, const spx_circuit* ckt, const spx_sweep_spec* spec,
                          const spx_nr_config* cfg, const double* d_x0, spx_sweep_result* d_out);

/* ------------------------------------------------------------------------- */
/* Transient sweep: Backward Euler by default (CapacitorState::stamp_be / update,
 * crates/spicier-solver/src/transient/companion.rs:41-66; solve_transient,
 * crates/spicier-solver/src/transient/solver.rs:41-160; PULSE sources by
 * eval_pulse, crates/spicier-devices/src/waveforms.rs:160-180).  Every
 * instance: DC operating point with capacitors open and PULSE sources at V1
 * (the spx_sweep_dcop semantics), then `steps` fixed steps of `tstep` at
 * t = k*tstep, each capacitor replaced by G = C/h with I = G*v_prev from its
 * negative to its positive node.  The reference's transient is linear (one
 * solve per step); circuits with diodes or MOSFETs run plain Newton at every
 * step from the previous step's solution (no step cutting).  FP64. */
/* Integration methods (IntegrationMethod, transient/types.rs:50-63) and
 * their capacitor companions (companion.rs:41-130):
 *   SPX_TRAN_BE      G = C/h, I = G*v_prev;
 *   SPX_TRAN_TRAP    G = 2C/h, I = G*v_prev + i_prev;
 *   SPX_TRAN_TRBDF2  a trapezoidal stage of gamma*h (gamma = 2 - sqrt 2) to an
 *                    intermediate point, then a variable-step BDF2 stage of
 *                    (1 - gamma)*h; both stages evaluate sources at the end
 *                    of the step, and i_prev follows the reference's update
 *                    rules (companion.rs:62-95), as solve_transient
 *                    (solver.rs:184-248) does. */
enum { SPX_TRAN_BE = 0, SPX_TRAN_TRAP = 1, SPX_TRAN_TRBDF2 = 2 };

/* Adaptive trapezoidal stepping (AdaptiveTransientParams, types.rs:76-117;
 * solve_transient_adaptive, solver.rs:461-676): each instance keeps its own
 * time grid.  A step of h from t solves the trapezoidal companion with the
 * sources at t (as the reference stamps them), estimates each capacitor's
 * truncation error |i_trap - i_be| / 3 (estimate_lte, companion.rs:132-150),
 * and is rejected when max LTE > max(abstol, reltol * max|v_cap|) while
 * h > h_min (h *= clamp(sqrt(tol/lte), 0.1, 0.5)); an accepted step may grow
 * h by up to 1.5x when max LTE < tol/2.  h is clamped to [h_min, h_max] and
 * cut to end exactly at tstop.  An instance needing more than max_points
 * time points (t = 0 included) stops with SPX_FAILED. */
typedef struct spx_tran_adaptive {
    double tstop;     /* > 0 */
    double h_init;    /* > 0 */
    double h_min;     /* > 0 */
    double h_max;     /* >= h_min */
    double reltol;    /* >= 0 */
    double abstol;    /* >= 0 */
    int32_t max_points;
} spx_tran_adaptive;

typedef struct spx_tran_config {
    double tstep;            /* step h (s), > 0 (fixed steps) */
    int32_t steps;           /* number of steps after t = 0 (fixed steps) */
    int32_t num_probes;      /* unknowns recorded at every step */
    const int32_t* probes;   /* [num_probes] unknown indices (host array) */
    int32_t method;          /* SPX_TRAN_*; ignored (trapezoidal) when adaptive */
    const spx_tran_adaptive* adaptive;  /* NULL: fixed steps */
} spx_tran_config;

typedef struct spx_tran_result {
    double* x;           /* [B][n] solution at the last step reached */
    int32_t* status;     /* [B] spx_conv_status of the whole run */
    int32_t* iterations; /* [B] Newton iterations, operating point and steps */
    int32_t* flags;      /* [B] SPX_FLAG_* (KCL test of the last step) */
    double* residual;    /* [B] KCL residual ratio of the last step */
    double* params;      /* [B][P] sampled parameters (may be NULL) */
    double* waveforms;   /* [B][points][num_probes] probed unknowns at every time point:
                            points = steps + 1, or max_points when adaptive */
    int32_t* fail_step;  /* [B] first step that failed (0 = operating point), -1 if none */
    /* adaptive only (each may be NULL): */
    double* times;           /* [B][max_points] time of every point */
    int32_t* num_points;     /* [B] points recorded, t = 0 included */
    int32_t* total_steps;    /* [B] steps tried (accepted + rejected) */
    int32_t* rejected_steps; /* [B] */
    double* min_step;        /* [B] smallest accepted step (inf if none) */
    double* max_step;        /* [B] largest accepted step (0 if none) */
} spx_tran_result;

int spx_transient_sweep(spx_handle* h, const spx_circuit* ckt, const spx_sweep_spec* spec,
                        const spx_nr_config* cfg, const spx_tran_config* tran,
                        spx_tran_result* out);
/* Device-pointer variant (result buffers and explicit parameters on the
 * device); returns once the work is enqueued. */
int spx_transient_sweep_device(spx_handle* h, const spx_circuit* ckt, const spx_sweep_spec* spec,
                               const spx_nr_config* cfg, const spx_tran_config* tran,
                               spx_tran_result* d_out);

/* Mixed-size transient batch: several circuits (size classes), each with its
 * own sweep spec, transient settings and device result buffers, run as one
 * launch.  Each class runs on a persistent kernel shaped for it: the
 * register-LU kernel (n <= 64), the shared-symbolic sparse kernel with 64
 * or 256 threads per instance, or the same plan with one lane per instance
 * (chain-like elimination trees).  With SPX_MIXED_GRAPH every class's sampling
 * and transient kernel is captured into one hipGraph on the first call for a
 * batch description and replayed by later calls with the same description
 * (buffers, sizes, seeds, settings).  With SPX_MIXED_AUTOTUNE a class whose
 * kernel is SPX_KERNEL_AUTO is timed once on a sample (up to 1,024
 * instances, 20 steps) with every candidate kernel (register LU when n <= 64,
 * sparse 64, sparse 256, lanes), and the fastest is kept for that circuit on
 * this handle.  `kernels` (may be NULL): per class,
 * in = the requested kind, out = the kind that ran.  No counterpart in the
 * reference, whose transient runs one circuit at a time on the host. */
typedef struct spx_tran_class {
    const spx_circuit* circuit;
    spx_sweep_spec spec;     /* explicit_params, if used, on the device */
    spx_tran_config tran;
    spx_tran_result result;  /* device buffers */
} spx_tran_class;

enum { SPX_MIXED_GRAPH = 1, SPX_MIXED_AUTOTUNE = 2 };
enum {
    SPX_KERNEL_AUTO = 0,      /* register LU if n <= 64, else sparse 256 (or the tuned choice) */
    SPX_KERNEL_DENSE = 1,     /* register LU, one wave per instance (n <= 64) */
    SPX_KERNEL_SPARSE64 = 2,  /* shared-symbolic LU, 64 threads per instance */
    SPX_KERNEL_SPARSE256 = 3, /* shared-symbolic LU, 256 threads per instance */
    SPX_KERNEL_LANES = 4      /* shared-symbolic LU, one lane per instance, 64 per wave */
};

int spx_transient_mixed_device(spx_handle* h, const spx_tran_class* classes, int32_t num_classes,
                               const spx_nr_config* cfg, uint32_t options, int32_t* kernels);

/* Unknown count n of a circuit (num_nodes + number of voltage sources). */
int32_t spx_circuit_size(const spx_circuit* ckt);

/* ------------------------------------------------------------------------- */
/* Learned warm start: an MLP  n_in -> n_hidden -> n_hidden -> n_out  with SiLU
 * hidden activations.  Inputs are chosen parameter slots of the [B][P]
 * parameter rows, outputs are the first n_out unknowns; both are
 * standardised with statistics of the training set.  Matmuls run on BF16
 * MFMA with FP32 accumulation (weights and activations rounded to BF16 as
 * they enter the matrix cores); FP32 master weights are updated by Adam.
 * Training uses only the library's own kernels.  No counterpart in the
 * reference (its sweeps always start Newton from zero, newton.rs:80-83). */
typedef struct spx_surrogate spx_surrogate;

int spx_surrogate_create(spx_handle* h, int32_t n_in, int32_t n_hidden, int32_t n_out,
                         uint64_t seed, spx_surrogate** out);
int spx_surrogate_destroy(spx_surrogate* s);
/* Parameter slots feeding the inputs (n_in entries, each in [0, P)). */
int spx_surrogate_set_inputs(spx_surrogate* s, const int32_t* slots, int32_t P);
/* Number of trainable FP32 values (W1,b1,W2,b2,W3,b3 in that order). */
int64_t spx_surrogate_num_weights(const spx_surrogate* s);
int spx_surrogate_get_weights(spx_surrogate* s, float* host, int64_t count);
int spx_surrogate_set_weights(spx_surrogate* s, const float* host, int64_t count);
/* Normalisation (host arrays): in_mean/in_std [n_in], out_mean/out_std [n_out]. */
int spx_surrogate_set_normalization(spx_surrogate* s, const float* in_mean, const float* in_std,
                                    const float* out_mean, const float* out_std);
/* Fits the normalisation to the rows whose mask byte is non-zero (mask may be
 * NULL), then trains with Adam for `epochs` passes in minibatches of `batch`
 * rows shuffled by `shuffle_seed`.  d_params [B][P], d_x [B][n_x] (n_x >= n_out),
 * device pointers.  With tol_rel or tol_abs > 0, output j's squared error is
 * weighted by (std_j / (tol_rel |mean_j| + tol_abs))^2 (normalised to mean 1),
 * i.e. measured in units of the Newton tolerance the warm start has to meet.
 * loss_history (host, may be NULL) receives each epoch's mean loss.
 * The 32 -> 256 -> 256 -> 64 network trains in three fused launches per step
 * (row pass, weight-gradient tiles, reduce + Adam); other shapes take the
 * layered GEMM path.  Synchronous. */
int spx_surrogate_train_device(spx_surrogate* s, const double* d_params, const double* d_x,
                               int32_t n_x, const uint8_t* d_mask, int64_t B, int32_t epochs,
                               int32_t batch, float lr, float tol_rel, float tol_abs,
                               uint64_t shuffle_seed, float* loss_history);
/* x0[B][n_x] = prediction for the first n_out unknowns, zero for the rest.
 * Enqueued on the handle's stream. */
int spx_surrogate_predict_device(spx_surrogate* s, const double* d_params, int64_t B,
                                 int32_t n_x, double* d_x0);
/* Loss and gradient of the standardised MSE for one minibatch given already
 * standardised FP32 inputs [B][n_in] and targets [B][n_out] (device
 * pointers); d_grad receives num_weights values.  For gradient checks. */
int spx_surrogate_gradient_device(spx_surrogate* s, const float* d_in, const float* d_target,
                                  int64_t B, float* d_grad, float* loss);
/* Attach (or detach with NULL) the surrogate used by SPX_WARM_SURROGATE. */
int spx_set_surrogate(spx_handle* h, spx_surrogate* s);

/* ------------------------------------------------------------------------- */
/* Statistics reduction over a [B][n] buffer (device pointers), skipping
 * instances whose mask byte is 0 (mask may be NULL): out[n][5] =
 * count,sum,sumsq,min,max (StatisticsAccumulator::add/merge,
 * crates/spicier-batched-sweep/src/statistics.rs:135-154). */
int spx_stats_reduce_device(spx_handle* h, const double* d_values, const uint8_t* d_mask,
                            int64_t batch, int32_t n, double* d_out);

/* ------------------------------------------------------------------------- */
/* Convergence tracking over device arrays (ConvergenceTracker,
 * crates/spicier-batched-sweep/src/convergence.rs:76-387): status[B] holds
 * spx_conv_status codes, iterations[B] the per-instance counts.  Counts of
 * newly converged instances are written to device memory (may be NULL), so no
 * call below waits for the device except the summary. */

/* check_convergence (convergence.rs:274-314): an ACTIVE instance converges when
 * |change_j| <= abstol + reltol * |solution_j| for every j (a NaN change
 * passes, as in the reference's `delta > tol` test).  Buffers are [B][n]. */
int spx_tracker_check_convergence_device(spx_handle* h, int32_t* d_status, int64_t batch,
                                         int32_t n, const double* d_changes,
                                         const double* d_solutions, double abstol, double reltol,
                                         int64_t* d_newly_converged);
/* check_residual_convergence (convergence.rs:327-357): ACTIVE instances whose
 * residual 2-norm (summed in index order) is below tolerance converge. */
int spx_tracker_check_residual_device(spx_handle* h, int32_t* d_status, int64_t batch, int32_t n,
                                      const double* d_residuals, double tolerance,
                                      int64_t* d_newly_converged);
/* increment_all_active (convergence.rs:197-205) with indices NULL: ACTIVE
 * instances count one more iteration and fail on reaching max_iterations.
 * With host indices: increment_iteration (convergence.rs:184-190) on each
 * listed instance, whatever its status. */
int spx_tracker_increment_device(spx_handle* h, int32_t* d_status, int32_t* d_iterations,
                                 int64_t batch, int32_t max_iterations, const int64_t* indices,
                                 int64_t count);
/* mark_converged / mark_failed / mark_singular (convergence.rs:150-182) on the
 * listed instances (host indices); only ACTIVE ones change.  *marked (host,
 * may be NULL) = how many changed.  Waits for the handle's stream. */
int spx_tracker_mark_device(spx_handle* h, int32_t* d_status, int64_t batch,
                            const int64_t* indices, int64_t count, int32_t new_status,
                            int64_t* marked);

/* ConvergenceSummary (convergence.rs:360-400); average and maximum over the
 * converged instances. */
typedef struct spx_tracker_summary {
    int64_t total_points;
    int64_t converged_count;
    int64_t failed_count;      /* FAILED + SINGULAR */
    int64_t active_count;
    int64_t iteration_sum;     /* over converged instances */
    int32_t max_iterations;    /* over converged instances */
    double average_iterations;
} spx_tracker_summary;
/* Host result; waits for the handle's stream.  d_iterations may be NULL. */
int spx_tracker_summary_device(spx_handle* h, const int32_t* d_status, const int32_t* d_iterations,
                               int64_t batch, spx_tracker_summary* out);

/* active_indices / converged_indices / failed_indices (convergence.rs:209-238):
 * ascending indices of the selected instances into d_indices (room for
 * `batch`), their number into *d_count (device). */
enum { SPX_SELECT_ACTIVE = 0, SPX_SELECT_CONVERGED = 1, SPX_SELECT_FAILED = 2 };
int spx_select_indices_device(spx_handle* h, const int32_t* d_status, int64_t batch,
                              int32_t which, int64_t* d_indices, int64_t* d_count);
/* compact_active (convergence.rs:432-441): out[k] = data[indices[k]], rows of
 * item_size doubles. */
int spx_compact_rows_device(spx_handle* h, const double* d_data, int64_t item_size,
                            const int64_t* d_indices, int64_t count, double* d_out);
/* expand_active (convergence.rs:453-471): out = fill everywhere, then
 * out[indices[k]] = compacted[k].  Indices must be distinct. */
int spx_expand_rows_device(spx_handle* h, const double* d_compacted, int64_t item_size,
                           const int64_t* d_indices, int64_t count, int64_t batch,
                           double fill_value, double* d_out);

/* ------------------------------------------------------------------------- */
/* Histogram of one node over a [B][n] solution buffer (Histogram::new and
 * Histogram::from_sweep, crates/spicier-batched-sweep/src/statistics.rs:244-293).
 * Bin = floor((v - min) / width), NaN to bin 0, clamped to the bins.  With
 * from_sweep the range is the column's min/max (NaN skipped), widened by 0.5
 * each way if they are within f64::EPSILON; otherwise max must exceed min.
 * d_counts: num_bins int64 (device); d_range (device, may be NULL): the
 * range used. */
int spx_histogram_device(spx_handle* h, const double* d_solutions, int64_t batch, int32_t n,
                         int32_t node, int32_t num_bins, double min, double max,
                         int32_t from_sweep, int64_t* d_counts, double* d_range);

/* YieldSpec (statistics.rs:331-376): lower <= v <= upper passes. */
typedef struct spx_yield_spec {
    int32_t node;
    double lower_limit;
    double upper_limit;
} spx_yield_spec;
/* YieldAnalysis::compute_yield / individual_yields (statistics.rs:379-425):
 * d_counts[0] = instances passing every spec, d_counts[1 + s] = instances
 * passing spec s (device, num_specs + 1 int64). */
int spx_yield_device(spx_handle* h, const double* d_solutions, int64_t batch, int32_t n,
                     const spx_yield_spec* specs, int32_t num_specs, int64_t* d_counts);

/* StreamingStatistics::process_chunk (statistics.rs:470-587): reduces a chunk
 * [B][n] (mask as spx_stats_reduce_device) and merges it into the running
 * accumulators d_acc[n][5] (count,sum,sumsq,min,max; start from
 * 0,0,0,+inf,-inf). */
int spx_stats_accumulate_device(spx_handle* h, const double* d_values, const uint8_t* d_mask,
                                int64_t batch, int32_t n, double* d_acc);

/* StreamingStatistics::merge (statistics.rs:545-556): d_acc[n][5] merged
 * with d_other[n][5] (StatisticsAccumulator::merge, :148-154). */
int spx_stats_merge_device(spx_handle* h, double* d_acc, const double* d_other, int32_t n);

/* ------------------------------------------------------------------------- */
/* BatchLayout packing (crates/spicier-batched-sweep/src/batch_layout.rs:43-199):
 * column-major FP64 systems to row-major FP32 rows of row_stride floats (0:
 * n rounded up to a multiple of 32, BatchLayout::new), padding zero. */
int spx_pack_matrices_f32_device(spx_handle* h, const double* d_matrices, int64_t n,
                                 int64_t batch, int64_t row_stride, float* d_packed);
/* pack_rhs_f32 / unpack_solutions_f64 (batch_layout.rs:190-199). */
int spx_pack_rhs_f32_device(spx_handle* h, const double* d_rhs, int64_t count, float* d_packed);
int spx_unpack_solutions_f64_device(spx_handle* h, const float* d_packed, int64_t count,
                                    double* d_solutions);

/* ------------------------------------------------------------------------- */
/* Pipelined batched solve from host buffers (PipelinedSweep,
 * crates/spicier-batched-sweep/src/pipeline.rs:130-361; PipelineConfig
 * :7-87).  Same inputs, outputs and semantics as spx_solve_batch_f64; the
 * batch runs in chunks over num_buffers pinned staging slots, each with its
 * own stream, so the host copies, both transfer directions and the solves of
 * different chunks overlap. */
typedef struct spx_pipeline_config {
    int64_t chunk_size;
    int32_t num_buffers;
} spx_pipeline_config;
/* PipelineConfig::auto (pipeline.rs:68-87). */
void spx_pipeline_config_auto(int64_t n, int64_t total_points, spx_pipeline_config* cfg);
/* cfg NULL: spx_pipeline_config_auto. */
int spx_solve_batch_f64_pipelined(spx_handle* h, const double* matrices, const double* rhs,
                                  int64_t n, int64_t batch, const spx_pipeline_config* cfg,
                                  double* solutions, uint8_t* singular, int64_t* num_singular,
                                  int64_t* chunks_processed);

/* ------------------------------------------------------------------------- */
/* Batch spectral analysis of sweep waveforms: windowed FFT spectra
 * (compute_fft / compute_fft_from_samples,
 * crates/spicier-solver/src/spectral/fft.rs:98-367), THD with harmonic peak
 * search (compute_thd / compute_thd_from_samples, spectral/thd.rs:60-312),
 * and their batch forms over a sweep (compute_batch_thd,
 * compute_batch_thd_with_stats, compute_thd_statistics,
 * compute_batch_fft_stats, crates/spicier-batched-sweep/src/spectral.rs:46-330).
 * One workgroup runs one waveform: resampling, window, DFT in LDS (a
 * power-of-two Stockham transform, two real samples per complex point; other
 * lengths by chirp-z), spectrum, harmonic search and the per-bin ensemble
 * statistics are one kernel, for lengths up to 16384 (powers of two) or 4096
 * (others).  Longer transforms, up to 2^22 points (2^21 for lengths that are
 * not powers of two), run as four-step transforms through HBM in chunks of
 * the batch. */
enum {
    SPX_WINDOW_RECTANGULAR = 0,
    SPX_WINDOW_HANNING = 1, /* the reference's default */
    SPX_WINDOW_HAMMING = 2,
    SPX_WINDOW_BLACKMAN = 3
};

/* SpectralConfig (fft.rs:10-29); fft_size 0 = None (next power of two). */
typedef struct spx_spectral_config {
    int32_t window;
    int32_t zero_padding;
    int64_t fft_size;
} spx_spectral_config;
void spx_spectral_config_default(spx_spectral_config* cfg);

/* One node's waveform for every instance of a sweep.  Point t of waveform b
 * is data[b*batch_stride + t*point_stride] (device memory), so the probe
 * column of spx_tran_result.waveforms is read in place.  times (host,
 * num_points, strictly increasing and finite): the waveform is resampled to
 * uniform steps first, as compute_fft does (resample_uniform with its
 * automatic rate, fft.rs:98-147, interpolate_at, transient/result.rs:41-68).
 * times NULL: the points are already uniform at sample_rate
 * (compute_fft_from_samples). */
typedef struct spx_waveform_batch {
    const double* data;
    int64_t batch;
    int64_t batch_stride;
    int32_t num_points;
    int32_t point_stride;
    const double* times;
    double sample_rate;
} spx_waveform_batch;

/* Sizes of the spectrum a batch produces.  fft_size 0: the reference's empty
 * SpectralResult (no samples, or sample_rate <= 0). */
typedef struct spx_spectrum_info {
    int64_t num_samples;     /* samples after resampling */
    int64_t fft_size;        /* DFT length */
    int64_t num_bins;        /* fft_size/2 + 1 (DC to Nyquist) */
    double sample_rate;
    double freq_resolution;  /* frequencies[i] = i * freq_resolution */
} spx_spectrum_info;
int spx_spectrum_info_query(spx_handle* h, const spx_waveform_batch* wf,
                            const spx_spectral_config* cfg, spx_spectrum_info* info);

/* Device outputs of spx_batch_fft_device; any pointer may be NULL.
 * Per waveform [batch][num_bins]: SpectralResult's magnitude, magnitude_db,
 * phase and power_spectral_density.  Per bin [num_bins]:
 * BatchSpectralResult's mean/std/min/max magnitude over the batch
 * (population standard deviation, spectral.rs:262-330). */
typedef struct spx_spectrum_out {
    double* magnitude;
    double* magnitude_db;
    double* phase;
    double* psd;
    double* mean_magnitude;
    double* std_magnitude;
    double* min_magnitude;
    double* max_magnitude;
} spx_spectrum_out;
/* info (may be NULL) receives the sizes.  Returns once the work is enqueued. */
int spx_batch_fft_device(spx_handle* h, const spx_waveform_batch* wf,
                         const spx_spectral_config* cfg, const spx_spectrum_out* out,
                         spx_spectrum_info* info);

/* Device outputs of spx_batch_thd_device; any pointer may be NULL.  Per
 * waveform [batch]: ThdResult's thd_percent, thd_db, fundamental_freq (the
 * bin found), fundamental_magnitude and harmonic_rms.  harmonics
 * [batch][num_harmonics+1][5]: per HarmonicInfo (harmonic h+1) frequency,
 * magnitude, magnitude_db, phase and relative_percent; rows past
 * found_harmonics are left untouched. */
typedef struct spx_thd_out {
    double* thd_percent;
    double* thd_db;
    double* fundamental_freq;
    double* fundamental_magnitude;
    double* harmonic_rms;
    double* harmonics;
} spx_thd_out;
/* compute_batch_thd (spectral.rs:46-71): Hanning window, zero padding to the
 * next power of two.  found_harmonics (host, may be NULL) receives
 * ThdResult::num_harmonics, the harmonics at or below Nyquist (the same for
 * every waveform of the batch). */
int spx_batch_thd_device(spx_handle* h, const spx_waveform_batch* wf, double fundamental_freq,
                         int32_t num_harmonics, const spx_thd_out* out, int32_t* found_harmonics);

/* ThdStatistics (spectral.rs:12-31). */
typedef struct spx_thd_statistics {
    double mean_thd;
    double std_thd;
    double min_thd;
    double max_thd;
    double median_thd;
    double yield_percent;
    int64_t num_samples;
    double thd_spec;
} spx_thd_statistics;
/* compute_thd_statistics (spectral.rs:103-160) over count device values:
 * population std, median of the sorted values, yield = share <= thd_spec.
 * A NaN value is an SPX_ERR_INVALID_ARGUMENT error (the reference's sort
 * panics on it).  Blocks until the statistics are on the host. */
int spx_thd_statistics_device(spx_handle* h, const double* d_thd_percent, int64_t count,
                              double thd_spec, spx_thd_statistics* stats);

/* WindowFunction::apply's weights and coherent_gain (window.rs:24-108);
 * host arithmetic, the same expressions the spectra use. */
int spx_window_coefficients(int32_t window, int64_t n, double* out);
double spx_window_coherent_gain(int32_t window, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* SPICIER_HIP_H */
