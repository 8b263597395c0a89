/*
 * oracle.h — CPU FP64 restatement of spicier's sweep / DC-operating-point path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only the tests, the smoke check and the benchmark's CPU-baseline leg load it.
 * The engine (libspicier_hip.so) never links or calls it.
 *
 * The reference (rjwalters/spicier) is Rust on faer 0.24.0 / nalgebra 0.33.2 /
 * rayon 1.11.0 and cannot be compiled here (no cargo/rustc, no network), so each
 * function below restates the cited reference code in C++ with the same
 * operation order; it is built with -ffp-contract=off and uses glibc libm, the
 * libm Rust's f64::{ln,exp,cos,ln_1p} call on Linux.
 */
#ifndef SPICIER_ORACLE_H
#define SPICIER_ORACLE_H

#include <stdint.h>
#include "../include/spicier_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rng.rs:39-150 */
uint64_t orc_splitmix64(uint64_t x);
uint64_t orc_combine_indices(uint64_t seed, uint32_t sweep_idx, uint32_t param_idx);
double orc_uniform(uint64_t seed, uint32_t sweep_idx, uint32_t param_idx);
float orc_uniform_f32(uint64_t seed, uint32_t sweep_idx, uint32_t param_idx);
double orc_gaussian(uint64_t seed, uint32_t sweep_idx, uint32_t param_idx);
float orc_gaussian_f32(uint64_t seed, uint32_t sweep_idx, uint32_t param_idx);
double orc_gaussian_scaled(uint64_t seed, uint32_t sweep_idx, uint32_t param_idx,
                           double mean, double sigma);
/* rng.rs:208-227 */
void orc_generate_gaussian_parameters(uint64_t seed, int64_t num_sweeps, int32_t num_params,
                                      const double* means, const double* sigmas, double* out);
/* spicier-solver/src/sweep.rs:170-200 (sequential LCG), 204-227, 231-266 */
void orc_monte_carlo_lcg(uint64_t seed, int64_t num_samples, int32_t num_vars,
                         const double* nominal, const double* sigma, const double* vmin,
                         const double* vmax, double* out);
/* Any spx_generator into out[B][P]; returns spx_status. */
int orc_generate_params(const spx_sweep_spec* spec, int32_t num_params, double* out);

/* nalgebra LU::new + solve (solver.rs:370-380): returns 1 if singular (x zeroed). */
int orc_lu_solve(int32_t n, const double* a_colmajor, const double* b, double* x);
/* Its factorisation alone in LAPACK getrf format (L strictly below the
 * diagonal, U on and above, 0-based ipiv); returns 1 if singular (zeroed). */
int orc_lu_factor(int32_t n, const double* a_colmajor, double* lu, int32_t* ipiv);
/* Linear solver of the Newton and transient solves: 0 (default) the
 * reference's size rule (dense LU for n < 50, the CachedSparseLu restatement
 * of sparse_lu.h for n >= 50, linear.rs:22 / newton.rs:92-108), 1 dense LU
 * at every size. */
void orc_set_linear_solver(int32_t mode);
/* CachedSparseLu::new + solve on triplets (duplicates summed); returns 1 if
 * singular (x zeroed). */
int orc_sparse_lu_solve(int32_t n, int64_t nnz, const int32_t* rows, const int32_t* cols,
                        const double* vals, const double* b, double* x);
/* CpuBatchedSolver::solve_batch (solver.rs:327-385) plus the non-finite rule of
 * FaerBatchedSolver (faer_solver.rs:74-86). */
int orc_solve_batch(const double* matrices, const double* rhs, int64_t n, int64_t batch,
                    double* solutions, uint8_t* singular, int64_t* num_singular);

/* Device models. */
double orc_thermal_voltage(double temp_k);                                /* diode.rs:41-45 */
double orc_diode_limit(double vd, double nvt);                             /* diode.rs:139-148 */
void orc_diode_evaluate(double is, double n, double vd, int32_t limiter,
                        double* id, double* gd);                          /* diode.rs:91-108 */
void orc_mosfet_evaluate(int32_t pmos, double vto, double kp, double lambda, double w, double l,
                         double vgs, double vds, double* ids, double* gds, double* gm,
                         int32_t* region);                                 /* level1.rs:154-191 */

/* MNA assembly of the whole circuit linearised at x, dense column-major J (n*n)
 * and rhs (n): netlist.rs:441-460 + mna.rs:62-169. */
int32_t orc_circuit_size(const spx_circuit* ckt);
void orc_stamp(const spx_circuit* ckt, const double* params, const double* x, double gmin,
               double source_factor, int32_t limiter, double* j_colmajor, double* rhs);

/* solve_newton_raphson (newton.rs:70-131) at fixed gmin / source factor.
 * Returns 1 converged, 0 not converged, -1 singular. */
int orc_newton(const spx_circuit* ckt, const double* params, const spx_nr_config* cfg,
               double gmin, double source_factor, const double* x0, double* x,
               int32_t* iterations);

/* max_i |KCL_i| / tol_i of x (<= 1 passes the SPICE test). */
double orc_kcl_residual(const spx_circuit* ckt, const double* params, const spx_nr_config* cfg,
                        const double* x, double gmin);

/* One instance: plain NR -> Gmin stepping (newton.rs:373-459) -> source
 * stepping (newton.rs:261-356).  Returns the spx_conv_status. */
int orc_dcop(const spx_circuit* ckt, const double* params, const spx_nr_config* cfg,
             const double* x0, double* x, int32_t* iterations, int32_t* flags,
             double* residual);

/* The whole sweep over `nthreads` host threads (one task per instance, as
 * rayon par_iter in parallel.rs:150-155).  Host buffers. */
int orc_sweep_dcop(const spx_circuit* ckt, const spx_sweep_spec* spec, const spx_nr_config* cfg,
                   const double* x0, spx_sweep_result* out, int32_t nthreads);

/* Same, on caller-supplied parameters [B][P] (no generator). */
int orc_sweep_dcop_params(const spx_circuit* ckt, const double* params, int64_t batch,
                          const spx_nr_config* cfg, const double* x0, spx_sweep_result* out,
                          int32_t nthreads);

/* StatisticsAccumulator add/merge (statistics.rs:135-154): out[n][5]. */
void orc_stats_accumulate(const double* values, const uint8_t* mask, int64_t batch, int32_t n,
                          double* out);
/* SweepStatistics::from_moments (statistics.rs:54-74): out = mean,variance,std_dev */
void orc_stats_from_moments(int64_t count, double sum, double sum_sq, double* out3);
/* spicier-solver SweepStatistics::from_samples (sweep.rs:109-136):
 * out = mean, std_dev, min, max. */
void orc_sweep_statistics(const double* samples, int64_t count, double* out4);

/* Transient options: fixed steps of h with an integration method, or
 * adaptive trapezoidal stepping (see spx_tran_adaptive). */
typedef struct orc_tran_opts {
    int32_t method;   /* SPX_TRAN_* (fixed steps) */
    double h;
    int32_t steps;
    int32_t adaptive;
    double tstop, h_init, h_min, h_max, reltol, abstol;
    int32_t max_points;
} orc_tran_opts;
/* Transient of one instance with orc_tran_opts (transient/solver.rs:41-249,
 * 461-676): waves [points][np] and times [points] (points = steps+1, or
 * max_points when adaptive).  Returns the spx_conv_status of the run. */
int orc_transient_ex(const spx_circuit* c, const double* params, const spx_nr_config* cfg,
                     const orc_tran_opts* o, int32_t np, const int32_t* probes, double* x,
                     double* waves, double* times, int32_t* num_points, int32_t* iterations,
                     int32_t* flags, double* residual, int32_t* fail_step, int32_t* total_steps,
                     int32_t* rejected_steps, double* min_step, double* max_step);
int orc_transient_ex_sweep(const spx_circuit* c, const double* params, int64_t B,
                           const spx_nr_config* cfg, const orc_tran_opts* o, int32_t np,
                           const int32_t* probes, double* x, int32_t* status, double* waves,
                           double* times, int32_t* num_points, int32_t* iterations,
                           int32_t* flags, double* residual, int32_t* fail_step,
                           int32_t* total_steps, int32_t* rejected_steps, double* min_step,
                           double* max_step, int32_t nthreads);

/* Backward-Euler transient of one instance (transient/solver.rs:41-160);
 * waves is [(steps+1)][np].  Returns the spx_conv_status of the run. */
int orc_transient(const spx_circuit* c, const double* params, const spx_nr_config* cfg, double h,
                  int32_t steps, int32_t np, const int32_t* probes, double* x, double* waves,
                  int32_t* iterations, int32_t* flags, double* residual, int32_t* fail_step);
int orc_transient_sweep(const spx_circuit* c, const double* params, int64_t B,
                        const spx_nr_config* cfg, double h, int32_t steps, int32_t np,
                        const int32_t* probes, double* x, int32_t* status, double* waves,
                        int32_t* iterations, int32_t* flags, double* residual,
                        int32_t* fail_step, int32_t nthreads);

/* Restarted GMRES(restart) with right Jacobi preconditioning (jacobi != 0)
 * or none: gmres/real.rs:30-400, preconditioner.rs:43-100.  Triplets with
 * duplicates summed; x starts at 0.  Returns 1 converged / 0 not. */
int orc_gmres(int32_t n, int64_t nnz, const int32_t* rows, const int32_t* cols,
              const double* vals, const double* b, int32_t restart, int32_t max_iter,
              double tol, int32_t jacobi, double* x, int32_t* iterations, double* residual);

#ifdef __cplusplus
}
#endif
#endif
