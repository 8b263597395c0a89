// oracle.cpp — CPU FP64 restatement of spicier's sweep / DC-OP path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
// the smoke check and bench.py's cpu_baseline leg load this library.
//
// Every function follows the reference file:line named above it, keeping the
// reference's operation order so that -ffp-contract=off reproduces its
// rounding.  Where the reference delegates to a third-party crate (nalgebra
// 0.33.2 LU, faer 0.24.0 sparse LU) the crate's published algorithm is
// restated and the deviation noted.
#include "oracle.h"
#include "sparse_lu.h"

#include <atomic>
#include <cmath>
#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <thread>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- rng.rs

}  // namespace

extern "C" uint64_t orc_splitmix64(uint64_t x) {  // rng.rs:39-44
  x = x + 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

extern "C" uint64_t orc_combine_indices(uint64_t seed, uint32_t s, uint32_t p) {  // rng.rs:51-56
  return seed ^ ((uint64_t)s * 0x517cc1b727220a95ULL) ^ ((uint64_t)p * 0x5851f42d4c957f2dULL);
}

extern "C" double orc_uniform(uint64_t seed, uint32_t s, uint32_t p) {  // rng.rs:80-85
  uint64_t h = orc_splitmix64(orc_combine_indices(seed, s, p));
  return (double)(h >> 11) * (1.0 / (double)(1ULL << 53));
}

extern "C" float orc_uniform_f32(uint64_t seed, uint32_t s, uint32_t p) {  // rng.rs:91-95
  uint64_t h = orc_splitmix64(orc_combine_indices(seed, s, p));
  return (float)(h >> 40) * (1.0f / (float)(1ULL << 24));
}

extern "C" double orc_gaussian(uint64_t seed, uint32_t s, uint32_t p) {  // rng.rs:120-131
  double u1 = orc_uniform(seed, s, p * 2u);
  double u2 = orc_uniform(seed, s, p * 2u + 1u);
  u1 = std::fmax(u1, 1e-10);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

extern "C" float orc_gaussian_f32(uint64_t seed, uint32_t s, uint32_t p) {  // rng.rs:136-143
  float u1 = orc_uniform_f32(seed, s, p * 2u);
  float u2 = orc_uniform_f32(seed, s, p * 2u + 1u);
  u1 = std::fmax(u1, 1e-7f);
  return std::sqrt(-2.0f * std::log(u1)) * std::cos(2.0f * 3.14159265358979323846f * u2);
}

extern "C" double orc_gaussian_scaled(uint64_t seed, uint32_t s, uint32_t p, double mean,
                                      double sigma) {  // rng.rs:148-150
  return mean + orc_gaussian(seed, s, p) * sigma;
}

extern "C" void orc_generate_gaussian_parameters(uint64_t seed, int64_t num_sweeps,
                                                 int32_t num_params, const double* means,
                                                 const double* sigmas, double* out) {
  // rng.rs:208-227: sweep-major.
  for (int64_t s = 0; s < num_sweeps; ++s)
    for (int32_t p = 0; p < num_params; ++p)
      out[s * num_params + p] =
          orc_gaussian_scaled(seed, (uint32_t)s, (uint32_t)p, means[p], sigmas[p]);
}

namespace {

// Rust f64::clamp(min, max).
inline double rclamp(double v, double lo, double hi) {
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return v;
}

// ------------------------------------------- spicier-solver/src/sweep.rs:170-200
struct Lcg {
  uint64_t s;
  double next_random() {  // :175-177
    s = s * 6364136223846793005ULL + 1ULL;
    return (double)(s >> 33) / (double)(1ULL << 31);
  }
  double next_normal() {  // :181-185
    double u1 = std::fmax(next_random(), 1e-10);
    double u2 = next_random();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
  }
};

bool is_varied(const spx_param_spec& p) { return p.kind != SPX_P_CONST; }

}  // namespace

extern "C" void orc_monte_carlo_lcg(uint64_t seed, int64_t num_samples, int32_t num_vars,
                                    const double* nominal, const double* sigma,
                                    const double* vmin, const double* vmax, double* out) {
  Lcg g{seed};
  for (int64_t i = 0; i < num_samples; ++i)
    for (int32_t v = 0; v < num_vars; ++v) {
      double z = g.next_normal();
      out[i * num_vars + v] = rclamp(nominal[v] + z * sigma[v] * nominal[v], vmin[v], vmax[v]);
    }
}

extern "C" int orc_generate_params(const spx_sweep_spec* spec, int32_t P, double* out) {
  const int64_t B = spec->batch;
  if (B < 0 || P < 0) return SPX_ERR_INVALID_ARGUMENT;
  const spx_param_spec* ps = spec->params;
  int nvar = 0;
  for (int p = 0; p < P; ++p) nvar += is_varied(ps[p]);
  switch (spec->generator) {
    case SPX_GEN_HASH:
      for (int64_t i = 0; i < B; ++i) {
        uint32_t s = (uint32_t)(spec->sweep_offset + i);
        for (int p = 0; p < P; ++p) {
          const spx_param_spec& q = ps[p];
          uint32_t r = (uint32_t)q.rng_index;
          double v;
          switch (q.kind) {
            case SPX_P_CONST: v = q.a; break;
            case SPX_P_GAUSS_REL: {
              double z = orc_gaussian(spec->seed, s, r);
              v = rclamp(q.a + z * q.b * q.a, q.c, q.d);
              break;
            }
            case SPX_P_GAUSS: v = orc_gaussian_scaled(spec->seed, s, r, q.a, q.b); break;
            case SPX_P_UNIFORM: v = q.a + (q.b - q.a) * orc_uniform(spec->seed, s, r); break;
            case SPX_P_LOGUNIFORM: {
              double la = std::log(q.a), lb = std::log(q.b);
              v = std::exp(la + (lb - la) * orc_uniform(spec->seed, s, r));
              break;
            }
            default: return SPX_ERR_INVALID_ARGUMENT;
          }
          out[i * P + p] = v;
        }
      }
      return SPX_OK;
    case SPX_GEN_LCG: {
      // The generator's state is sequential over (sample, variation); the
      // sweep offset skips offset*nvar normals (two draws each).
      Lcg g{spec->seed};
      for (int64_t k = 0; k < 2 * nvar * spec->sweep_offset; ++k) g.next_random();
      for (int64_t i = 0; i < B; ++i)
        for (int p = 0; p < P; ++p) {
          const spx_param_spec& q = ps[p];
          if (q.kind == SPX_P_CONST) { out[i * P + p] = q.a; continue; }
          if (q.kind != SPX_P_GAUSS_REL) return SPX_ERR_INVALID_ARGUMENT;
          double z = g.next_normal();
          out[i * P + p] = rclamp(q.a + z * q.b * q.a, q.c, q.d);
        }
      return SPX_OK;
    }
    case SPX_GEN_CORNER: {  // sweep.rs:204-227
      if (nvar >= 63 || B != (int64_t(1) << nvar)) return SPX_ERR_INVALID_DIMENSION;
      for (int64_t i = 0; i < B; ++i) {
        int j = 0;
        for (int p = 0; p < P; ++p) {
          const spx_param_spec& q = ps[p];
          if (!is_varied(q)) { out[i * P + p] = q.a; continue; }
          out[i * P + p] = ((i >> j) & 1) == 0 ? q.c : q.d;
          ++j;
        }
      }
      return SPX_OK;
    }
    case SPX_GEN_LINEAR:  // sweep.rs:231-266
      for (int64_t i = 0; i < B; ++i) {
        double t = B > 1 ? (double)i / (double)(B - 1) : 0.5;
        for (int p = 0; p < P; ++p) {
          const spx_param_spec& q = ps[p];
          out[i * P + p] = is_varied(q) ? q.c + t * (q.d - q.c) : q.a;
        }
      }
      return SPX_OK;
    case SPX_GEN_EXPLICIT:
      if (!spec->explicit_params) return SPX_ERR_INVALID_ARGUMENT;
      std::memcpy(out, spec->explicit_params, sizeof(double) * (size_t)(B * P));
      return SPX_OK;
    default:
      return SPX_ERR_INVALID_ARGUMENT;
  }
}

// ----------------------------------------------------------- dense LU
// nalgebra 0.33.2 LU::new (partial pivoting, first max |a| on ties, a zero
// pivot column is skipped) then LU::solve (permute, unit-lower forward
// substitution, upper back substitution with a zero-diagonal check), as used by
// CpuBatchedSolver (solver.rs:370-380) and solve_dense (linear.rs:506).
// The multiplier is formed as a * (1/pivot) and the update as (-u)*l + a,
// the order nalgebra's gauss_step / axpy use.
namespace {
// nalgebra LU::new (gauss_step): in place on the column-major a, perm_from[i]
// the row swapped with row i at step i; a zero pivot column is skipped.
void lu_factor_inplace(int n, double* a, int* perm_from) {
  auto A = [&](int r, int c) -> double& { return a[(size_t)c * n + r]; };
  for (int i = 0; i < n; ++i) {
    int piv = i;
    double best = std::fabs(A(i, i));
    for (int r = i + 1; r < n; ++r) {
      double v = std::fabs(A(r, i));
      if (v > best) { best = v; piv = r; }
    }
    perm_from[i] = piv;
    double diag = A(piv, i);
    if (diag == 0.0) continue;
    if (piv != i) {
      for (int c = 0; c < n; ++c) std::swap(A(i, c), A(piv, c));
    }
    double inv = 1.0 / diag;
    for (int r = i + 1; r < n; ++r) A(r, i) *= inv;
    for (int c = i + 1; c < n; ++c) {
      double u = A(i, c);
      for (int r = i + 1; r < n; ++r) A(r, c) = (-u) * A(r, i) + A(r, c);
    }
  }
}
}  // namespace

// The factorisation alone, in LAPACK getrf's format: lu column-major with the
// unit-lower L strictly below the diagonal and U on and above it, ipiv[i] the
// 0-based row swapped with row i at step i.  Returns 1 (singular) if a pivot
// is zero or a factor is not finite; lu and ipiv are then zeroed.
extern "C" int orc_lu_factor(int32_t n, const double* a_cm, double* lu, int32_t* ipiv) {
  std::vector<double> a(a_cm, a_cm + (size_t)n * n);
  std::vector<int> perm_from(n);
  lu_factor_inplace(n, a.data(), perm_from.data());
  bool singular = false;
  for (int i = 0; i < n; ++i) singular |= !(a[(size_t)i * n + i] != 0.0);
  for (size_t t = 0; t < a.size(); ++t) singular |= !std::isfinite(a[t]);
  for (size_t t = 0; t < a.size(); ++t) lu[t] = singular ? 0.0 : a[t];
  for (int i = 0; i < n; ++i) ipiv[i] = singular ? 0 : perm_from[i];
  return singular ? 1 : 0;
}

extern "C" int orc_lu_solve(int32_t n, const double* a_cm, const double* b, double* x) {
  std::vector<double> a(a_cm, a_cm + (size_t)n * n);  // column-major
  auto A = [&](int r, int c) -> double& { return a[(size_t)c * n + r]; };
  std::vector<double> y(b, b + n);
  std::vector<int> perm_from(n);
  lu_factor_inplace(n, a.data(), perm_from.data());
  // apply the row transpositions to b in order
  for (int i = 0; i < n; ++i)
    if (perm_from[i] != i) std::swap(y[i], y[perm_from[i]]);
  for (int i = 0; i < n; ++i) {
    double coeff = y[i];
    for (int r = i + 1; r < n; ++r) y[r] = (-coeff) * A(r, i) + y[r];
  }
  bool singular = false;
  for (int i = n - 1; i >= 0; --i) {
    double d = A(i, i);
    if (d == 0.0) { singular = true; break; }
    double coeff = y[i] / d;
    y[i] = coeff;
    for (int r = 0; r < i; ++r) y[r] = (-coeff) * A(r, i) + y[r];
  }
  if (!singular)
    for (int i = 0; i < n; ++i)
      if (!std::isfinite(y[i])) { singular = true; break; }
  for (int i = 0; i < n; ++i) x[i] = singular ? 0.0 : y[i];
  return singular ? 1 : 0;
}

extern "C" int orc_solve_batch(const double* m, const double* rhs, int64_t n, int64_t batch,
                               double* sol, uint8_t* singular, int64_t* num_singular) {
  if (n < 0 || batch < 0) return SPX_ERR_INVALID_DIMENSION;
  int64_t ns = 0;
  for (int64_t i = 0; i < batch; ++i) {
    int s = orc_lu_solve((int32_t)n, m + i * n * n, rhs + i * n, sol + i * n);
    if (singular) singular[i] = (uint8_t)s;
    ns += s;
  }
  if (num_singular) *num_singular = ns;
  return SPX_OK;
}

// ----------------------------------------------------------- devices
extern "C" double orc_thermal_voltage(double temp_k) {  // diode.rs:41-45
  const double k = 1.380649e-23, q = 1.602176634e-19;
  return k * temp_k / q;
}

extern "C" double orc_diode_limit(double vd, double nvt) {  // diode.rs:139-148
  double vcrit = nvt * std::log(nvt / (1.41421356237309504880 * 1e-14));
  if (vd > vcrit) {
    double arg = (vd - vcrit) / nvt;
    return vcrit + nvt * (1.0 + std::log1p(arg));
  }
  return vd;
}

extern "C" void orc_diode_evaluate(double is, double n, double vd, int32_t limiter, double* id,
                                   double* gd) {  // diode.rs:91-108
  double vt = orc_thermal_voltage(300.15);
  double nvt = n * vt;
  double v = limiter == SPX_LIMIT_REFERENCE ? orc_diode_limit(vd, nvt) : vd;
  double e = std::exp(v / nvt);
  *id = is * (e - 1.0);
  double g = is * e / nvt;
  *gd = std::fmax(g, 1e-12);
}

extern "C" void orc_mosfet_evaluate(int32_t pmos, double vto, double kp, double lambda, double w,
                                    double l, double vgs, double vds, double* ids, double* gds,
                                    double* gm, int32_t* region) {  // level1.rs:154-191
  if (pmos) { vgs = -vgs; vds = -vds; }
  double vth = std::fabs(vto);
  double beta = kp * 1.0 * w / l;  // level1.rs:68-70, nf = 1
  double i, gd, g;
  if (vgs < vth) {
    *ids = 0.0; *gds = 1e-12; *gm = 0.0; *region = 0;
    return;
  } else if (vds < vgs - vth) {
    double vov = vgs - vth;
    i = beta * (vov * vds - 0.5 * vds * vds) * (1.0 + lambda * vds);
    gd = beta * (vov - vds) * (1.0 + lambda * vds) + beta * (vov * vds - 0.5 * vds * vds) * lambda;
    g = beta * vds * (1.0 + lambda * vds);
    *region = 1;
  } else {
    double vov = vgs - vth;
    i = 0.5 * beta * vov * vov * (1.0 + lambda * vds);
    gd = 0.5 * beta * vov * vov * lambda;
    g = beta * vov * (1.0 + lambda * vds);
    *region = 2;
  }
  *ids = pmos ? -i : i;
  *gds = std::fmax(gd, 1e-12);
  *gm = g;
}

// ----------------------------------------------------------- MNA assembly
namespace {
// Devices with num_current_vars() == 1: voltage sources, Inductor, Vcvs, Ccvs
// (sources.rs, passive.rs:341-350, controlled.rs:111-113, 400-402).
bool has_branch(int kind) {
  return kind == SPX_EL_VSOURCE || kind == SPX_EL_VPULSE || kind == SPX_EL_VSIN ||
         kind == SPX_EL_VPWL || kind == SPX_EL_INDUCTOR || kind == SPX_EL_VCVS ||
         kind == SPX_EL_CCVS;
}
// The branch unknown of element e (which has one): its current_index, the
// count of branch devices before it (parser/elements.rs next_current_index).
int branch_row(const spx_circuit* c, int e) {
  int k = 0;
  for (int i = 0; i < e; ++i) k += has_branch(c->elements[i].kind);
  return c->num_nodes + k;
}
}  // namespace

extern "C" int32_t orc_circuit_size(const spx_circuit* c) {
  int32_t nv = 0;
  for (int e = 0; e < c->num_elements; ++e) nv += has_branch(c->elements[e].kind);
  return c->num_nodes + nv;
}

namespace {

// Triplet MNA (mna.rs:13-170) assembled straight into a dense column-major
// matrix: adding each triplet in push order onto zeros equals
// to_dense_matrix (mna.rs:62-71).
struct Mna {
  int n, nn;
  double* J;
  double* rhs;
  std::vector<orc_sparse::Triplet>* trips;  // non-null: the triplet list instead of J
  void add(int r, int c, double v) {
    if (trips)
      trips->push_back({r, c, v});
    else
      J[(size_t)c * n + r] += v;
  }
  void conductance(int i, int j, double g) {  // mna.rs:87-98
    if (i >= 0) add(i, i, g);
    if (j >= 0) add(j, j, g);
    if (i >= 0 && j >= 0) { add(i, j, -g); add(j, i, -g); }
  }
  void current(int i, int j, double I) {  // mna.rs:103-115
    if (i >= 0) rhs[i] += -I;
    if (j >= 0) rhs[j] += I;
  }
  void vsource(int i, int j, int k, double V) {  // mna.rs:124-145
    int row = nn + k;
    if (i >= 0) { add(i, row, 1.0); add(row, i, 1.0); }
    if (j >= 0) { add(j, row, -1.0); add(row, j, -1.0); }
    rhs[row] = V;
  }
};

inline int idx(int32_t node) { return node - 1; }  // NodeId -> unknown, ground -> -1

// Linear solver inside Newton and the transient: 0 the reference's rule
// (SPARSE_THRESHOLD = 50, linear.rs:22: dense LU below, CachedSparseLu at and
// above), 1 dense LU at every size.
std::atomic<int> g_linear_mode{0};
inline bool use_sparse(int n) { return g_linear_mode.load() == 0 && n >= 50; }
inline double volt(const double* x, int32_t node) { return node > 0 ? x[node - 1] : 0.0; }
inline double P(const double* params, int32_t slot) { return slot >= 0 ? params[slot] : 0.0; }

}  // namespace

namespace {

// SPICE3 DEVpnjlim: limits the new junction voltage against the previous
// iteration's when it is above vcrit and moved by more than 2*nvt.
double pnjlim(double vnew, double vold, double vt, double vcrit, int* limited) {
  if (vnew > vcrit && std::fabs(vnew - vold) > vt + vt) {
    if (vold > 0) {
      double arg = 1 + (vnew - vold) / vt;
      vnew = arg > 0 ? vold + vt * std::log(arg) : vcrit;
    } else {
      vnew = vt * std::log(vnew / vt);
    }
    if (limited) *limited = 1;
  }
  return vnew;
}

// Transient context: mode 1 is a step ending at time t; without it (or mode
// 0) capacitors are open and PULSE sources sit at V1 (waveforms.rs:150-160).
// comp picks the capacitor companion (companion.rs:41-130):
//   0 (BE, stamp_be)       G = C/h, I = G*vprev
//   1 (trap, stamp_trap)   G = 2C/h, I = G*vprev + iprev
//   2 (stamp_trbdf2_bdf2)  G = C/bden, I = G*(a1*vprev + a2*vpp)
// with h the companion's step (gamma*h for the TR-BDF2 first stage).
struct OTran {
  int mode;
  double t, h;
  const double* vprev;
  int comp = 0;
  const double* iprev = nullptr;
  const double* vpp = nullptr;
  double a1 = 0.0, a2 = 0.0, bden = 0.0;
};

// eval_pulse (waveforms.rs:160-180)
double eval_pulse(double v1, double v2, double td, double tr, double tf, double pw, double per,
                  double t) {
  if (t < td) return v1;
  const double tr_ = per > 0.0 ? std::fmod(t - td, per) : t - td;
  if (tr_ < tr) return v1 + (v2 - v1) * tr_ / tr;
  if (tr_ < tr + pw) return v2;
  if (tr_ < tr + pw + tf) return v2 - (v2 - v1) * (tr_ - tr - pw) / tf;
  return v1;
}

// eval_sin (waveforms.rs:185-201)
double eval_sin(double vo, double va, double freq, double td, double theta, double phase,
                double t) {
  if (t < td) return vo;
  const double t_rel = t - td;
  const double phase_rad = phase * M_PI / 180.0;
  const double damping = theta > 0.0 ? std::exp(-theta * t_rel) : 1.0;
  return vo + va * damping * std::sin(2.0 * M_PI * freq * t_rel + phase_rad);
}

// eval_pwl (waveforms.rs:204-232) over np (t, v) pairs in params[first..]
double eval_pwl(const double* params, int first, int np, double t) {
  const double* p = params + first;
  if (t <= p[0]) return p[1];
  if (t >= p[2 * (np - 1)]) return p[2 * (np - 1) + 1];
  for (int i = 0; i < np - 1; ++i) {
    const double t0 = p[2 * i], v0 = p[2 * i + 1], t1 = p[2 * i + 2], v1 = p[2 * i + 3];
    if (t >= t0 && t <= t1) {
      const double frac = (t - t0) / (t1 - t0);
      return v0 + frac * (v1 - v0);
    }
  }
  return p[2 * (np - 1) + 1];
}

void stamp_impl(const spx_circuit* c, const double* params, const double* x, double gmin,
                double sf, int32_t limiter, double* J, double* rhs, double* vstate,
                int* limited, const OTran* tc = nullptr,
                std::vector<orc_sparse::Triplet>* trips = nullptr) {
  const int n = orc_circuit_size(c);
  if (trips)
    trips->clear();
  else
    std::memset(J, 0, sizeof(double) * (size_t)n * n);
  std::memset(rhs, 0, sizeof(double) * (size_t)n);
  Mna m{n, c->num_nodes, J, rhs, trips};
  int k = 0;  // voltage-source branch counter
  for (int e = 0; e < c->num_elements; ++e) {
    const spx_element& el = c->elements[e];
    const int a = idx(el.node[0]), b = idx(el.node[1]);
    switch (el.kind) {
      case SPX_EL_RESISTOR:
        m.conductance(a, b, 1.0 / P(params, el.param[0]));
        break;
      case SPX_EL_CAPACITOR:  // open at DC (passive.rs:233-245); BE companion in a step
        if (tc && tc->mode == 1) {
          const double C = P(params, el.param[0]);
          double geq, ieq;
          if (tc->comp == 1) {  // stamp_trap, companion.rs:53-60
            geq = 2.0 * C / tc->h;
            ieq = geq * tc->vprev[e] + tc->iprev[e];
          } else if (tc->comp == 2) {  // stamp_trbdf2_bdf2, companion.rs:100-130
            geq = C / tc->bden;
            ieq = geq * (tc->a1 * tc->vprev[e] + tc->a2 * tc->vpp[e]);
          } else {  // stamp_be, companion.rs:41-48
            geq = C / tc->h;
            ieq = geq * tc->vprev[e];
          }
          m.conductance(a, b, geq);
          m.current(b, a, ieq);
        }
        break;
      case SPX_EL_VSOURCE:
        m.vsource(a, b, k++, P(params, el.param[0]) * sf);
        break;
      case SPX_EL_VPULSE: {
        double v = P(params, el.param[0]);
        if (tc && tc->mode == 1)
          v = eval_pulse(P(params, el.param[0]), P(params, el.param[1]), P(params, el.param[2]),
                         P(params, el.param[3]), P(params, el.param[4]), P(params, el.param[5]),
                         P(params, el.param[6]), tc->t);
        m.vsource(a, b, k++, v * sf);
        break;
      }
      case SPX_EL_VPWL: {  // DC value v1 (waveforms.rs:153)
        double v = params[el.param[0] + 1];
        if (tc && tc->mode == 1) v = eval_pwl(params, el.param[0], el.ctrl, tc->t);
        m.vsource(a, b, k++, v * sf);
        break;
      }
      case SPX_EL_VSIN: {
        double v = P(params, el.param[0]);
        if (tc && tc->mode == 1)
          v = eval_sin(P(params, el.param[0]), P(params, el.param[1]), P(params, el.param[2]),
                       P(params, el.param[3]), P(params, el.param[4]), P(params, el.param[5]),
                       tc->t);
        m.vsource(a, b, k++, v * sf);
        break;
      }
      case SPX_EL_INDUCTOR: {
        const int kb = k++;
        if (tc && tc->mode == 1) {
          // InductorState companions (companion.rs:197-243): the conductance
          // G with I from node 0 to node 1 and no branch current, as the
          // reference's transient MNA has none (stampers.rs:515-535); the
          // branch unknown stays as a decoupled row x = 0.
          const double L = P(params, el.param[0]);
          double geq, ieq;
          if (tc->comp == 1) {  // stamp_trap
            geq = tc->h / (2.0 * L);
            ieq = tc->iprev[e] + geq * tc->vprev[e];
          } else if (tc->comp == 2) {  // stamp_trbdf2_bdf2
            geq = tc->bden / L;
            ieq = tc->a1 * tc->iprev[e] + tc->a2 * tc->vpp[e];
          } else {  // stamp_be
            geq = tc->h / L;
            ieq = tc->iprev[e];
          }
          m.conductance(a, b, geq);
          m.current(a, b, ieq);
          m.add(c->num_nodes + kb, c->num_nodes + kb, 1.0);
        } else {
          m.vsource(a, b, kb, 0.0);  // a 0 V source at DC (passive.rs:318-325)
        }
        break;
      }
      case SPX_EL_VCVS: {  // Vcvs::stamp, controlled.rs:58-90
        const int br = c->num_nodes + k++;
        const int cp = idx(el.node[2]), cn = idx(el.node[3]);
        const double gain = P(params, el.param[0]);
        if (a >= 0) m.add(a, br, 1.0);
        if (b >= 0) m.add(b, br, -1.0);
        if (a >= 0) m.add(br, a, 1.0);
        if (b >= 0) m.add(br, b, -1.0);
        if (cp >= 0) m.add(br, cp, -gain);
        if (cn >= 0) m.add(br, cn, gain);
        break;
      }
      case SPX_EL_VCCS: {  // Vccs::stamp, controlled.rs:172-201
        const int cp = idx(el.node[2]), cn = idx(el.node[3]);
        const double gm = P(params, el.param[0]);
        if (a >= 0) {
          if (cp >= 0) m.add(a, cp, -gm);
          if (cn >= 0) m.add(a, cn, gm);
        }
        if (b >= 0) {
          if (cp >= 0) m.add(b, cp, gm);
          if (cn >= 0) m.add(b, cn, -gm);
        }
        break;
      }
      case SPX_EL_CCCS: {  // Cccs::stamp, controlled.rs:267-285
        const int br = branch_row(c, el.ctrl);
        const double gain = P(params, el.param[0]);
        if (a >= 0) m.add(a, br, gain);
        if (b >= 0) m.add(b, br, -gain);
        break;
      }
      case SPX_EL_CCVS: {  // Ccvs::stamp, controlled.rs:351-375
        const int br = c->num_nodes + k++;
        const int ctrl_br = branch_row(c, el.ctrl);
        if (a >= 0) m.add(a, br, 1.0);
        if (b >= 0) m.add(b, br, -1.0);
        if (a >= 0) m.add(br, a, 1.0);
        if (b >= 0) m.add(br, b, -1.0);
        m.add(br, ctrl_br, -P(params, el.param[0]));
        break;
      }
      case SPX_EL_ISOURCE:
        m.current(a, b, P(params, el.param[0]) * sf);
        break;
      case SPX_EL_DIODE: {  // diode.rs:113-131, 190-199
        double vd = volt(x, el.node[0]) - volt(x, el.node[1]);
        double is = P(params, el.param[0]), nn = P(params, el.param[1]);
        double nvt = nn * orc_thermal_voltage(300.15);
        double vl, id, gd;
        if (limiter == SPX_LIMIT_PNJLIM) {
          double vold = vstate ? vstate[e] : 0.0;
          double vcrit = nvt * std::log(nvt / (1.41421356237309504880 * is));
          vl = pnjlim(vd, vold, nvt, vcrit, limited);
          if (vstate) vstate[e] = vl;
          double ex = std::exp(vl / nvt);
          id = is * (ex - 1.0);
          gd = std::fmax(is * ex / nvt, 1e-12);
        } else {
          vl = limiter == SPX_LIMIT_REFERENCE ? orc_diode_limit(vd, nvt) : vd;
          orc_diode_evaluate(is, nn, vl, limiter, &id, &gd);
        }
        double ieq = id - gd * vl;
        m.conductance(a, b, gd);
        m.current(a, b, ieq);
        break;
      }
      case SPX_EL_NMOS:
      case SPX_EL_PMOS: {  // level1.rs:202-248, 305-318
        const int d = idx(el.node[0]), g = idx(el.node[1]), s = idx(el.node[2]);
        double vg = volt(x, el.node[1]), vd = volt(x, el.node[0]), vs = volt(x, el.node[2]);
        double vgs = vg - vs, vds = vd - vs;
        double ids, gds, gm;
        int32_t region;
        orc_mosfet_evaluate(el.kind == SPX_EL_PMOS, P(params, el.param[0]), P(params, el.param[1]),
                            P(params, el.param[2]), P(params, el.param[3]), P(params, el.param[4]),
                            vgs, vds, &ids, &gds, &gm, &region);
        m.conductance(d, s, gds);
        if (d >= 0) {
          if (g >= 0) m.add(d, g, gm);
          if (s >= 0) m.add(d, s, -gm);
        }
        if (s >= 0) {
          if (g >= 0) m.add(s, g, -gm);
          m.add(s, s, gm);
        }
        double ieq = ids - gds * vds - gm * vgs;
        m.current(d, s, ieq);
        break;
      }
      default:
        break;
    }
  }
  if (gmin != 0.0)  // GminStamper (newton.rs:393-398) + stamp_gmin (mna.rs:165-169)
    for (int i = 0; i < c->num_nodes; ++i) m.add(i, i, gmin);
}

}  // namespace

extern "C" void orc_set_linear_solver(int32_t mode) { g_linear_mode.store(mode); }

extern "C" int orc_sparse_lu_solve(int32_t n, int64_t nnz, const int32_t* rows,
                                   const int32_t* cols, const double* vals, const double* b,
                                   double* x) {
  std::vector<orc_sparse::Triplet> t((size_t)nnz);
  for (int64_t i = 0; i < nnz; ++i) t[i] = {rows[i], cols[i], vals[i]};
  const orc_sparse::Csc a = orc_sparse::from_triplets(n, t);
  return orc_sparse::solve(orc_sparse::analyse(a), a, b, x);
}

extern "C" void orc_stamp(const spx_circuit* c, const double* params, const double* x,
                          double gmin, double sf, int32_t limiter, double* J, double* rhs) {
  stamp_impl(c, params, x, gmin, sf, limiter, J, rhs, nullptr, nullptr);
}

namespace {

// check_convergence (newton.rs:134-159)
bool converged_dx(const double* old, const double* nw, int nn, int n, const spx_nr_config* cfg) {
  for (int i = 0; i < nn; ++i) {
    double delta = std::fabs(nw[i] - old[i]);
    double tol = cfg->v_reltol * std::fmax(std::fabs(nw[i]), std::fabs(old[i])) + cfg->v_abstol;
    if (delta > tol) return false;
  }
  for (int i = nn; i < n; ++i) {
    double delta = std::fabs(nw[i] - old[i]);
    double tol = cfg->v_reltol * std::fmax(std::fabs(nw[i]), std::fabs(old[i])) + cfg->i_abstol;
    if (delta > tol) return false;
  }
  return true;
}

}  // namespace

namespace {
int newton_impl(const spx_circuit* c, const double* params, const spx_nr_config* cfg, double gmin,
                double sf, const double* x0, double* x, int32_t* iterations, const OTran* tc);
double kcl_impl(const spx_circuit* c, const double* params, const spx_nr_config* cfg,
                const double* x, double gmin, const OTran* tc);
}  // namespace

extern "C" int orc_newton(const spx_circuit* c, const double* params, const spx_nr_config* cfg,
                          double gmin, double sf, const double* x0, double* x,
                          int32_t* iterations) {
  return newton_impl(c, params, cfg, gmin, sf, x0, x, iterations, nullptr);
}

namespace {
int newton_impl(const spx_circuit* c, const double* params, const spx_nr_config* cfg, double gmin,
                double sf, const double* x0, double* x, int32_t* iterations, const OTran* tc) {
  const int n = orc_circuit_size(c);
  // the linear solve by size, as solve_newton_raphson (newton.rs:92-108):
  // dense LU below SPARSE_THRESHOLD, the cached sparse LU from the first
  // iteration's pattern at and above it
  const bool sparse = use_sparse(n);
  std::vector<double> J(sparse ? 0 : (size_t)n * n), rhs(n), xn(n), sol(n, 0.0);
  std::vector<orc_sparse::Triplet> trips;
  orc_sparse::Symbolic sym;
  std::vector<double> vstate(c->num_elements, 0.0);  // pnjlim junction voltages
  if (x0) std::memcpy(sol.data(), x0, sizeof(double) * n);
  for (int it = 0; it < cfg->max_iterations; ++it) {
    int limited = 0;
    stamp_impl(c, params, sol.data(), gmin, sf, cfg->limiter, J.data(), rhs.data(),
               vstate.data(), &limited, tc, sparse ? &trips : nullptr);
    int singular;
    if (sparse) {
      const orc_sparse::Csc a = orc_sparse::from_triplets(n, trips);
      if (it == 0) sym = orc_sparse::analyse(a);
      singular = orc_sparse::solve(sym, a, rhs.data(), xn.data());
    } else {
      singular = orc_lu_solve(n, J.data(), rhs.data(), xn.data());
    }
    if (singular) {
      std::memcpy(x, sol.data(), sizeof(double) * n);
      *iterations = it + 1;
      return -1;
    }
    bool conv = converged_dx(sol.data(), xn.data(), c->num_nodes, n, cfg) && !limited;
    sol = xn;
    if (conv) {
      std::memcpy(x, sol.data(), sizeof(double) * n);
      *iterations = it + 1;
      return 1;
    }
  }
  std::memcpy(x, sol.data(), sizeof(double) * n);
  *iterations = cfg->max_iterations;
  return 0;
}
}  // namespace

extern "C" double orc_kcl_residual(const spx_circuit* c, const double* params,
                                   const spx_nr_config* cfg, const double* x, double gmin) {
  return kcl_impl(c, params, cfg, x, gmin, nullptr);
}

namespace {
double kcl_impl(const spx_circuit* c, const double* params, const spx_nr_config* cfg,
                const double* x, double gmin, const OTran* tc) {
  // KCL at x with devices linearised, unlimited, at x: r = J(x) x - b(x)
  // equals the sum of element currents leaving each node.  For each row,
  // tol = reltol * max(|J_ij x_j|, |b_i|) + abstol (node rows, amperes) or
  // + vntol (branch rows, volts).  Returns max |r_i| / tol_i.
  const int n = orc_circuit_size(c);
  std::vector<double> J((size_t)n * n), rhs(n);
  stamp_impl(c, params, x, gmin, 1.0, SPX_LIMIT_NONE, J.data(), rhs.data(), nullptr, nullptr, tc);
  double worst = 0.0;
  for (int i = 0; i < n; ++i) {
    double r = -rhs[i], mag = std::fabs(rhs[i]);
    for (int j = 0; j < n; ++j) {
      double t = J[(size_t)j * n + i] * x[j];
      r += t;
      mag = std::fmax(mag, std::fabs(t));
    }
    double tol = cfg->kcl_reltol * mag + (i < c->num_nodes ? cfg->kcl_abstol : cfg->kcl_vntol);
    double q = std::fabs(r) / tol;
    if (!(q <= worst)) worst = q;  // NaN propagates as failure
  }
  return worst;
}
}  // namespace

namespace {

bool has_nonlinear(const spx_circuit* c) {
  for (int e = 0; e < c->num_elements; ++e) {
    int k = c->elements[e].kind;
    if (k == SPX_EL_DIODE || k == SPX_EL_NMOS || k == SPX_EL_PMOS) return true;
  }
  return false;
}

}  // namespace

extern "C" int orc_dcop(const spx_circuit* c, const double* params, const spx_nr_config* cfg,
                        const double* x0, double* x, int32_t* iterations, int32_t* flags,
                        double* residual) {
  const int n = orc_circuit_size(c);
  int32_t total = 0, it = 0, fl = (cfg->warm_start != SPX_WARM_ZERO && x0) ? SPX_FLAG_WARM_START : 0;
  double final_gmin = 0.0;
  int status;

  if (!has_nonlinear(c)) {
    // A linear circuit takes the reference sweep's single solve per point
    // (spicier-batched-sweep/src/sweep.rs:159-217).
    std::vector<double> J((size_t)n * n), rhs(n);
    std::vector<double> zero(n, 0.0);
    stamp_impl(c, params, zero.data(), 0.0, 1.0, cfg->limiter, J.data(), rhs.data(), nullptr,
               nullptr);
    status = orc_lu_solve(n, J.data(), rhs.data(), x) ? SPX_SINGULAR : SPX_CONVERGED;
    total = 1;
  } else {
    int st = orc_newton(c, params, cfg, 0.0, 1.0, x0, x, &it);
    total += it;
    status = st == 1 ? SPX_CONVERGED : (st < 0 ? SPX_SINGULAR : SPX_FAILED);

    // Gmin stepping (newton.rs:373-459).  Retries at one level restart from the
    // same solution and are therefore identical replays of the first attempt;
    // the first attempt decides the level.  A level that does not converge
    // adds its iterations once per attempt (max_attempts_per_level of them,
    // newton.rs:411-426), as total_iterations does.  On failure the stage's
    // solution (last converged level, zeros if none) is what it returns.
    if (status != SPX_CONVERGED && cfg->gmin_stepping) {
      fl |= SPX_FLAG_GMIN_STEPPING;
      std::vector<double> sol(n, 0.0), trial(n);
      double g = cfg->gmin_initial;
      bool ok = true;
      int last = 0;
      while (g >= cfg->gmin_final * 0.99) {
        last = orc_newton(c, params, cfg, g, 1.0, sol.data(), trial.data(), &it);
        total += it;
        if (last == 0) total += (cfg->gmin_max_attempts - 1) * it;
        if (last != 1) { ok = false; break; }
        sol = trial;
        final_gmin = g;
        if (g <= cfg->gmin_final) break;
        g /= cfg->gmin_factor;
        if (g < cfg->gmin_final) g = cfg->gmin_final;
      }
      std::memcpy(x, sol.data(), sizeof(double) * n);
      if (ok) {
        status = SPX_CONVERGED;
      } else {
        status = last < 0 ? SPX_SINGULAR : SPX_FAILED;
        final_gmin = 0.0;
      }
    }

    // Source stepping (newton.rs:261-356), same replay and counting remarks
    // (newton.rs:300-315).
    if (status != SPX_CONVERGED && cfg->source_stepping) {
      fl |= SPX_FLAG_SOURCE_STEPPING;
      std::vector<double> sol(n, 0.0), trial(n);
      double factor = cfg->src_initial, step = cfg->src_step;
      bool ok = true;
      int last = 1;
      while (factor <= 1.0 + 1e-9) {
        double cur = std::fmin(factor, 1.0);
        last = orc_newton(c, params, cfg, 0.0, cur, sol.data(), trial.data(), &it);
        total += it;
        if (last == 0) total += (cfg->src_max_attempts - 1) * it;
        if (last != 1) {
          if (step > cfg->src_min_step) {
            step /= 2.0;
            factor -= step;
            continue;
          }
          ok = false;
          break;
        }
        sol = trial;
        if (std::fabs(cur - 1.0) < 1e-9) break;
        factor += step;
        if (step < cfg->src_step) step = std::fmin(step * 1.5, cfg->src_step);
      }
      std::memcpy(x, sol.data(), sizeof(double) * n);
      if (ok) {
        status = SPX_CONVERGED;
        final_gmin = 0.0;
      } else {
        status = last < 0 ? SPX_SINGULAR : SPX_FAILED;
      }
    }
  }

  double res = orc_kcl_residual(c, params, cfg, x, final_gmin);
  if (status == SPX_CONVERGED && !(res <= 1.0)) fl |= SPX_FLAG_KCL_FAIL;
  if (status == SPX_SINGULAR)
    for (int i = 0; i < n; ++i) x[i] = 0.0;
  if (iterations) *iterations = total;
  if (flags) *flags = fl;
  if (residual) *residual = res;
  return status;
}

namespace {
void run_pool(int64_t count, int nthreads, const std::function<void(int64_t)>& body) {
  if (nthreads <= 1) {
    for (int64_t i = 0; i < count; ++i) body(i);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&] {
      for (;;) {
        int64_t i = next.fetch_add(16);
        if (i >= count) break;
        int64_t e = std::min<int64_t>(count, i + 16);
        for (; i < e; ++i) body(i);
      }
    });
  for (auto& t : th) t.join();
}
}  // namespace

extern "C" int orc_sweep_dcop_params(const spx_circuit* c, const double* params, int64_t B,
                                     const spx_nr_config* cfg, const double* x0,
                                     spx_sweep_result* out, int32_t nthreads) {
  const int n = orc_circuit_size(c);
  const int P = c->num_params;
  run_pool(B, nthreads, [&](int64_t i) {
    int32_t it = 0, fl = 0;
    double res = 0.0;
    int st = orc_dcop(c, params + i * P, cfg,
                      (cfg->warm_start != SPX_WARM_ZERO && x0) ? x0 + i * n : nullptr,
                      out->x + i * n, &it, &fl, &res);
    if (out->status) out->status[i] = st;
    if (out->iterations) out->iterations[i] = it;
    if (out->flags) out->flags[i] = fl;
    if (out->residual) out->residual[i] = res;
  });
  if (out->node_stats) {
    std::vector<uint8_t> mask(B);
    for (int64_t i = 0; i < B; ++i)
      mask[i] = out->status ? out->status[i] == SPX_CONVERGED : 1;
    orc_stats_accumulate(out->x, mask.data(), B, n, out->node_stats);
  }
  return SPX_OK;
}

extern "C" int orc_sweep_dcop(const spx_circuit* c, const spx_sweep_spec* spec,
                              const spx_nr_config* cfg, const double* x0, spx_sweep_result* out,
                              int32_t nthreads) {
  const int P = c->num_params;
  std::vector<double> local;
  double* params = out->params;
  if (!params) {
    local.resize((size_t)spec->batch * P);
    params = local.data();
  }
  int rc = orc_generate_params(spec, P, params);
  if (rc) return rc;
  return orc_sweep_dcop_params(c, params, spec->batch, cfg, x0, out, nthreads);
}

// ----------------------------------------------------------- statistics.rs
extern "C" void orc_stats_accumulate(const double* v, const uint8_t* mask, int64_t B, int32_t n,
                                     double* out) {
  for (int j = 0; j < n; ++j) {
    double cnt = 0, sum = 0, sq = 0, mn = INFINITY, mx = -INFINITY;
    for (int64_t i = 0; i < B; ++i) {
      if (mask && !mask[i]) continue;
      double x = v[i * n + j];
      cnt += 1;
      sum += x;
      sq += x * x;
      mn = std::fmin(mn, x);
      mx = std::fmax(mx, x);
    }
    out[j * 5 + 0] = cnt;
    out[j * 5 + 1] = sum;
    out[j * 5 + 2] = sq;
    out[j * 5 + 3] = mn;
    out[j * 5 + 4] = mx;
  }
}

extern "C" void orc_stats_from_moments(int64_t count, double sum, double sum_sq, double* o) {
  double mean = count > 0 ? sum / (double)count : 0.0;  // statistics.rs:54-74
  double var = count > 0 ? (sum_sq / (double)count) - mean * mean : 0.0;
  var = std::fmax(var, 0.0);
  o[0] = mean;
  o[1] = var;
  o[2] = std::sqrt(var);
}

extern "C" void orc_sweep_statistics(const double* s, int64_t count, double* o) {
  // spicier-solver/src/sweep.rs:109-136 (two-pass).
  if (count == 0) { o[0] = o[1] = o[2] = o[3] = 0.0; return; }
  double sum = 0.0;
  for (int64_t i = 0; i < count; ++i) sum += s[i];
  double mean = sum / (double)count, acc = 0.0;
  for (int64_t i = 0; i < count; ++i) acc += (s[i] - mean) * (s[i] - mean);
  double mn = INFINITY, mx = -INFINITY;
  for (int64_t i = 0; i < count; ++i) { mn = std::fmin(mn, s[i]); mx = std::fmax(mx, s[i]); }
  o[0] = mean;
  o[1] = std::sqrt(acc / (double)count);
  o[2] = mn;
  o[3] = mx;
}

// ----------------------------------------------------------- GMRES
// Restarted GMRES with right Jacobi preconditioning:
// solve_gmres_real_preconditioned (spicier-solver/src/gmres/real.rs:208-400)
// with JacobiPreconditioner (preconditioner.rs:43-100: M = diag(A), entries
// below 1e-30 in magnitude replaced by 1); jacobi == 0 is solve_gmres_real
// (real.rs:30-205), i.e. M = I.  Givens rotations as helpers.rs:40-48.  The
// operator sums duplicate triplets per (row, col) in triplet order and
// applies rows in column order; dot products are sequential (the reference's
// SIMD dot product sums in a different order, so agreement is to rounding).
extern "C" int orc_gmres(int32_t n, int64_t nnz, const int32_t* rows, const int32_t* cols,
                         const double* vals, const double* b, int32_t restart, int32_t max_iter,
                         double tol, int32_t jacobi, double* x_out, int32_t* iterations,
                         double* residual) {
  std::map<std::pair<int, int>, double> m;
  for (int64_t t = 0; t < nnz; ++t) m[{rows[t], cols[t]}] += vals[t];
  std::vector<int> rp(n + 1, 0), ci;
  std::vector<double> av;
  std::vector<double> diag(n, 0.0);
  for (auto& kv : m) {
    rp[kv.first.first + 1]++;
    ci.push_back(kv.first.second);
    av.push_back(kv.second);
    if (kv.first.first == kv.first.second) diag[kv.first.first] = kv.second;
  }
  for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
  std::vector<double> inv(n, 1.0);
  if (jacobi)
    for (int i = 0; i < n; ++i) inv[i] = std::fabs(diag[i]) < 1e-30 ? 1.0 : 1.0 / diag[i];
  auto apply = [&](const std::vector<double>& v, std::vector<double>& y) {
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int q = rp[i]; q < rp[i + 1]; ++q) s += av[q] * v[ci[q]];
      y[i] = s;
    }
  };
  auto dot = [&](const std::vector<double>& a, const std::vector<double>& c) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * c[i];
    return s;
  };
  std::vector<double> bv(b, b + n), x(n, 0.0);
  const double bn = std::sqrt(dot(bv, bv));
  *iterations = 0;
  if (bn < 1e-30) {
    for (int i = 0; i < n; ++i) x_out[i] = 0.0;
    *residual = 0.0;
    return 1;
  }
  int total = 0;
  std::vector<double> ax(n);
  for (int cycle = 0; cycle < max_iter; ++cycle) {
    apply(x, ax);
    std::vector<double> r(n);
    for (int i = 0; i < n; ++i) r[i] = bv[i] - ax[i];
    const double rn = std::sqrt(dot(r, r));
    if (rn / bn < tol) {
      for (int i = 0; i < n; ++i) x_out[i] = x[i];
      *iterations = total;
      *residual = rn / bn;
      return 1;
    }
    const int mk = std::min(restart, n);
    std::vector<std::vector<double>> v, z;
    std::vector<std::vector<double>> h(mk, std::vector<double>(mk + 1, 0.0));
    for (int i = 0; i < n; ++i) r[i] *= 1.0 / rn;
    v.push_back(r);
    std::vector<double> g(mk + 1, 0.0), cs(mk, 0.0), sn(mk, 0.0);
    g[0] = rn;
    int k = 0;
    while (k < mk) {
      ++total;
      if (total > max_iter) break;
      std::vector<double> zk(n);
      for (int i = 0; i < n; ++i) zk[i] = v[k][i] * inv[i];
      z.push_back(zk);
      std::vector<double> w(n);
      apply(z[k], w);
      for (int j = 0; j <= k; ++j) {
        const double hij = dot(v[j], w);
        h[k][j] = hij;
        for (int i = 0; i < n; ++i) w[i] -= hij * v[j][i];
      }
      const double wn = std::sqrt(dot(w, w));
      h[k][k + 1] = wn;
      if (wn < 1e-30) {
        ++k;
        break;
      }
      for (int i = 0; i < n; ++i) w[i] *= 1.0 / wn;
      v.push_back(w);
      for (int j = 0; j < k; ++j) {
        const double t = cs[j] * h[k][j] + sn[j] * h[k][j + 1];
        h[k][j + 1] = -sn[j] * h[k][j] + cs[j] * h[k][j + 1];
        h[k][j] = t;
      }
      double c = 1.0, s = 0.0;
      if (!(std::fabs(h[k][k + 1]) < 1e-30)) {
        const double rr = std::sqrt(h[k][k] * h[k][k] + h[k][k + 1] * h[k][k + 1]);
        c = h[k][k] / rr;
        s = h[k][k + 1] / rr;
      }
      cs[k] = c;
      sn[k] = s;
      h[k][k] = c * h[k][k] + s * h[k][k + 1];
      h[k][k + 1] = 0.0;
      const double tg = c * g[k] + s * g[k + 1];
      g[k + 1] = -s * g[k] + c * g[k + 1];
      g[k] = tg;
      ++k;
      if (std::fabs(g[k]) / bn < tol) break;
    }
    std::vector<double> y(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
      double sum = g[i];
      for (int j = i + 1; j < k; ++j) sum -= h[j][i] * y[j];
      if (std::fabs(h[i][i]) > 1e-30) y[i] = sum / h[i][i];
    }
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < n; ++j) x[j] += z[i][j] * y[i];
    apply(x, ax);
    double fr = 0.0;
    for (int i = 0; i < n; ++i) fr += (bv[i] - ax[i]) * (bv[i] - ax[i]);
    fr = std::sqrt(fr);
    if (fr / bn < tol || total >= max_iter) {
      for (int i = 0; i < n; ++i) x_out[i] = x[i];
      *iterations = total;
      *residual = fr / bn;
      return fr / bn < tol ? 1 : 0;
    }
  }
  for (int i = 0; i < n; ++i) x_out[i] = x[i];
  *iterations = total;
  *residual = NAN;
  return 0;
}

// ----------------------------------------------------------- transient
namespace {

// TRBDF2_GAMMA (types.rs:11) and the BDF2-stage coefficients of
// stamp_trbdf2_bdf2 (companion.rs:100-124) for a full step h.
const double kGamma = 2.0 - M_SQRT2;

void bdf2_coeffs(double h, double* a1, double* a2, double* bden) {
  const double h2 = (1.0 - kGamma) * h;
  const double h1 = kGamma * h;
  const double rho = h2 / h1;
  const double denom = 1.0 + 2.0 * rho;
  *a1 = (1.0 + rho) * (1.0 + rho) / denom;
  *a2 = -rho * rho / denom;
  const double b0 = (1.0 + rho) / denom;
  *bden = b0 * h2;
}

struct CapState {
  std::vector<double> vprev, iprev, vpp;
  explicit CapState(int ne) : vprev(ne, 0.0), iprev(ne, 0.0), vpp(ne, 0.0) {}
};

double cap_volt(const spx_circuit* c, int e, const double* x) {
  return volt(x, c->elements[e].node[0]) - volt(x, c->elements[e].node[1]);
}

// InductorState::update (companion.rs:219-239) for rules 0-2 and
// update_trbdf2_intermediate (companion.rs:241-250) for rule 3, followed by
// the main loop's v_prev = v_gamma (solver.rs:199-203); vp = v_prev, ip =
// i_prev, iq = i_prev_prev.  The TR-BDF2 final update reads v_prev in both
// history terms, as the reference writes it.
void ind_update_rule(double L, double v, double h, int rule, double& vp, double& ip,
                     double& iq) {
  if (rule == 3) {
    const double h_gamma = kGamma * h;
    iq = ip;
    ip += h_gamma / (2.0 * L) * (v + vp);
    vp = v;
    return;
  }
  if (rule == 0) {
    ip += h / L * v;
  } else if (rule == 1) {
    ip += h / (2.0 * L) * (v + vp);
  } else {
    const double gamma = kGamma;
    const double alpha = (1.0 - gamma) / (gamma * (2.0 - gamma));
    const double di = h / L * ((1.0 + alpha) * v - (1.0 + 2.0 * alpha) * vp + alpha * vp);
    ip += di;
  }
  iq = ip;
  vp = v;
}

// CapacitorState::update (companion.rs:62-85), rule 0 BE, 1 trapezoidal,
// 2 TR-BDF2 final; rule 3 update_trbdf2_intermediate (companion.rs:87-98).
void cap_update_rule(const spx_circuit* c, const double* params, const double* x, double h,
                     int rule, CapState& cs) {
  for (int e = 0; e < c->num_elements; ++e) {
    if (c->elements[e].kind == SPX_EL_INDUCTOR) {
      ind_update_rule(P(params, c->elements[e].param[0]), cap_volt(c, e, x), h, rule,
                      cs.vprev[e], cs.iprev[e], cs.vpp[e]);
      continue;
    }
    if (c->elements[e].kind != SPX_EL_CAPACITOR) continue;
    const double C = P(params, c->elements[e].param[0]);
    const double v = cap_volt(c, e, x);
    double& vp = cs.vprev[e];
    double& ip = cs.iprev[e];
    double& vq = cs.vpp[e];
    if (rule == 3) {
      const double h_gamma = kGamma * h;
      vq = vp;
      ip = 2.0 * C / h_gamma * (v - vq) - ip;
      vp = v;
      continue;
    }
    if (rule == 0) {
      ip = C / h * (v - vp);
    } else if (rule == 1) {
      ip = 2.0 * C / h * (v - vp) - ip;
    } else {
      const double alpha = (1.0 - kGamma) / (kGamma * (2.0 - kGamma));
      ip = C / h * ((1.0 + alpha) * v - (1.0 + 2.0 * alpha) * vp + alpha * vq);
    }
    vq = vp;
    vp = v;
  }
}

}  // namespace

// One solve of a transient stage: linear circuits one LU solve (the
// reference's), nonlinear ones plain Newton from x.  Returns 1 converged,
// 0 failed, -1 singular; *it the iterations.
static int tran_solve(const spx_circuit* c, const double* params, const spx_nr_config* cfg,
                      bool nonlin, const OTran& tc, double* x, int32_t* it) {
  const int n = orc_circuit_size(c);
  std::vector<double> rhs(n), xn(n);
  if (!nonlin) {  // solve_transient's solve by size (transient/solver.rs:131-141)
    std::vector<double> zero(n, 0.0);
    *it = 1;
    if (use_sparse(n)) {
      std::vector<orc_sparse::Triplet> trips;
      stamp_impl(c, params, zero.data(), 0.0, 1.0, cfg->limiter, nullptr, rhs.data(), nullptr,
                 nullptr, &tc, &trips);
      const orc_sparse::Csc a = orc_sparse::from_triplets(n, trips);
      return orc_sparse::solve(orc_sparse::analyse(a), a, rhs.data(), x) ? -1 : 1;
    }
    std::vector<double> J((size_t)n * n);
    stamp_impl(c, params, zero.data(), 0.0, 1.0, cfg->limiter, J.data(), rhs.data(), nullptr,
               nullptr, &tc);
    return orc_lu_solve(n, J.data(), rhs.data(), x) ? -1 : 1;
  }
  const int st = newton_impl(c, params, cfg, 0.0, 1.0, x, xn.data(), it, &tc);
  std::memcpy(x, xn.data(), sizeof(double) * n);
  return st;
}

// Transient of one instance: DC operating point (orc_dcop semantics,
// capacitors open, PULSE at V1), capacitor states from it (v_prev = its
// voltage, i_prev = 0, solver.rs:49-55), then either fixed steps at
// t = k*h with o->method (solve_transient, solver.rs:41-249) or adaptive
// trapezoidal steps (solve_transient_adaptive, solver.rs:461-676).  Linear
// circuits take one solve per stage as the reference does; nonlinear ones
// plain Newton from the previous solution.  waves[k*np + p] = x[probes[p]]
// at point k (k = 0 the operating point); times[k] (adaptive).  The KCL
// residual is taken at the last point with the companion of its stage.
extern "C" int orc_transient_ex(const spx_circuit* c, const double* params,
                                const spx_nr_config* cfg, const orc_tran_opts* o, int32_t np,
                                const int32_t* probes, double* x, double* waves, double* times,
                                int32_t* num_points, int32_t* iterations, int32_t* flags,
                                double* residual, int32_t* fail_step, int32_t* total_steps,
                                int32_t* rejected_steps, double* min_step, double* max_step) {
  const int n = orc_circuit_size(c);
  int32_t it = 0, fl = 0;
  double res = 0.0;
  int status = orc_dcop(c, params, cfg, nullptr, x, &it, &fl, &res);
  fl &= ~SPX_FLAG_KCL_FAIL;
  int32_t total = it, fail = -1, points = 0, tried = 0, rejected = 0;
  double hmin_used = INFINITY, hmax_used = 0.0;
  // xr: x with each inductor's branch entry set to its current G v + I under
  // the companion of the latest solved stage (the transient stages leave
  // that entry 0)
  std::vector<double> xr(x, x + n);
  bool solved = false;  // a transient stage solved: x takes xr's branch currents
  std::vector<int> brow(c->num_elements, -1);  // inductor branch rows
  for (int e = 0; e < c->num_elements; ++e)
    if (c->elements[e].kind == SPX_EL_INDUCTOR) brow[e] = branch_row(c, e);
  auto probe = [&](int k, double t, const double* xp) {
    if (waves)
      for (int p = 0; p < np; ++p) waves[(size_t)k * np + p] = xp[probes[p]];
    if (times) times[k] = t;
  };
  if (status != SPX_CONVERGED) {
    fail = 0;
  } else {
    CapState cs(c->num_elements);
    for (int e = 0; e < c->num_elements; ++e) {
      if (c->elements[e].kind == SPX_EL_CAPACITOR) cs.vprev[e] = cap_volt(c, e, x);
      if (c->elements[e].kind == SPX_EL_INDUCTOR) {  // solver.rs:60-75
        cs.vprev[e] = cap_volt(c, e, x);
        cs.iprev[e] = x[brow[e]];
      }
    }
    probe(0, 0.0, x);
    points = 1;
    const bool nonlin = has_nonlinear(c);
    auto ctx = [&](double t, double h, int comp) {
      OTran tc{1, t, h, cs.vprev.data()};
      tc.comp = comp;
      tc.iprev = cs.iprev.data();
      tc.vpp = cs.vpp.data();
      return tc;
    };
    auto branch_currents = [&](const OTran& tc) {
      solved = true;
      std::memcpy(xr.data(), x, sizeof(double) * n);
      for (int e = 0; e < c->num_elements; ++e) {
        if (brow[e] < 0) continue;
        const double L = P(params, c->elements[e].param[0]);
        double geq, ieq;
        if (tc.comp == 1) {
          geq = tc.h / (2.0 * L);
          ieq = tc.iprev[e] + geq * tc.vprev[e];
        } else if (tc.comp == 2) {
          geq = tc.bden / L;
          ieq = tc.a1 * tc.iprev[e] + tc.a2 * tc.vpp[e];
        } else {
          geq = tc.h / L;
          ieq = tc.iprev[e];
        }
        xr[brow[e]] = geq * cap_volt(c, e, x) + ieq;
      }
    };
    if (!o->adaptive) {
      const double h = o->h;
      for (int k = 1; k <= o->steps; ++k) {
        const double t = k * h;
        int st;
        if (o->method == SPX_TRAN_TRBDF2) {
          const OTran t1 = ctx(t, kGamma * h, 1);
          st = tran_solve(c, params, cfg, nonlin, t1, x, &it);
          total += it;
          if (st == 1) {
            branch_currents(t1);
            cap_update_rule(c, params, x, h, 3, cs);
            OTran t2 = ctx(t, h, 2);
            bdf2_coeffs(h, &t2.a1, &t2.a2, &t2.bden);
            st = tran_solve(c, params, cfg, nonlin, t2, x, &it);
            total += it;
            if (st == 1 && k == o->steps) res = kcl_impl(c, params, cfg, x, 0.0, &t2);
            if (st == 1) branch_currents(t2);
          }
          if (st == 1) cap_update_rule(c, params, x, h, 2, cs);
        } else {
          const int comp = o->method == SPX_TRAN_TRAP ? 1 : 0;
          const OTran tc = ctx(t, h, comp);
          st = tran_solve(c, params, cfg, nonlin, tc, x, &it);
          total += it;
          if (st == 1) {
            if (k == o->steps) res = kcl_impl(c, params, cfg, x, 0.0, &tc);
            branch_currents(tc);
            cap_update_rule(c, params, x, h, comp, cs);
          }
        }
        if (st != 1) {
          status = st < 0 ? SPX_SINGULAR : SPX_FAILED;
          fail = k;
          break;
        }
        probe(k, t, xr.data());
        points = k + 1;
      }
    } else {
      // Newton (nonlinear circuits) starts from the latest solve's result,
      // accepted or rejected; the capacitor states move only on acceptance.
      double t = 0.0, h = o->h_init;
      while (t < o->tstop) {
        h = h < o->h_min ? o->h_min : (h > o->h_max ? o->h_max : h);
        if (t + h > o->tstop) h = o->tstop - t;
        const OTran tc = ctx(t, h, 1);  // sources at t, as the reference stamps them
        const int st = tran_solve(c, params, cfg, nonlin, tc, x, &it);
        total += it;
        ++tried;
        if (st != 1) {
          status = st < 0 ? SPX_SINGULAR : SPX_FAILED;
          fail = points;
          break;
        }
        branch_currents(tc);
        double max_lte = 0.0, max_ref = 0.0;
        for (int e = 0; e < c->num_elements; ++e) {
          if (c->elements[e].kind != SPX_EL_CAPACITOR) continue;
          const double C = P(params, c->elements[e].param[0]);
          const double v = cap_volt(c, e, x);
          const double i_trap = 2.0 * C / h * (v - cs.vprev[e]) - cs.iprev[e];
          const double i_be = C / h * (v - cs.vprev[e]);
          max_lte = std::fmax(max_lte, std::fabs(i_trap - i_be) / 3.0);
          max_ref = std::fmax(max_ref, std::fabs(v));
        }
        for (int e = 0; e < c->num_elements; ++e) {  // solver.rs:612-617
          if (c->elements[e].kind != SPX_EL_INDUCTOR) continue;
          const double L = P(params, c->elements[e].param[0]);
          const double v = cap_volt(c, e, x);
          const double di_trap = h / (2.0 * L) * (v + cs.vprev[e]);
          const double di_be = h / L * v;
          max_lte = std::fmax(max_lte, std::fabs(di_trap - di_be) / 3.0);
          max_ref = std::fmax(max_ref, std::fabs(cs.iprev[e]));
        }
        const double tol = std::fmax(o->abstol, o->reltol * max_ref);
        if (max_lte > tol && h > o->h_min) {
          ++rejected;
          const double factor = std::fmin(std::sqrt(tol / max_lte), 0.5);
          h *= std::fmax(factor, 0.1);
          continue;
        }
        if (points >= o->max_points) {
          status = SPX_FAILED;
          fail = points;
          break;
        }
        t += h;
        if (!(t < o->tstop)) res = kcl_impl(c, params, cfg, x, 0.0, &tc);
        cap_update_rule(c, params, x, h, 1, cs);
        hmin_used = std::fmin(hmin_used, h);
        hmax_used = std::fmax(hmax_used, h);
        probe(points, t, xr.data());
        ++points;
        if (max_lte < tol * 0.5 && h < o->h_max) {
          const double factor = std::fmin(std::sqrt(tol / std::fmax(max_lte, 1e-20)), 2.0);
          h *= std::fmin(factor, 1.5);
        }
      }
    }
  }
  if (status == SPX_CONVERGED && !(res <= 1.0)) fl |= SPX_FLAG_KCL_FAIL;
  if (status != SPX_SINGULAR && solved)
    for (int e = 0; e < c->num_elements; ++e)
      if (brow[e] >= 0) x[brow[e]] = xr[brow[e]];
  if (status == SPX_SINGULAR)
    for (int i = 0; i < n; ++i) x[i] = 0.0;
  if (iterations) *iterations = total;
  if (flags) *flags = fl;
  if (residual) *residual = res;
  if (fail_step) *fail_step = fail;
  if (num_points) *num_points = points;
  if (total_steps) *total_steps = tried;
  if (rejected_steps) *rejected_steps = rejected;
  if (min_step) *min_step = hmin_used;
  if (max_step) *max_step = hmax_used;
  return status;
}

// Backward-Euler fixed steps (the original entry; orc_transient_ex with
// method BE).
extern "C" int orc_transient(const spx_circuit* c, const double* params, const spx_nr_config* cfg,
                             double h, int32_t steps, int32_t np, const int32_t* probes,
                             double* x, double* waves, int32_t* iterations, int32_t* flags,
                             double* residual, int32_t* fail_step) {
  orc_tran_opts o{};
  o.method = SPX_TRAN_BE;
  o.h = h;
  o.steps = steps;
  return orc_transient_ex(c, params, cfg, &o, np, probes, x, waves, nullptr, nullptr, iterations,
                          flags, residual, fail_step, nullptr, nullptr, nullptr, nullptr);
}

extern "C" int orc_transient_ex_sweep(const spx_circuit* c, const double* params, int64_t B,
                                      const spx_nr_config* cfg, const orc_tran_opts* o,
                                      int32_t np, const int32_t* probes, double* x,
                                      int32_t* status, double* waves, double* times,
                                      int32_t* num_points, int32_t* iterations, int32_t* flags,
                                      double* residual, int32_t* fail_step, int32_t* total_steps,
                                      int32_t* rejected_steps, double* min_step,
                                      double* max_step, int32_t nthreads) {
  const int n = orc_circuit_size(c);
  const int P = c->num_params;
  const int64_t pts = o->adaptive ? o->max_points : o->steps + 1;
  run_pool(B, nthreads, [&](int64_t i) {
    status[i] = orc_transient_ex(
        c, params + i * P, cfg, o, np, probes, x + i * n, waves ? waves + i * pts * np : nullptr,
        times ? times + i * pts : nullptr, num_points ? num_points + i : nullptr, iterations + i,
        flags + i, residual + i, fail_step + i, total_steps ? total_steps + i : nullptr,
        rejected_steps ? rejected_steps + i : nullptr, min_step ? min_step + i : nullptr,
        max_step ? max_step + i : nullptr);
  });
  return SPX_OK;
}

extern "C" int orc_transient_sweep(const spx_circuit* c, const double* params, int64_t B,
                                   const spx_nr_config* cfg, double h, int32_t steps, int32_t np,
                                   const int32_t* probes, double* x, int32_t* status,
                                   double* waves, int32_t* iterations, int32_t* flags,
                                   double* residual, int32_t* fail_step, int32_t nthreads) {
  const int n = orc_circuit_size(c);
  const int P = c->num_params;
  run_pool(B, nthreads, [&](int64_t i) {
    status[i] = orc_transient(c, params + i * P, cfg, h, steps, np, probes, x + i * n,
                              waves ? waves + i * (int64_t)(steps + 1) * np : nullptr,
                              iterations + i, flags + i, residual + i, fail_step + i);
  });
  return SPX_OK;
}
