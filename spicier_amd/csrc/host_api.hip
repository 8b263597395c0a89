// host_api.hip — the C ABI of libspicier_hip.so (include/spicier_hip.h).
//
// A handle owns one HIP stream on one device and a grow-only set of device
// workspaces; calls on a handle are serialized by its mutex.  Host-pointer
// entry points stage through the workspaces and synchronize; "_device"
// entry points only enqueue on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spicier_hip.h"
#include "handle.h"
#include "program.h"
#include "gmres.h"
#include "sparse_plan.h"

namespace spx {
hipError_t launch_newton(const DevProgram& g, const DevNrCfg& cfg, const double* params,
                         const double* x0, int64_t B, const DevOut& out, hipStream_t st);
hipError_t launch_sample(const spx_param_spec* d_specs, int P, int gen, uint64_t seed, int64_t B,
                         int64_t offset, int nvar, double* out, hipStream_t st);
hipError_t launch_stats(const double* v, const uint8_t* mask, const int* status, int64_t B, int n,
                        double* part, double* out, hipStream_t st);
int64_t stats_num_blocks(int64_t B);
hipError_t launch_solve_batch(const double* M, const double* rhs, int n, int64_t B, double* x,
                              uint8_t* sing, double* work, int64_t chunk, hipStream_t st);
hipError_t sparse_solve_shape(const DevSparse& sp, int n, int64_t B, int* grid, int64_t* slot,
                              size_t* smem, int* lu_in_lds);
hipError_t launch_sparse_solve(const DevSparse& sp, int n, int num_pos, const int* pos_off,
                               const int* pos_trip, int64_t nnz, const double* values,
                               const double* rhs, int64_t B, double* x, uint8_t* singular,
                               double* ws, int grid, int64_t slot, size_t smem, int lu_in_lds,
                               hipStream_t st);
hipError_t launch_transient_8(const DevProgram&, const DevNrCfg&, const DevTran&, const double*,
                              int64_t, const DevOut&, const DevTranOut&, hipStream_t, int);
hipError_t launch_transient_16(const DevProgram&, const DevNrCfg&, const DevTran&, const double*,
                               int64_t, const DevOut&, const DevTranOut&, hipStream_t, int);
hipError_t launch_transient_32(const DevProgram&, const DevNrCfg&, const DevTran&, const double*,
                               int64_t, const DevOut&, const DevTranOut&, hipStream_t, int);
hipError_t launch_transient_64(const DevProgram&, const DevNrCfg&, const DevTran&, const double*,
                               int64_t, const DevOut&, const DevTranOut&, hipStream_t, int);
hipError_t sparse_transient_shape(const DevProgram& g, const DevSparse& sp, int64_t B, int* grid,
                                  int64_t* slot, size_t* smem, int* lu_in_lds, int threads,
                                  int force_lu);
hipError_t launch_transient_sparse(const DevProgram& g, const DevSparse& sp, const DevNrCfg& cfg,
                                   const DevTran& tr, const double* params, int64_t B,
                                   const DevOut& out, const DevTranOut& tout, double* ws,
                                   int grid, int64_t slot, size_t smem, int lu_in_lds,
                                   int threads, hipStream_t st,
                                   int* next = nullptr);
hipError_t lanes_transient_shape(const DevProgram& g, const DevSparse& sp, int64_t B, int* grid,
                                 int64_t* slot);
hipError_t launch_transient_lanes(const DevProgram& g, const DevSparse& sp, const DevNrCfg& cfg,
                                  const DevTran& tr, const double* params, int64_t B,
                                  const DevOut& out, const DevTranOut& tout, double* ws, int grid,
                                  int64_t slot, hipStream_t st);
hipError_t sparse_launch_shape(const DevProgram& g, const DevSparse& sp, int64_t B, int ir,
                               int* grid, int64_t* slot, size_t* smem, int* lu_in_lds);
hipError_t launch_newton_sparse(const DevProgram& g, const DevSparse& sp, const DevNrCfg& cfg,
                                const double* params, const double* x0, int64_t B,
                                const DevOut& out, double* ws, int grid, int64_t slot,
                                size_t smem, int lu_in_lds, int ir, int escalated_only,
                                hipStream_t st);
size_t solve_generic_workspace(int n, int64_t chunk);
hipError_t launch_factor_batch(const double* M, int n, int64_t B, double* F, int32_t* piv,
                               uint8_t* sing, hipStream_t st);
int noalloc_scratch_init(int device);
}  // namespace spx

namespace spx {
std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}
}  // namespace spx

using namespace spx;

namespace {

// Unknown counts: the register-LU kernels take n <= kDenseMaxN, the
// shared-symbolic sparse kernel n <= kSparseMaxN.
constexpr int kDenseMaxN = 64;
constexpr int kSparseMaxN = 4096;

// FP32+IR form of the dense transient kernels (phase bit 4); the sparse and
// lane kernels factor in FP64 whatever the precision.
int ir_phase(const DevNrCfg& dc) { return dc.precision == SPX_PREC_F32_IR ? 4 : 0; }

// ------------------------------------------------------------- circuit compile
constexpr int kMaxPwlPoints = 4096;

// Element kinds with a branch current unknown (numbered in element order).
inline bool has_branch(int kind) {
  return kind == SPX_EL_VSOURCE || kind == SPX_EL_VPULSE || kind == SPX_EL_VSIN ||
         kind == SPX_EL_VPWL || kind == SPX_EL_INDUCTOR || kind == SPX_EL_VCVS ||
         kind == SPX_EL_CCVS;
}

struct HostProgram {
  int n = 0, nn = 0, P = 0, num_elems = 0, num_vals = 2, has_nonlinear = 0, has_inductor = 0;
  bool want_sparse = false;  // build the sparse plan even when n fits the register LU
  std::vector<int> elems, order, pos_rc, pos_off, row_ptr, contrib, rhs_off, rhs_contrib;
  std::vector<char> weak_diag;  // nodes with no diagonal entry besides gmin
  int max_row_nnz = 0;
};

// transient: capacitors get companion slots (G = C/h, I = G v_prev) instead
// of being left open.
int compile_circuit(const spx_circuit* c, HostProgram& hp, int max_n, bool transient = false) {
  if (!c || c->num_nodes < 0 || c->num_elements < 0 || c->num_params < 0 ||
      (c->num_elements > 0 && !c->elements))
    return fail(SPX_ERR_INVALID_ARGUMENT, "invalid circuit description");
  const int nn = c->num_nodes, P = c->num_params;
  int nv = 0;
  for (int e = 0; e < c->num_elements; ++e)
    nv += has_branch(c->elements[e].kind);
  const int n = nn + nv;
  if (n < 1) return fail(SPX_ERR_INVALID_DIMENSION, "circuit has no unknowns");
  if (n > max_n)
    return fail(SPX_ERR_INVALID_DIMENSION,
                "circuit has " + std::to_string(n) + " unknowns; the sweep path takes <= " +
                    std::to_string(max_n));
  hp.n = n;
  hp.nn = nn;
  hp.P = P;
  hp.num_elems = c->num_elements;
  std::map<std::pair<int, int>, std::vector<int>> pos;
  std::vector<std::vector<int>> rhs(n);
  auto add = [&](int r, int col, int vid, bool neg) { pos[{r, col}].push_back(vid << 1 | neg); };
  auto cond = [&](int i, int j, int vid) {  // mna.rs:87-98
    if (i >= 0) add(i, i, vid, false);
    if (j >= 0) add(j, j, vid, false);
    if (i >= 0 && j >= 0) { add(i, j, vid, true); add(j, i, vid, true); }
  };
  auto cur = [&](int i, int j, int vid) {  // mna.rs:103-115
    if (i >= 0) rhs[i].push_back(vid << 1 | 1);
    if (j >= 0) rhs[j].push_back(vid << 1);
  };
  int branch = 0;
  std::vector<int> brow(c->num_elements, -1);  // branch row of each branch element
  std::set<int> dc_zero_vid;                   // value slots that are 0 at DC
  for (int e = 0; e < c->num_elements; ++e) {
    const spx_element& el = c->elements[e];
    int u[4];
    for (int k = 0; k < 4; ++k) {
      if (el.node[k] < 0 || el.node[k] > nn)
        return fail(SPX_ERR_INVALID_ARGUMENT, "element " + std::to_string(e) +
                                                  " references node " + std::to_string(el.node[k]));
      u[k] = el.node[k] - 1;
    }
    if (has_branch(el.kind)) brow[e] = nn + branch;  // the case below takes branch++
    int cb = -1;  // controlling branch row (CCCS / CCVS)
    if (el.kind == SPX_EL_CCCS || el.kind == SPX_EL_CCVS) {
      if (el.ctrl < 0 || el.ctrl >= e || brow[el.ctrl] < 0)
        return fail(SPX_ERR_INVALID_ARGUMENT,
                    "element " + std::to_string(e) +
                        " needs an earlier controlling element with a branch current");
      cb = brow[el.ctrl];
    }
    int need = 0;
    switch (el.kind) {
      case SPX_EL_RESISTOR: case SPX_EL_CAPACITOR: case SPX_EL_VSOURCE: case SPX_EL_ISOURCE:
      case SPX_EL_INDUCTOR: case SPX_EL_VCVS: case SPX_EL_VCCS: case SPX_EL_CCCS:
      case SPX_EL_CCVS:
        need = 1; break;
      case SPX_EL_DIODE: need = 2; break;
      case SPX_EL_NMOS: case SPX_EL_PMOS: need = 5; break;
      case SPX_EL_VPULSE: need = 7; break;
      case SPX_EL_VSIN: need = 6; break;
      case SPX_EL_VPWL:
        if (el.ctrl < 1 || el.ctrl > kMaxPwlPoints || el.param[0] < 0 ||
            (int64_t)el.param[0] + 2 * (int64_t)el.ctrl > P)
          return fail(SPX_ERR_INVALID_ARGUMENT,
                      "element " + std::to_string(e) + ": PWL needs 1 to " +
                          std::to_string(kMaxPwlPoints) + " points in parameter slots");
        need = 1;
        break;
      default:
        return fail(SPX_ERR_INVALID_ARGUMENT, "element " + std::to_string(e) + " has unknown kind");
    }
    for (int k = 0; k < need; ++k)
      if (el.param[k] < 0 || el.param[k] >= P)
        return fail(SPX_ERR_INVALID_ARGUMENT, "element " + std::to_string(e) +
                                                  " parameter slot out of range");
    int vid = -1;
    switch (el.kind) {
      case SPX_EL_RESISTOR:
        vid = hp.num_vals++;
        cond(u[0], u[1], vid);
        break;
      case SPX_EL_CAPACITOR:  // open at DC; BE companion in a transient (companion.rs:41-48)
        if (transient) {
          vid = hp.num_vals;
          dc_zero_vid.insert(vid);
          hp.num_vals += 2;
          cond(u[0], u[1], vid);
          cur(u[1], u[0], vid + 1);
        }
        break;
      case SPX_EL_VSOURCE:
      case SPX_EL_VPULSE:
      case SPX_EL_VSIN:
      case SPX_EL_VPWL: {  // mna.rs:124-145
        if (el.kind == SPX_EL_VPWL) u[2] = el.ctrl;  // the record carries the point count
        vid = hp.num_vals++;
        const int row = nn + branch++;
        if (u[0] >= 0) { add(u[0], row, kValOne, false); add(row, u[0], kValOne, false); }
        if (u[1] >= 0) { add(u[1], row, kValOne, true); add(row, u[1], kValOne, true); }
        rhs[row].clear();  // rhs[row] = voltage (assignment)
        rhs[row].push_back(vid << 1);
        break;
      }
      case SPX_EL_INDUCTOR: {
        // KCL: +i at node 0, -i at node 1, as a voltage source (mna.rs:124-145).
        // Branch row: c (v0 - v1) + d i = r with (c, d, r) = (1, 0, 0) at DC
        // (the reference's 0 V source, passive.rs:318-325) and (-G, 1, I) in a
        // transient step, which is the reference's conductance companion
        // (G, I from node 0 to node 1) with i kept as an unknown.
        vid = hp.num_vals;
        hp.num_vals += 3;
        const int row = nn + branch++;
        u[2] = row;  // the element record carries its branch row
        if (u[0] >= 0) { add(u[0], row, kValOne, false); add(row, u[0], vid, false); }
        if (u[1] >= 0) { add(u[1], row, kValOne, true); add(row, u[1], vid, true); }
        if (transient) {
          add(row, row, vid + 1, false);
          rhs[row].push_back((vid + 2) << 1);
        }
        hp.has_inductor = 1;
        break;
      }
      case SPX_EL_VCVS: {  // controlled.rs:58-90, in its push order
        vid = hp.num_vals++;
        const int row = nn + branch++;
        if (u[0] >= 0) add(u[0], row, kValOne, false);
        if (u[1] >= 0) add(u[1], row, kValOne, true);
        if (u[0] >= 0) add(row, u[0], kValOne, false);
        if (u[1] >= 0) add(row, u[1], kValOne, true);
        if (u[2] >= 0) add(row, u[2], vid, true);
        if (u[3] >= 0) add(row, u[3], vid, false);
        break;
      }
      case SPX_EL_VCCS:  // controlled.rs:172-201
        vid = hp.num_vals++;
        if (u[0] >= 0) {
          if (u[2] >= 0) add(u[0], u[2], vid, true);
          if (u[3] >= 0) add(u[0], u[3], vid, false);
        }
        if (u[1] >= 0) {
          if (u[2] >= 0) add(u[1], u[2], vid, false);
          if (u[3] >= 0) add(u[1], u[3], vid, true);
        }
        break;
      case SPX_EL_CCCS:  // controlled.rs:267-285
        vid = hp.num_vals++;
        if (u[0] >= 0) add(u[0], cb, vid, false);
        if (u[1] >= 0) add(u[1], cb, vid, true);
        break;
      case SPX_EL_CCVS: {  // controlled.rs:351-375
        vid = hp.num_vals++;
        const int row = nn + branch++;
        if (u[0] >= 0) add(u[0], row, kValOne, false);
        if (u[1] >= 0) add(u[1], row, kValOne, true);
        if (u[0] >= 0) add(row, u[0], kValOne, false);
        if (u[1] >= 0) add(row, u[1], kValOne, true);
        add(row, cb, vid, true);
        break;
      }
      case SPX_EL_ISOURCE:
        vid = hp.num_vals++;
        cur(u[0], u[1], vid);
        break;
      case SPX_EL_DIODE:
        vid = hp.num_vals;
        hp.num_vals += 2;
        cond(u[0], u[1], vid);
        cur(u[0], u[1], vid + 1);
        hp.has_nonlinear = 1;
        break;
      case SPX_EL_NMOS:
      case SPX_EL_PMOS: {  // level1.rs:209-247
        vid = hp.num_vals;
        hp.num_vals += 3;
        const int d = u[0], g = u[1], s = u[2];
        cond(d, s, vid);
        if (d >= 0) {
          if (g >= 0) add(d, g, vid + 1, false);
          if (s >= 0) add(d, s, vid + 1, true);
        }
        if (s >= 0) {
          if (g >= 0) add(s, g, vid + 1, true);
          add(s, s, vid + 1, false);
        }
        cur(d, s, vid + 2);
        hp.has_nonlinear = 1;
        break;
      }
    }
    int rec[kElemInts] = {el.kind, u[0], u[1], u[2], el.param[0], el.param[1], el.param[2],
                          el.param[3], el.param[4], el.param[5], el.param[6], vid};
    hp.elems.insert(hp.elems.end(), rec, rec + kElemInts);
  }
  // weak: a node diagonal with nothing but capacitor companions (zero at the
  // operating point) and, below, gmin
  hp.weak_diag.assign(n, 0);
  for (int i = 0; i < nn; ++i) {
    auto it = pos.find({i, i});
    bool weak = true;
    if (it != pos.end())
      for (int t : it->second) weak = weak && dc_zero_vid.count(t >> 1);
    hp.weak_diag[i] = weak;
  }
  for (int i = 0; i < nn; ++i) add(i, i, kValGmin, false);  // GminStamper, newton.rs:393-398
  std::vector<int> rownnz(n, 0);
  hp.pos_off.push_back(0);
  for (auto& kv : pos) {
    hp.pos_rc.push_back(kv.first.first | (kv.first.second << 16));
    hp.contrib.insert(hp.contrib.end(), kv.second.begin(), kv.second.end());
    hp.pos_off.push_back((int)hp.contrib.size());
    rownnz[kv.first.first]++;
  }
  hp.max_row_nnz = *std::max_element(rownnz.begin(), rownnz.end());
  hp.row_ptr.assign(1, 0);
  for (int r = 0; r < n; ++r) hp.row_ptr.push_back(hp.row_ptr.back() + rownnz[r]);
  hp.rhs_off.push_back(0);
  for (int r = 0; r < n; ++r) {
    hp.rhs_contrib.insert(hp.rhs_contrib.end(), rhs[r].begin(), rhs[r].end());
    hp.rhs_off.push_back((int)hp.rhs_contrib.size());
  }
  // evaluation order: nonlinear devices first, grouped by kind
  auto rank = [](int k) {
    switch (k) {
      case SPX_EL_NMOS: case SPX_EL_PMOS: return 0;
      case SPX_EL_DIODE: return 1;
      case SPX_EL_RESISTOR: return 2;
      default: return 3;
    }
  };
  hp.order.resize(c->num_elements);
  for (int e = 0; e < c->num_elements; ++e) hp.order[e] = e;
  std::stable_sort(hp.order.begin(), hp.order.end(), [&](int a, int b) {
    return rank(c->elements[a].kind) < rank(c->elements[b].kind);
  });
  return SPX_OK;
}

inline bool sp_always(const HostProgram& hp) { return hp.want_sparse; }

// Uploads the program into the handle's "program<slot>" workspace, and for n
// past the register LU the shared-symbolic LU plan into "sparse_plan<slot>".
int upload_program(spx_handle* h, const HostProgram& hp, DevProgram& g, DevSparse* sp,
                   const std::string& slot = "") {
  auto& ps = h->programs[slot];
  const std::string wprog = "program" + slot, wplan = "sparse_plan" + slot;
  std::vector<int> blob;
  auto put = [&](const std::vector<int>& v) {
    size_t off = blob.size();
    blob.insert(blob.end(), v.begin(), v.end());
    while (blob.size() % 4) blob.push_back(0);
    return off;
  };
  size_t o_el = put(hp.elems), o_or = put(hp.order), o_rc = put(hp.pos_rc),
         o_po = put(hp.pos_off), o_co = put(hp.contrib), o_ro = put(hp.rhs_off),
         o_rh = put(hp.rhs_contrib), o_rp = put(hp.row_ptr);
  // lane-major element table and gather program of one 64-lane wave
  const int rounds = (hp.num_elems + 63) / 64;
  std::vector<int> el_lm((size_t)rounds * kElemLm * 64, 0);
  for (int t = 0; t < rounds; ++t)
    for (int l = 0; l < 64; ++l) {
      const int i = t * 64 + l;
      int* dst = el_lm.data() + (size_t)t * kElemLm * 64 + l;
      if (i >= hp.num_elems) {
        dst[0] = -1;
        continue;
      }
      const int e = hp.order[i];
      for (int f = 0; f < kElemInts; ++f) dst[f * 64] = hp.elems[(size_t)e * kElemInts + f];
      dst[kElemInts * 64] = e;
    }
  const int num_pos = (int)hp.pos_rc.size();
  std::vector<std::vector<int>> lane_ops(64);
  bool fits = hp.num_vals <= 0x3fff && num_pos < 0x8000 && hp.n < 0x8000;
  auto emit = [&](int lane, const int* c0, const int* c1, int dst) {
    auto& ops = lane_ops[lane];
    if (c0 == c1) {  // an empty sum: +1 - 1 leaves +0.0, as the loop over nothing
      ops.push_back(kValOne);
      ops.push_back(kValOne | 1 << 14 | 1 << 15 | dst << 16);
      return;
    }
    for (const int* c = c0; c < c1; ++c) {
      const int slot = *c >> 1, neg = *c & 1;
      fits = fits && slot <= 0x3fff;
      ops.push_back(slot | neg << 14 | (c + 1 == c1 ? 1 << 15 : 0) | dst << 16);
    }
  };
  for (int q = 0; q < num_pos && fits; ++q)
    emit(q % 64, hp.contrib.data() + hp.pos_off[q], hp.contrib.data() + hp.pos_off[q + 1], q);
  for (int r = 0; r < hp.n && fits; ++r)
    emit(r % 64, hp.rhs_contrib.data() + hp.rhs_off[r], hp.rhs_contrib.data() + hp.rhs_off[r + 1],
         0x8000 | r);
  size_t gk = 0;
  for (auto& v : lane_ops) gk = std::max(gk, v.size());
  if (!fits) gk = 0;
  std::vector<int> gath(std::max<size_t>(1, gk * 64), -1);
  for (int l = 0; l < 64 && fits; ++l)
    for (size_t k = 0; k < lane_ops[l].size(); ++k) gath[k * 64 + l] = lane_ops[l][k];
  std::vector<int> rowmask(2 * (size_t)hp.n, 0);
  for (int r = 0; r < hp.n && r < 64; ++r)
    for (int q = hp.row_ptr[r]; q < hp.row_ptr[r + 1]; ++q) {
      const int col = hp.pos_rc[q] >> 16;
      if (col < 64) rowmask[2 * r + (col >> 5)] |= (int)(1u << (col & 31));
    }
  size_t o_lm = put(el_lm), o_ga = put(gath), o_rm = put(rowmask);
  // A sweep loop over one circuit uploads its program once: the upload (and
  // the stream wait its pageable source needs) is skipped while the program
  // in the workspace is byte-identical.
  const size_t bytes = blob.size() * sizeof(int) + 16;
  auto it = h->ws.find(wprog);
  const bool cached = it != h->ws.end() && it->second.second >= bytes && blob == ps.blob;
  void* d = nullptr;
  int rc = workspace(h, wprog.c_str(), bytes, &d);
  if (rc) return rc;
  if (!cached) {
    ps.blob.clear();
    ps.has_sparse = false;
    SPX_HIP(hipMemcpyAsync(d, blob.data(), blob.size() * sizeof(int), hipMemcpyHostToDevice,
                           h->stream),
            SPX_ERR_TRANSFER);
    // the host blob dies at return: wait for the copy
    SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
    ps.blob = std::move(blob);
  }
  if (hp.n > kDenseMaxN || sp_always(hp)) {
    if (!ps.has_sparse) {
      const char* why = build_sparse_plan(hp.n, hp.nn, hp.pos_rc, ps.plan, &hp.weak_diag);
      if (*why) return fail(SPX_ERR_INVALID_DIMENSION, why);
      void* ds = nullptr;
      const auto& pb = ps.plan.blob;
      if ((rc = workspace(h, wplan.c_str(), pb.size() * sizeof(int), &ds))) return rc;
      SPX_HIP(hipMemcpyAsync(ds, pb.data(), pb.size() * sizeof(int), hipMemcpyHostToDevice,
                             h->stream),
              SPX_ERR_TRANSFER);
      SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
      ps.has_sparse = true;
    }
    void* ds = nullptr;
    if ((rc = workspace(h, wplan.c_str(), ps.plan.blob.size() * sizeof(int), &ds))) return rc;
    if (sp) *sp = device_sparse(ps.plan, static_cast<const int*>(ds));
  }
  const int* base = static_cast<const int*>(d);
  g.n = hp.n;
  g.num_nodes = hp.nn;
  g.P = hp.P;
  g.num_elems = hp.num_elems;
  g.num_vals = hp.num_vals;
  g.num_pos = (int)hp.pos_rc.size();
  g.has_nonlinear = hp.has_nonlinear;
  g.has_inductor = hp.has_inductor;
  g.max_row_nnz = hp.max_row_nnz;
  g.elems = base + o_el;
  g.order = base + o_or;
  g.pos_rc = base + o_rc;
  g.pos_off = base + o_po;
  g.row_ptr = base + o_rp;
  g.contrib = base + o_co;
  g.rhs_off = base + o_ro;
  g.rhs_contrib = base + o_rh;
  g.el_rounds = rounds;
  g.el_lm = base + o_lm;
  g.gath_k = (int)gk;
  g.gath = base + o_ga;
  g.rowmask = base + o_rm;
  return SPX_OK;
}

int count_varied(const spx_param_spec* ps, int P) {
  int v = 0;
  for (int p = 0; p < P; ++p) v += ps[p].kind != SPX_P_CONST;
  return v;
}

int check_spec(const spx_sweep_spec* spec, int P) {
  if (!spec) return fail(SPX_ERR_INVALID_ARGUMENT, "null sweep spec");
  if (spec->batch < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative batch");
  if (spec->batch > 0x7fffffffLL)
    return fail(SPX_ERR_BATCH_TOO_LARGE, "batch " + std::to_string(spec->batch) + " exceeds " +
                                             std::to_string(0x7fffffffLL));
  if (P > 0 && !spec->params && spec->generator != SPX_GEN_EXPLICIT)
    return fail(SPX_ERR_INVALID_ARGUMENT, "null parameter specs");
  if (spec->generator < SPX_GEN_HASH || spec->generator > SPX_GEN_EXPLICIT)
    return fail(SPX_ERR_INVALID_ARGUMENT, "unknown generator");
  if (spec->generator == SPX_GEN_EXPLICIT && P > 0 && spec->batch > 0 && !spec->explicit_params)
    return fail(SPX_ERR_INVALID_ARGUMENT, "explicit generator without parameters");
  if (spec->generator == SPX_GEN_LCG)
    for (int p = 0; p < P; ++p)
      if (spec->params[p].kind != SPX_P_CONST && spec->params[p].kind != SPX_P_GAUSS_REL)
        return fail(SPX_ERR_INVALID_ARGUMENT,
                    "the LCG generator draws ParameterVariation (Gaussian-relative) slots only");
  if (spec->generator == SPX_GEN_CORNER && P > 0) {
    int nv = count_varied(spec->params, P);
    if (nv >= 62 || spec->batch != (int64_t(1) << nv))
      return fail(SPX_ERR_INVALID_DIMENSION, "corner sweep needs batch == 2^(varied slots)");
  }
  return SPX_OK;
}

// Enqueue parameter generation into d_out [B][P].
int sample_device(spx_handle* h, const spx_sweep_spec* spec, int P, double* d_out) {
  const int64_t B = spec->batch;
  if (B == 0 || P == 0) return SPX_OK;
  if (spec->generator == SPX_GEN_EXPLICIT) {
    SPX_HIP(hipMemcpyAsync(d_out, spec->explicit_params, sizeof(double) * B * P,
                           hipMemcpyDeviceToDevice, h->stream),
            SPX_ERR_TRANSFER);
    return SPX_OK;
  }
  void* d_specs = nullptr;
  int rc = workspace(h, "specs", sizeof(spx_param_spec) * P, &d_specs);
  if (rc) return rc;
  SPX_HIP(hipMemcpyAsync(d_specs, spec->params, sizeof(spx_param_spec) * P,
                         hipMemcpyHostToDevice, h->stream),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  SPX_HIP(launch_sample(static_cast<spx_param_spec*>(d_specs), P, spec->generator, spec->seed, B,
                        spec->sweep_offset, count_varied(spec->params, P), d_out, h->stream),
          SPX_ERR_BACKEND);
  return SPX_OK;
}

DevNrCfg to_dev(const spx_nr_config* c) {
  DevNrCfg d;
  d.v_abstol = c->v_abstol;
  d.v_reltol = c->v_reltol;
  d.i_abstol = c->i_abstol;
  d.max_iterations = c->max_iterations;
  d.gmin_stepping = c->gmin_stepping;
  d.gmin_initial = c->gmin_initial;
  d.gmin_final = c->gmin_final;
  d.gmin_factor = c->gmin_factor;
  d.gmin_attempts = c->gmin_max_attempts;
  d.source_stepping = c->source_stepping;
  d.src_initial = c->src_initial;
  d.src_step = c->src_step;
  d.src_min_step = c->src_min_step;
  d.src_attempts = c->src_max_attempts;
  d.precision = c->precision;
  d.limiter = c->limiter;
  d.warm = c->warm_start;
  d.ir_max_steps = c->ir_max_steps;
  d.ir_tol = c->ir_tol;
  d.kcl_reltol = c->kcl_reltol;
  d.kcl_abstol = c->kcl_abstol;
  d.kcl_vntol = c->kcl_vntol;
  return d;
}

int check_cfg(const spx_nr_config* c) {
  if (!c) return fail(SPX_ERR_INVALID_ARGUMENT, "null Newton configuration");
  if (c->max_iterations < 1 || c->max_iterations > 100000)
    return fail(SPX_ERR_INVALID_ARGUMENT, "max_iterations must be 1..100000");
  if (c->precision != SPX_PREC_F64 && c->precision != SPX_PREC_F32_IR)
    return fail(SPX_ERR_INVALID_ARGUMENT, "unknown precision");
  if (c->limiter < SPX_LIMIT_REFERENCE || c->limiter > SPX_LIMIT_PNJLIM)
    return fail(SPX_ERR_INVALID_ARGUMENT, "unknown limiter");
  if (c->warm_start < SPX_WARM_ZERO || c->warm_start > SPX_WARM_SURROGATE)
    return fail(SPX_ERR_INVALID_ARGUMENT, "unknown warm start");
  if (c->precision == SPX_PREC_F32_IR && (c->ir_max_steps < 1 || c->ir_max_steps > 1000))
    return fail(SPX_ERR_INVALID_ARGUMENT, "ir_max_steps must be 1..1000");
  if (c->gmin_stepping && !(c->gmin_factor > 1.0 && c->gmin_final > 0.0))
    return fail(SPX_ERR_INVALID_ARGUMENT, "Gmin stepping needs factor > 1 and final > 0");
  if (c->source_stepping && !(c->src_min_step > 0.0 && c->src_step > 0.0))
    return fail(SPX_ERR_INVALID_ARGUMENT, "source stepping needs positive steps");
  if ((c->gmin_stepping && (c->gmin_max_attempts < 1 || c->gmin_max_attempts > 1000)) ||
      (c->source_stepping && (c->src_max_attempts < 1 || c->src_max_attempts > 1000)))
    return fail(SPX_ERR_INVALID_ARGUMENT, "max attempts per stepping level must be 1..1000");
  return SPX_OK;
}

int sweep_device_locked(spx_handle* h, const spx_circuit* ckt, const spx_sweep_spec* spec,
                        const spx_nr_config* cfg, const double* d_x0, spx_sweep_result* out) {
  int rc;
  if (!out || !out->x) return fail(SPX_ERR_INVALID_ARGUMENT, "result.x is required");
  if ((rc = check_cfg(cfg))) return rc;
  HostProgram hp;
  if ((rc = compile_circuit(ckt, hp, kSparseMaxN))) return rc;
  if ((rc = check_spec(spec, hp.P))) return rc;
  const bool sparse = hp.n > kDenseMaxN;
  if (cfg->warm_start == SPX_WARM_GIVEN && !d_x0)
    return fail(SPX_ERR_INVALID_ARGUMENT, "warm start given without x0");
  if (cfg->warm_start == SPX_WARM_SURROGATE && !h->surrogate)
    return fail(SPX_ERR_INVALID_ARGUMENT, "surrogate warm start without an attached surrogate");
  const int64_t B = spec->batch;
  if (B == 0) return SPX_OK;
  DevProgram g;
  DevSparse sp{};
  if ((rc = upload_program(h, hp, g, &sp))) return rc;
  // Parameters the kernels read: the caller's explicit table itself when no
  // copy was asked for, else sampled (or copied) into result.params or a
  // workspace.
  const double* params = out->params;
  if (!out->params && spec->generator == SPX_GEN_EXPLICIT) {
    params = spec->explicit_params;
  } else {
    double* dst = out->params;
    if (!dst) {
      void* p = nullptr;
      if ((rc = workspace(h, "params", sizeof(double) * std::max<int64_t>(1, B * hp.P), &p)))
        return rc;
      dst = static_cast<double*>(p);
    }
    if ((rc = sample_device(h, spec, hp.P, dst))) return rc;
    params = dst;
  }
  int* status = out->status;
  if (!status) {  // the dense kernels' second pass and the statistics read it
    void* p = nullptr;
    if ((rc = workspace(h, "status", sizeof(int) * B, &p))) return rc;
    status = static_cast<int*>(p);
  }
  DevOut o{out->x, status, out->iterations, out->flags, out->residual, out->factorizations,
           out->refinements};
  DevNrCfg dc = to_dev(cfg);
  const double* x0 = cfg->warm_start == SPX_WARM_GIVEN ? d_x0 : nullptr;
  if (cfg->warm_start == SPX_WARM_SURROGATE) {
    void* p = nullptr;
    if ((rc = workspace(h, "x0_surrogate", sizeof(double) * B * hp.n, &p))) return rc;
    if ((rc = surrogate_predict_locked(h->surrogate, params, B, hp.n, hp.P,
                                       static_cast<double*>(p))))
      return rc;
    x0 = static_cast<double*>(p);
  }
  if (sparse) {
    // FP64: one pass.  FP32+IR: the refinement form, then the FP64 kernel on
    // the instances it handed over (the same workspace, used in turn).
    const bool ir = cfg->precision == SPX_PREC_F32_IR;
    int grid = 0, lu_lds = 0, grid32 = 0, lu_lds32 = 0;
    int64_t slot = 0, slot32 = 0;
    size_t smem = 0, smem32 = 0;
    SPX_HIP(sparse_launch_shape(g, sp, B, 0, &grid, &slot, &smem, &lu_lds), SPX_ERR_BACKEND);
    if (ir)
      SPX_HIP(sparse_launch_shape(g, sp, B, 1, &grid32, &slot32, &smem32, &lu_lds32),
              SPX_ERR_BACKEND);
    const size_t ws_doubles = std::max((size_t)grid * slot, (size_t)grid32 * slot32);
    void* ws = nullptr;
    if ((rc = workspace(h, "sparse_ws", sizeof(double) * ws_doubles, &ws))) return rc;
    if (ir)
      SPX_HIP(launch_newton_sparse(g, sp, dc, params, x0, B, o, static_cast<double*>(ws), grid32,
                                   slot32, smem32, lu_lds32, 1, 0, h->stream),
              SPX_ERR_BACKEND);
    SPX_HIP(launch_newton_sparse(g, sp, dc, params, x0, B, o, static_cast<double*>(ws), grid,
                                 slot, smem, lu_lds, 0, ir ? 1 : 0, h->stream),
            SPX_ERR_BACKEND);
  } else {
    SPX_HIP(launch_newton(g, dc, params, x0, B, o, h->stream), SPX_ERR_BACKEND);
  }
  if (out->node_stats) {
    void* part = nullptr;
    if ((rc = workspace(h, "stats_part", sizeof(double) * 5 * hp.n * stats_num_blocks(B), &part)))
      return rc;
    SPX_HIP(launch_stats(out->x, nullptr, status, B, hp.n, static_cast<double*>(part),
                         out->node_stats, h->stream),
            SPX_ERR_BACKEND);
  }
  return SPX_OK;
}

// Batched sparse solve on device pointers (caller holds the mutex).
int sparse_solve_locked(spx_handle* h, int64_t n, int64_t nnz, const int32_t* rows,
                        const int32_t* cols, const double* d_values, const double* d_rhs,
                        int64_t B, double* d_x, uint8_t* d_sing) {
  if (n < 1 || n > kSparseMaxN)
    return fail(SPX_ERR_INVALID_DIMENSION, "sparse solve needs 1 <= n <= " +
                                               std::to_string(kSparseMaxN));
  if (nnz < 1 || nnz > 0x7fffffffLL || !rows || !cols)
    return fail(SPX_ERR_INVALID_ARGUMENT, "sparse solve needs a non-empty triplet pattern");
  if (B < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative batch");
  if (B > 0 && (!d_values || !d_rhs || !d_x))
    return fail(SPX_ERR_INVALID_ARGUMENT, "null values, rhs or solutions");
  for (int64_t t = 0; t < nnz; ++t)
    if (rows[t] < 0 || rows[t] >= n || cols[t] < 0 || cols[t] >= n)
      return fail(SPX_ERR_INVALID_ARGUMENT, "triplet " + std::to_string(t) + " outside the matrix");
  if (B == 0) return SPX_OK;
  int rc;
  std::vector<int> key;
  key.reserve(2 + 2 * nnz);
  key.push_back((int)n);
  key.push_back((int)nnz);
  key.insert(key.end(), rows, rows + nnz);
  key.insert(key.end(), cols, cols + nnz);
  void* dp = nullptr;
  if (key != h->lin_key) {
    std::map<std::pair<int, int>, std::vector<int>> pos;
    for (int64_t t = 0; t < nnz; ++t) pos[{rows[t], cols[t]}].push_back((int)t);
    std::vector<int> pos_rc, pos_off(1, 0), pos_trip;
    for (auto& kv : pos) {
      pos_rc.push_back(kv.first.first | (kv.first.second << 16));
      pos_trip.insert(pos_trip.end(), kv.second.begin(), kv.second.end());
      pos_off.push_back((int)pos_trip.size());
    }
    h->lin_key.clear();
    const char* why = build_sparse_plan((int)n, (int)n, pos_rc, h->lin_plan);
    if (*why) return fail(SPX_ERR_INVALID_DIMENSION, why);
    std::vector<int>& blob = h->lin_plan.blob;
    h->lin_o_off = blob.size();
    blob.insert(blob.end(), pos_off.begin(), pos_off.end());
    h->lin_o_trip = blob.size();
    blob.insert(blob.end(), pos_trip.begin(), pos_trip.end());
    blob.push_back(0);
    h->lin_num_pos = (int)pos_rc.size();
    if ((rc = workspace(h, "lin_plan", blob.size() * sizeof(int), &dp))) return rc;
    SPX_HIP(hipMemcpyAsync(dp, blob.data(), blob.size() * sizeof(int), hipMemcpyHostToDevice,
                           h->stream),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
    h->lin_key = std::move(key);
  }
  if ((rc = workspace(h, "lin_plan", h->lin_plan.blob.size() * sizeof(int), &dp))) return rc;
  const int* base = static_cast<const int*>(dp);
  DevSparse sp = device_sparse(h->lin_plan, base);
  int grid = 0, lu_lds = 0;
  int64_t slot = 0;
  size_t smem = 0;
  SPX_HIP(sparse_solve_shape(sp, (int)n, B, &grid, &slot, &smem, &lu_lds), SPX_ERR_BACKEND);
  void* ws = nullptr;
  if ((rc = workspace(h, "lin_ws", sizeof(double) * std::max<int64_t>(1, (int64_t)grid * slot),
                      &ws)))
    return rc;
  SPX_HIP(launch_sparse_solve(sp, (int)n, h->lin_num_pos, base + h->lin_o_off,
                              base + h->lin_o_trip, nnz, d_values, d_rhs, B, d_x, d_sing,
                              static_cast<double*>(ws), grid, slot, smem, lu_lds, h->stream),
          SPX_ERR_BACKEND);
  return SPX_OK;
}

int gmres_locked(spx_handle* h, int64_t n, int64_t nnz, const int32_t* rows, const int32_t* cols,
                 const double* d_values, const double* d_rhs, const double* d_x0, int64_t B,
                 const spx_gmres_config* cfg, double* d_x, int32_t* d_it, double* d_res,
                 uint8_t* d_conv) {
  if (!cfg) return fail(SPX_ERR_INVALID_ARGUMENT, "null GMRES configuration");
  if (cfg->restart < 1 || cfg->restart > 256 || cfg->max_iter < 0 || !(cfg->tol >= 0.0))
    return fail(SPX_ERR_INVALID_ARGUMENT, "GMRES needs 1 <= restart <= 256, max_iter >= 0, tol >= 0");
  if (cfg->precision != SPX_PREC_F64 && cfg->precision != SPX_PREC_F32_IR)
    return fail(SPX_ERR_INVALID_ARGUMENT, "unknown GMRES precision");
  const int prec = cfg->precision == SPX_PREC_F32_IR ? 1 : 0;
  if (n < 1 || n > kSparseMaxN)
    return fail(SPX_ERR_INVALID_DIMENSION, "GMRES needs 1 <= n <= " + std::to_string(kSparseMaxN));
  if (nnz < 1 || nnz > 0x7fffffffLL || !rows || !cols)
    return fail(SPX_ERR_INVALID_ARGUMENT, "GMRES needs a non-empty triplet pattern");
  if (B < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative batch");
  if (B > 0 && (!d_values || !d_rhs || !d_x))
    return fail(SPX_ERR_INVALID_ARGUMENT, "null values, rhs or solutions");
  for (int64_t t = 0; t < nnz; ++t)
    if (rows[t] < 0 || rows[t] >= n || cols[t] < 0 || cols[t] >= n)
      return fail(SPX_ERR_INVALID_ARGUMENT, "triplet " + std::to_string(t) + " outside the matrix");
  if (B == 0) return SPX_OK;
  int rc;
  std::vector<int> key;
  key.reserve(2 + 2 * nnz);
  key.push_back((int)n);
  key.push_back((int)nnz);
  key.insert(key.end(), rows, rows + nnz);
  key.insert(key.end(), cols, cols + nnz);
  void* dp = nullptr;
  if (key != h->gm_key) {
    std::map<std::pair<int, int>, std::vector<int>> pos;
    for (int64_t t = 0; t < nnz; ++t) pos[{rows[t], cols[t]}].push_back((int)t);
    std::vector<int> row_ptr(n + 1, 0), col, off(1, 0), trip, diag(n, -1);
    int q = 0;
    for (auto& kv : pos) {
      row_ptr[kv.first.first + 1]++;
      col.push_back(kv.first.second);
      if (kv.first.first == kv.first.second) diag[kv.first.first] = q;
      trip.insert(trip.end(), kv.second.begin(), kv.second.end());
      off.push_back((int)trip.size());
      ++q;
    }
    for (int64_t i = 0; i < n; ++i) row_ptr[i + 1] += row_ptr[i];
    h->gm_key.clear();
    std::vector<int>& blob = h->gm_blob;
    blob.clear();
    auto put = [&](const std::vector<int>& v) {
      size_t o = blob.size();
      blob.insert(blob.end(), v.begin(), v.end());
      while (blob.size() % 4) blob.push_back(0);
      return o;
    };
    h->gm_o_rowptr = put(row_ptr);
    h->gm_o_col = put(col);
    h->gm_o_off = put(off);
    h->gm_o_trip = put(trip);
    h->gm_o_diag = put(diag);
    blob.push_back(0);
    h->gm_num_pos = q;
    if ((rc = workspace(h, "gm_pattern", blob.size() * sizeof(int), &dp))) return rc;
    SPX_HIP(hipMemcpyAsync(dp, blob.data(), blob.size() * sizeof(int), hipMemcpyHostToDevice,
                           h->stream),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
    h->gm_key = std::move(key);
  }
  if ((rc = workspace(h, "gm_pattern", h->gm_blob.size() * sizeof(int), &dp))) return rc;
  const int* base = static_cast<const int*>(dp);
  DevGmres p{(int)n, h->gm_num_pos, base + h->gm_o_rowptr, base + h->gm_o_col,
             base + h->gm_o_off, base + h->gm_o_trip, base + h->gm_o_diag};
  if (gmres_smem_bytes((int)n, cfg->restart, prec) > 160 * 1024)
    return fail(SPX_ERR_INVALID_DIMENSION, "GMRES workgroup state exceeds the LDS");
  int grid = 0;
  SPX_HIP(gmres_grid((int)n, cfg->restart, prec, B, &grid), SPX_ERR_BACKEND);
  void* ws = nullptr;
  const int64_t slot = gmres_slot_doubles((int)n, h->gm_num_pos, cfg->restart, prec);
  if ((rc = workspace(h, "gm_ws", sizeof(double) * (size_t)grid * slot, &ws))) return rc;
  SPX_HIP(launch_gmres(p, nnz, d_values, d_rhs, d_x0, B, cfg->restart, cfg->max_iter, cfg->tol,
                       cfg->jacobi, prec, d_x, d_it, d_res, d_conv, static_cast<double*>(ws), grid,
                       h->stream),
          SPX_ERR_BACKEND);
  return SPX_OK;
}

int check_tran(const spx_tran_config* t, int n) {
  if (!t) return fail(SPX_ERR_INVALID_ARGUMENT, "null transient configuration");
  if (t->method < SPX_TRAN_BE || t->method > SPX_TRAN_TRBDF2)
    return fail(SPX_ERR_INVALID_ARGUMENT, "unknown integration method " + std::to_string(t->method));
  if (const spx_tran_adaptive* a = t->adaptive) {
    auto pos = [](double v) { return std::isfinite(v) && v > 0.0; };
    if (!pos(a->tstop) || !pos(a->h_init) || !pos(a->h_min) || !pos(a->h_max) ||
        a->h_max < a->h_min)
      return fail(SPX_ERR_INVALID_ARGUMENT,
                  "adaptive transient needs finite tstop, h_init, h_min > 0 and h_max >= h_min");
    if (!(a->reltol >= 0.0) || !(a->abstol >= 0.0) || !std::isfinite(a->reltol) ||
        !std::isfinite(a->abstol))
      return fail(SPX_ERR_INVALID_ARGUMENT, "adaptive transient needs reltol, abstol >= 0");
    if (a->max_points < 1 || a->max_points > 10000000)
      return fail(SPX_ERR_INVALID_ARGUMENT, "adaptive transient needs 1 <= max_points <= 1e7");
  } else if (!(t->tstep > 0.0) || t->steps < 0 || t->steps > 1000000) {
    return fail(SPX_ERR_INVALID_ARGUMENT, "transient needs tstep > 0 and 0 <= steps <= 1e6");
  }
  if (t->num_probes < 0 || (t->num_probes > 0 && !t->probes))
    return fail(SPX_ERR_INVALID_ARGUMENT, "bad probe list");
  for (int p = 0; p < t->num_probes; ++p)
    if (t->probes[p] < 0 || t->probes[p] >= n)
      return fail(SPX_ERR_INVALID_ARGUMENT, "probe " + std::to_string(p) + " outside the unknowns");
  return SPX_OK;
}

// Points per instance in the waveform and time arrays.
int64_t tran_points(const spx_tran_config* t) {
  return t->adaptive ? t->adaptive->max_points : (int64_t)t->steps + 1;
}

DevTran make_dev_tran(const spx_tran_config* t, const int* dprobe) {
  DevTran tr{};
  tr.h = t->tstep;
  tr.steps = t->steps;
  tr.num_probes = t->num_probes;
  tr.probes = dprobe;
  tr.method = t->method;
  if (const spx_tran_adaptive* a = t->adaptive) {
    tr.adaptive = 1;
    tr.method = SPX_TRAN_TRAP;
    tr.tstop = a->tstop;
    tr.h_init = a->h_init;
    tr.h_min = a->h_min;
    tr.h_max = a->h_max;
    tr.reltol = a->reltol;
    tr.abstol = a->abstol;
    tr.max_points = a->max_points;
  }
  return tr;
}

DevTranOut make_dev_tran_out(const spx_tran_result& r, bool adaptive) {
  DevTranOut to{};
  to.waves = r.waveforms;
  to.fail_step = r.fail_step;
  if (adaptive) {
    to.times = r.times;
    to.num_points = r.num_points;
    to.total_steps = r.total_steps;
    to.rejected_steps = r.rejected_steps;
    to.min_step = r.min_step;
    to.max_step = r.max_step;
  }
  return to;
}

int transient_device_locked(spx_handle* h, const spx_circuit* ckt, const spx_sweep_spec* spec,
                            const spx_nr_config* cfg, const spx_tran_config* tran,
                            spx_tran_result* out) {
  int rc;
  if (!out || !out->x) return fail(SPX_ERR_INVALID_ARGUMENT, "result.x is required");
  if ((rc = check_cfg(cfg))) return rc;
  HostProgram hp;
  if ((rc = compile_circuit(ckt, hp, kSparseMaxN, true))) return rc;
  if ((rc = check_spec(spec, hp.P))) return rc;
  if ((rc = check_tran(tran, hp.n))) return rc;
  const int64_t B = spec->batch;
  if (B == 0) return SPX_OK;
  DevProgram g;
  DevSparse sp{};
  if ((rc = upload_program(h, hp, g, &sp))) return rc;
  const double* params = out->params;
  if (!out->params && spec->generator == SPX_GEN_EXPLICIT) {
    params = spec->explicit_params;
  } else {
    double* dst = out->params;
    if (!dst) {
      void* p = nullptr;
      if ((rc = workspace(h, "params", sizeof(double) * std::max<int64_t>(1, B * hp.P), &p)))
        return rc;
      dst = static_cast<double*>(p);
    }
    if ((rc = sample_device(h, spec, hp.P, dst))) return rc;
    params = dst;
  }
  void* dprobe = nullptr;
  if ((rc = workspace(h, "tran_probes", sizeof(int) * std::max(1, tran->num_probes), &dprobe)))
    return rc;
  if (tran->num_probes > 0) {
    SPX_HIP(hipMemcpyAsync(dprobe, tran->probes, sizeof(int) * tran->num_probes,
                           hipMemcpyHostToDevice, h->stream),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  }
  DevTran tr = make_dev_tran(tran, static_cast<const int*>(dprobe));
  DevOut o{out->x, out->status, out->iterations, out->flags, out->residual, nullptr, nullptr};
  DevTranOut to = make_dev_tran_out(*out, tran->adaptive != nullptr);
  DevNrCfg dc = to_dev(cfg);
  if (!o.status && hp.n <= kDenseMaxN) {
    void* st = nullptr;  // the exact FP64 hand-over pass reads the statuses
    if ((rc = workspace(h, "tran_status_ir", sizeof(int) * B, &st))) return rc;
    o.status = static_cast<int*>(st);
  }
  if (hp.n > kDenseMaxN) {
    int grid = 0, lu_lds = 0;
    int64_t slot = 0;
    size_t smem = 0;
    SPX_HIP(sparse_transient_shape(g, sp, B, &grid, &slot, &smem, &lu_lds, 256, -1),
            SPX_ERR_BACKEND);
    void* ws = nullptr;
    if ((rc = workspace(h, "sparse_ws", sizeof(double) * (size_t)grid * slot, &ws))) return rc;
    SPX_HIP(launch_transient_sparse(g, sp, dc, tr, params, B, o, to, static_cast<double*>(ws),
                                    grid, slot, smem, lu_lds, 256, h->stream),
            SPX_ERR_BACKEND);
  } else {
    auto f = hp.n <= 8 ? launch_transient_8 : hp.n <= 16 ? launch_transient_16
           : hp.n <= 32 ? launch_transient_32 : launch_transient_64;
    SPX_HIP(f(g, dc, tr, params, B, o, to, h->stream, 3 | ir_phase(dc)), SPX_ERR_BACKEND);
  }
  return SPX_OK;
}

// ------------------------------------------------------------ mixed batches
struct MixedClass {
  HostProgram hp;
  DevProgram g;
  DevSparse sp{};
  DevTran tr;
  DevOut o;
  DevTranOut to;
  const double* params = nullptr;
  double* sample_dst = nullptr;  // sampled into here (null: explicit parameters read in place)
  void* d_specs = nullptr;
  int gen = 0, nvar = 0;
  uint64_t seed = 0;
  int64_t B = 0, offset = 0;
  int kernel = SPX_KERNEL_AUTO, grid = 0, lu_lds = 0;
  int64_t slot = 0;
  size_t smem = 0;
  double* ws = nullptr;
  int* queue = nullptr;  // sparse kernels: the instance queue (shared GPU, dynamic balance)
  const void* prog = nullptr;
};

using TranFn = hipError_t (*)(const DevProgram&, const DevNrCfg&, const DevTran&, const double*,
                              int64_t, const DevOut&, const DevTranOut&, hipStream_t, int);
TranFn dense_tran(int n) {
  return n <= 8 ? launch_transient_8 : n <= 16 ? launch_transient_16
       : n <= 32 ? launch_transient_32 : launch_transient_64;
}

// Shapes class `mc` for `kernel` over B instances: LDS attributes, grid and
// the "mx_ws<i>" workspace (name given).
int shape_class(spx_handle* h, MixedClass& mc, int kernel, int64_t B, const DevNrCfg& dc,
                const std::string& wsname) {
  int rc;
  mc.kernel = kernel;
  if (kernel == SPX_KERNEL_DENSE) {
    SPX_HIP(dense_tran(mc.hp.n)(mc.g, dc, mc.tr, nullptr, B, mc.o, mc.to, h->stream,
                                1 | ir_phase(dc)),
            SPX_ERR_BACKEND);
    mc.grid = (int)B;
    mc.ws = nullptr;
    mc.queue = nullptr;
    return SPX_OK;
  }
  if (kernel == SPX_KERNEL_LANES) {
    SPX_HIP(lanes_transient_shape(mc.g, mc.sp, B, &mc.grid, &mc.slot), SPX_ERR_BACKEND);
    mc.smem = 0;
    mc.lu_lds = 0;
  } else {
    const int threads = kernel == SPX_KERNEL_SPARSE64 ? 64 : 256;
    SPX_HIP(sparse_transient_shape(mc.g, mc.sp, B, &mc.grid, &mc.slot, &mc.smem, &mc.lu_lds,
                                   threads, -1),
            SPX_ERR_BACKEND);
  }
  void* ws = nullptr;
  if ((rc = workspace(h, wsname.c_str(), sizeof(double) * (size_t)mc.grid * mc.slot, &ws)))
    return rc;
  mc.ws = static_cast<double*>(ws);
  mc.queue = nullptr;
  if (kernel != SPX_KERNEL_LANES) {
    void* q = nullptr;
    if ((rc = workspace(h, (wsname + "_q").c_str(), sizeof(int), &q))) return rc;
    mc.queue = static_cast<int*>(q);
  }
  return SPX_OK;
}

hipError_t enqueue_class(const MixedClass& mc, const DevNrCfg& dc, const DevTran& tr,
                         const DevOut& o, const DevTranOut& to, int64_t B, hipStream_t st) {
  if (mc.kernel == SPX_KERNEL_DENSE)
    return dense_tran(mc.hp.n)(mc.g, dc, tr, mc.params, B, o, to, st, 2 | ir_phase(dc));
  if (mc.kernel == SPX_KERNEL_LANES)
    return launch_transient_lanes(mc.g, mc.sp, dc, tr, mc.params, B, o, to, mc.ws, mc.grid,
                                  mc.slot, st);
  const int threads = mc.kernel == SPX_KERNEL_SPARSE64 ? 64 : 256;
  return launch_transient_sparse(mc.g, mc.sp, dc, tr, mc.params, B, o, to, mc.ws, mc.grid,
                                 mc.slot, mc.smem, mc.lu_lds, threads, st, mc.queue);
}

// Times each candidate kernel on a sample of the class; returns the fastest.
// The sample is the class's own instances up to 8,192 (20 steps): the
// candidates' occupancy differs (one-wave workgroups against four-wave ones
// against 64 instances per wave), and a sample smaller than the GPU ranks
// them as it would not at full size (1,024 instances picked the 256-thread
// sparse kernel for 256-unknown ladders, 20 % slower on the full class).
int autotune_class(spx_handle* h, MixedClass& mc, const DevNrCfg& dc, int* best) {
  const int64_t Bs = std::min<int64_t>(mc.B, 8192);
  DevTran trs = mc.tr;
  if (trs.adaptive) {  // an adaptive class is timed on fixed trapezoidal steps of h_init
    trs.adaptive = 0;
    trs.h = trs.h_init;
    trs.steps = 20;
  }
  trs.steps = std::min(trs.steps, 20);
  trs.num_probes = 0;
  void *dx, *dst, *dit, *dfl, *dres;
  int rc;
  if ((rc = workspace(h, "tune_x", sizeof(double) * Bs * mc.hp.n, &dx))) return rc;
  if ((rc = workspace(h, "tune_status", sizeof(int) * Bs, &dst))) return rc;
  if ((rc = workspace(h, "tune_iter", sizeof(int) * Bs, &dit))) return rc;
  if ((rc = workspace(h, "tune_flags", sizeof(int) * Bs, &dfl))) return rc;
  if ((rc = workspace(h, "tune_res", sizeof(double) * Bs, &dres))) return rc;
  DevOut o{(double*)dx, (int*)dst, (int*)dit, (int*)dfl, (double*)dres, nullptr, nullptr};
  DevTranOut to{};
  std::vector<int> cands;
  if (mc.hp.n <= kDenseMaxN) cands.push_back(SPX_KERNEL_DENSE);
  cands.push_back(SPX_KERNEL_SPARSE64);
  cands.push_back(SPX_KERNEL_SPARSE256);
  cands.push_back(SPX_KERNEL_LANES);
  hipEvent_t e0, e1;
  SPX_HIP(hipEventCreate(&e0), SPX_ERR_BACKEND);
  SPX_HIP(hipEventCreate(&e1), SPX_ERR_BACKEND);
  float best_ms = 1e30f;
  *best = cands[0];
  for (int k : cands) {
    MixedClass t = mc;
    if ((rc = shape_class(h, t, k, Bs, dc, "tune_ws"))) break;
    float ms = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {  // first run warms caches and code
      if (hipEventRecord(e0, h->stream) != hipSuccess ||
          enqueue_class(t, dc, trs, o, to, Bs, h->stream) != hipSuccess ||
          hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
        rc = fail(SPX_ERR_BACKEND, "autotune launch failed");
        break;
      }
      float m = 0.0f;
      (void)hipEventElapsedTime(&m, e0, e1);
      if (rep > 0) ms = m;
    }
    if (rc) break;
    if (ms < best_ms) {
      best_ms = ms;
      *best = k;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// The handle's side streams and fork/join events (mixed batches).
int ensure_side_streams(spx_handle* h) {
  if (h->fork_ev) return SPX_OK;
  for (int s = 0; s < spx_handle::kSideStreams; ++s) {
    SPX_HIP(hipStreamCreateWithFlags(&h->side[s], hipStreamNonBlocking), SPX_ERR_BACKEND_INIT);
    SPX_HIP(hipEventCreateWithFlags(&h->join_ev[s], hipEventDisableTiming), SPX_ERR_BACKEND_INIT);
  }
  SPX_HIP(hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming), SPX_ERR_BACKEND_INIT);
  return SPX_OK;
}

int mixed_locked(spx_handle* h, const spx_tran_class* cls, int nc, const spx_nr_config* cfg,
                 uint32_t options, int32_t* kernels) {
  int rc;
  if (nc < 0 || (nc > 0 && !cls)) return fail(SPX_ERR_INVALID_ARGUMENT, "bad class list");
  if ((rc = check_cfg(cfg))) return rc;
  const DevNrCfg dc = to_dev(cfg);
  std::vector<MixedClass> mcs(nc);
  std::vector<uint64_t> key;
  auto kput = [&](const void* p, size_t bytes) {
    std::vector<uint64_t> w((bytes + 7) / 8, 0);
    std::memcpy(w.data(), p, bytes);
    key.insert(key.end(), w.begin(), w.end());
  };
  kput(&dc, sizeof(dc));
  for (int i = 0; i < nc; ++i) {
    const spx_tran_class& c = cls[i];
    MixedClass& mc = mcs[i];
    const std::string tag = "_m" + std::to_string(i);
    if (!c.result.x) return fail(SPX_ERR_INVALID_ARGUMENT, "class result.x is required");
    if ((rc = compile_circuit(c.circuit, mc.hp, kSparseMaxN, true))) return rc;
    if ((rc = check_spec(&c.spec, mc.hp.P))) return rc;
    if ((rc = check_tran(&c.tran, mc.hp.n))) return rc;
    int req = kernels ? kernels[i] : SPX_KERNEL_AUTO;
    if (req < SPX_KERNEL_AUTO || req > SPX_KERNEL_LANES ||
        (req == SPX_KERNEL_DENSE && mc.hp.n > kDenseMaxN))
      return fail(SPX_ERR_INVALID_ARGUMENT, "class " + std::to_string(i) + ": bad kernel kind");
    mc.hp.want_sparse = true;
    if ((rc = upload_program(h, mc.hp, mc.g, &mc.sp, tag))) return rc;
    mc.prog = mc.g.elems;
    mc.B = c.spec.batch;
    mc.gen = c.spec.generator;
    mc.seed = c.spec.seed;
    mc.offset = c.spec.sweep_offset;
    const int P = mc.hp.P;
    if (c.spec.generator == SPX_GEN_EXPLICIT) {
      mc.params = c.spec.explicit_params;
      if (c.result.params && mc.B * P > 0)  // copy out for the caller, outside any graph
        SPX_HIP(hipMemcpyAsync(c.result.params, c.spec.explicit_params,
                               sizeof(double) * mc.B * P, hipMemcpyDeviceToDevice, h->stream),
                SPX_ERR_TRANSFER);
    } else {
      double* dst = c.result.params;
      if (!dst) {
        void* p = nullptr;
        if ((rc = workspace(h, ("mx_params" + tag).c_str(),
                            sizeof(double) * std::max<int64_t>(1, mc.B * P), &p)))
          return rc;
        dst = static_cast<double*>(p);
      }
      mc.sample_dst = dst;
      mc.params = dst;
      if (P > 0) {
        if ((rc = workspace(h, ("mx_specs" + tag).c_str(), sizeof(spx_param_spec) * P,
                            &mc.d_specs)))
          return rc;
        SPX_HIP(hipMemcpyAsync(mc.d_specs, c.spec.params, sizeof(spx_param_spec) * P,
                               hipMemcpyHostToDevice, h->stream),
                SPX_ERR_TRANSFER);
        mc.nvar = count_varied(c.spec.params, P);
      }
    }
    void* dprobe = nullptr;
    if ((rc = workspace(h, ("mx_probes" + tag).c_str(), sizeof(int) * std::max(1, c.tran.num_probes),
                        &dprobe)))
      return rc;
    if (c.tran.num_probes > 0)
      SPX_HIP(hipMemcpyAsync(dprobe, c.tran.probes, sizeof(int) * c.tran.num_probes,
                             hipMemcpyHostToDevice, h->stream),
              SPX_ERR_TRANSFER);
    SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
    mc.tr = make_dev_tran(&c.tran, static_cast<const int*>(dprobe));
    mc.o = DevOut{c.result.x, c.result.status, c.result.iterations, c.result.flags,
                  c.result.residual, nullptr, nullptr};
    if (!mc.o.status) {
      void* st = nullptr;  // the exact FP64 hand-over pass reads the statuses
      if ((rc = workspace(h, ("mx_status" + tag).c_str(), sizeof(int) * std::max<int64_t>(1, mc.B),
                          &st)))
        return rc;
      mc.o.status = static_cast<int*>(st);
    }
    mc.to = make_dev_tran_out(c.result, c.tran.adaptive != nullptr);
    // the fastest kernel depends on the precision form too
    std::vector<int> tkey = h->programs[tag].blob;
    tkey.push_back(dc.precision);
    int kind = req;
    if (kind == SPX_KERNEL_AUTO) {
      auto it = h->tuned.find(tkey);
      if (it != h->tuned.end()) {
        kind = it->second;
      } else if ((options & SPX_MIXED_AUTOTUNE) && mc.B > 0) {
        if (mc.sample_dst && mc.hp.P > 0)  // the sample needs parameters now
          SPX_HIP(launch_sample(static_cast<spx_param_spec*>(mc.d_specs), mc.hp.P, mc.gen,
                                mc.seed, mc.B, mc.offset, mc.nvar, mc.sample_dst, h->stream),
                  SPX_ERR_BACKEND);
        if ((rc = autotune_class(h, mc, dc, &kind))) return rc;
        h->tuned[tkey] = kind;
      } else {
        kind = mc.hp.n <= kDenseMaxN ? SPX_KERNEL_DENSE : SPX_KERNEL_SPARSE256;
      }
    }
    if ((rc = shape_class(h, mc, kind, mc.B, dc, "mx_ws" + tag))) return rc;
    if (kernels) kernels[i] = kind;
    const uint64_t words[] = {(uint64_t)kind, (uint64_t)mc.B, (uint64_t)mc.hp.n,
                              (uint64_t)mc.tr.steps, (uint64_t)mc.tr.num_probes,
                              (uint64_t)mc.gen, mc.seed, (uint64_t)mc.offset, (uint64_t)mc.grid,
                              (uint64_t)mc.slot, (uint64_t)mc.smem, (uint64_t)mc.lu_lds,
                              (uint64_t)mc.tr.method, (uint64_t)mc.tr.adaptive,
                              (uint64_t)mc.tr.max_points};
    kput(words, sizeof(words));
    const double dwords[] = {mc.tr.h,      mc.tr.tstop,  mc.tr.h_init, mc.tr.h_min,
                             mc.tr.h_max,  mc.tr.reltol, mc.tr.abstol};
    kput(dwords, sizeof(dwords));
    const void* ptrs[] = {mc.params, mc.sample_dst, mc.d_specs, dprobe, mc.ws, mc.queue, mc.prog,
                          mc.sp.perm, c.result.x, mc.o.status, c.result.iterations,
                          c.result.flags, c.result.residual, c.result.waveforms,
                          c.result.fail_step, mc.to.times, mc.to.num_points,
                          mc.to.total_steps, mc.to.rejected_steps, mc.to.min_step,
                          mc.to.max_step};
    kput(ptrs, sizeof(ptrs));
  }
  // Classes share nothing, so the register-LU (dense) classes run beside the
  // others: the sparse and lane classes, whose grids fill the GPU, stay in
  // order on the main stream, and the dense classes (a workgroup per
  // instance) go to up to kSideStreams side streams (longest first onto the
  // least-loaded one, cost ~ B n^1.5), forked from and joined back into the
  // main stream with events.  Their waves fill the slots the persistent
  // kernels leave between instances.  Captured, the streams become parallel
  // branches of one graph.
  int nstreams = 1;
  std::vector<int> lane_of(nc, 0);
  std::vector<int> dense;
  for (int i = 0; i < nc; ++i)
    if (mcs[i].kernel == SPX_KERNEL_DENSE) dense.push_back(i);
  if (nc > 1 && !dense.empty()) {
    if ((rc = ensure_side_streams(h))) return rc;
    const bool others = (int)dense.size() < nc;
    const int nside = std::min((int)dense.size(), spx_handle::kSideStreams);
    nstreams = 1 + nside;
    auto cost = [&](int i) { return (double)mcs[i].B * std::pow((double)mcs[i].hp.n, 1.5); };
    std::stable_sort(dense.begin(), dense.end(), [&](int a, int b) { return cost(a) > cost(b); });
    // without sparse classes the main stream takes a dense share too
    std::vector<double> load(nstreams, 0.0);
    if (others) load[0] = 1e300;
    for (int i : dense) {
      const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      lane_of[i] = s;
      load[s] += cost(i);
    }
  }
  auto enqueue_all = [&](hipStream_t st) -> hipError_t {
    hipError_t e;
    auto stream_of = [&](int lane) { return lane == 0 ? st : h->side[lane - 1]; };
    if (nstreams > 1) {
      if ((e = hipEventRecord(h->fork_ev, st)) != hipSuccess) return e;
      for (int s = 1; s < nstreams; ++s)
        if ((e = hipStreamWaitEvent(h->side[s - 1], h->fork_ev, 0)) != hipSuccess) return e;
    }
    for (int i = 0; i < nc; ++i) {
      const MixedClass& mc = mcs[i];
      hipStream_t cs = stream_of(lane_of[i]);
      if (mc.sample_dst && mc.hp.P > 0 &&
          (e = launch_sample(static_cast<spx_param_spec*>(mc.d_specs), mc.hp.P, mc.gen, mc.seed,
                             mc.B, mc.offset, mc.nvar, mc.sample_dst, cs)) != hipSuccess)
        return e;
      if ((e = enqueue_class(mc, dc, mc.tr, mc.o, mc.to, mc.B, cs)) != hipSuccess) return e;
    }
    for (int s = 1; s < nstreams; ++s) {
      if ((e = hipEventRecord(h->join_ev[s - 1], h->side[s - 1])) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(st, h->join_ev[s - 1], 0)) != hipSuccess) return e;
    }
    return hipSuccess;
  };
  if (!(options & SPX_MIXED_GRAPH)) {
    SPX_HIP(enqueue_all(h->stream), SPX_ERR_BACKEND);
    return SPX_OK;
  }
  if (!h->mixed_exec || key != h->mixed_key) {
    if (h->mixed_exec) {
      (void)hipGraphExecDestroy(h->mixed_exec);
      h->mixed_exec = nullptr;
    }
    h->mixed_key.clear();
    hipGraph_t graph = nullptr;
    SPX_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed), SPX_ERR_BACKEND);
    hipError_t e = enqueue_all(h->stream);
    hipError_t e2 = hipStreamEndCapture(h->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return fail(SPX_ERR_BACKEND, std::string("graph capture failed: ") +
                                       hipGetErrorString(e != hipSuccess ? e : e2));
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess)
      return fail(SPX_ERR_BACKEND, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    h->mixed_exec = exec;
    h->mixed_key = std::move(key);
  }
  SPX_HIP(hipGraphLaunch(h->mixed_exec, h->stream), SPX_ERR_BACKEND);
  return SPX_OK;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

int spx_device_count(int32_t* count) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    if (count) *count = 0;
    return fail(SPX_ERR_NO_BACKEND, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  if (count) *count = c;
  return SPX_OK;
}

int spx_create(int32_t device_id, spx_handle** out) {
  if (!out) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle pointer");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(SPX_ERR_NO_BACKEND, "no HIP device available");
  if (device_id < 0 || device_id >= count)
    return fail(SPX_ERR_BACKEND_INIT, "device " + std::to_string(device_id) + " out of range");
  DeviceGuard dg(device_id);
  if (int rc = noalloc_scratch_init(device_id)) return rc;
  auto* h = new spx_handle;
  h->device = device_id;
  e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return fail(SPX_ERR_BACKEND_INIT, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *out = h;
  return SPX_OK;
}

int spx_destroy(spx_handle* h) {
  if (!h) return SPX_OK;
  {
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    hipStreamSynchronize(h->stream);
    if (h->mixed_exec) hipGraphExecDestroy(h->mixed_exec);
    for (int s = 0; s < spx_handle::kSideStreams; ++s) {
      if (h->side[s]) hipStreamSynchronize(h->side[s]);
      if (h->side[s]) hipStreamDestroy(h->side[s]);
      if (h->join_ev[s]) hipEventDestroy(h->join_ev[s]);
    }
    if (h->fork_ev) hipEventDestroy(h->fork_ev);
    for (auto& kv : h->ws)
      if (kv.second.first) hipFree(kv.second.first);
    hipStreamDestroy(h->stream);
  }
  delete h;
  return SPX_OK;
}

const char* spx_last_error(void) { return last_error().c_str(); }

int spx_synchronize(spx_handle* h) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  DeviceGuard dg(h->device);
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_BACKEND);
  return SPX_OK;
}

void* spx_stream(spx_handle* h) { return h ? (void*)h->stream : nullptr; }

int64_t spx_max_batch_per_launch(void) { return 0x7fffffffLL; }

int spx_solve_batch_f64_device(spx_handle* h, const double* dM, const double* drhs, int64_t n,
                               int64_t B, double* dx, uint8_t* dsing) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  if (B < 0 || n < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (B == 0) return SPX_OK;
  if (n < 1 || n > 4096)
    return fail(SPX_ERR_INVALID_DIMENSION, "n must be 1..4096, got " + std::to_string(n));
  if (B > spx_max_batch_per_launch())
    return fail(SPX_ERR_BATCH_TOO_LARGE, "batch " + std::to_string(B) + " exceeds " +
                                             std::to_string(spx_max_batch_per_launch()));
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  double* work = nullptr;
  int64_t chunk = 1;
  if (n > 64) {
    chunk = std::max<int64_t>(1, std::min<int64_t>(B, (int64_t)(256ull << 20) /
                                                          (int64_t)(sizeof(double) * n * n)));
    void* w = nullptr;
    int rc = workspace(h, "lu_generic", solve_generic_workspace((int)n, chunk), &w);
    if (rc) return rc;
    work = static_cast<double*>(w);
  }
  SPX_HIP(launch_solve_batch(dM, drhs, (int)n, B, dx, dsing, work, chunk, h->stream),
          SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_lu_factor_batch_f64_device(spx_handle* h, const double* dM, int64_t n, int64_t B,
                                   double* dF, int32_t* dpiv, uint8_t* dsing) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  if (B < 0 || n < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (B == 0) return SPX_OK;
  if (!dM || !dF || !dpiv) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  if (n < 1 || n > 4096)
    return fail(SPX_ERR_INVALID_DIMENSION, "n must be 1..4096, got " + std::to_string(n));
  if (B > spx_max_batch_per_launch())
    return fail(SPX_ERR_BATCH_TOO_LARGE, "batch " + std::to_string(B) + " exceeds " +
                                             std::to_string(spx_max_batch_per_launch()));
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  SPX_HIP(launch_factor_batch(dM, (int)n, B, dF, dpiv, dsing, h->stream), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_lu_factor_batch_f64(spx_handle* h, const double* M, int64_t n, int64_t B, double* F,
                            int32_t* piv, uint8_t* sing, int64_t* num_singular) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  if (B < 0 || n < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (num_singular) *num_singular = 0;
  if (B == 0) return SPX_OK;
  if (!M || !F || !piv) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  if (n < 1 || n > 4096)
    return fail(SPX_ERR_INVALID_DIMENSION, "n must be 1..4096, got " + std::to_string(n));
  void *dM, *dF, *dp, *ds;
  {
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    int rc;
    if ((rc = workspace(h, "sb_M", sizeof(double) * B * n * n, &dM))) return rc;
    if ((rc = workspace(h, "lf_F", sizeof(double) * B * n * n, &dF))) return rc;
    if ((rc = workspace(h, "lf_piv", sizeof(int32_t) * B * n, &dp))) return rc;
    if ((rc = workspace(h, "sb_sing", B, &ds))) return rc;
    SPX_HIP(hipMemcpyAsync(dM, M, sizeof(double) * B * n * n, hipMemcpyHostToDevice, h->stream),
            SPX_ERR_TRANSFER);
  }
  int rc = spx_lu_factor_batch_f64_device(h, (double*)dM, n, B, (double*)dF, (int32_t*)dp,
                                          (uint8_t*)ds);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  std::vector<uint8_t> hs(B);
  SPX_HIP(hipMemcpyAsync(F, dF, sizeof(double) * B * n * n, hipMemcpyDeviceToHost, h->stream),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipMemcpyAsync(piv, dp, sizeof(int32_t) * B * n, hipMemcpyDeviceToHost, h->stream),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipMemcpyAsync(hs.data(), ds, B, hipMemcpyDeviceToHost, h->stream), SPX_ERR_TRANSFER);
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  int64_t ns = 0;
  for (int64_t i = 0; i < B; ++i) ns += hs[i];
  if (sing) std::memcpy(sing, hs.data(), B);
  if (num_singular) *num_singular = ns;
  return SPX_OK;
}

int spx_solve_batch_f64(spx_handle* h, const double* M, const double* rhs, int64_t n, int64_t B,
                        double* x, uint8_t* sing, int64_t* num_singular) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  if (B < 0 || n < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative size");
  if (num_singular) *num_singular = 0;
  if (B == 0) return SPX_OK;  // batched_lu.rs:380-387
  if (!M || !rhs || !x) return fail(SPX_ERR_INVALID_ARGUMENT, "null buffer");
  if (n < 1 || n > 4096)
    return fail(SPX_ERR_INVALID_DIMENSION, "n must be 1..4096, got " + std::to_string(n));
  void *dM, *dr, *dx, *ds;
  {
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    int rc;
    if ((rc = workspace(h, "sb_M", sizeof(double) * B * n * n, &dM))) return rc;
    if ((rc = workspace(h, "sb_rhs", sizeof(double) * B * n, &dr))) return rc;
    if ((rc = workspace(h, "sb_x", sizeof(double) * B * n, &dx))) return rc;
    if ((rc = workspace(h, "sb_sing", B, &ds))) return rc;
    SPX_HIP(hipMemcpyAsync(dM, M, sizeof(double) * B * n * n, hipMemcpyHostToDevice, h->stream),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipMemcpyAsync(dr, rhs, sizeof(double) * B * n, hipMemcpyHostToDevice, h->stream),
            SPX_ERR_TRANSFER);
  }
  int rc = spx_solve_batch_f64_device(h, (double*)dM, (double*)dr, n, B, (double*)dx,
                                      (uint8_t*)ds);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  std::vector<uint8_t> hs(B);
  SPX_HIP(hipMemcpyAsync(x, dx, sizeof(double) * B * n, hipMemcpyDeviceToHost, h->stream),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipMemcpyAsync(hs.data(), ds, B, hipMemcpyDeviceToHost, h->stream), SPX_ERR_TRANSFER);
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  int64_t ns = 0;
  for (int64_t i = 0; i < B; ++i) ns += hs[i];
  if (sing) std::memcpy(sing, hs.data(), B);
  if (num_singular) *num_singular = ns;
  return SPX_OK;
}

int spx_sample_params_device(spx_handle* h, const spx_sweep_spec* spec, int32_t P,
                             double* d_params) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  int rc = check_spec(spec, P);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  return sample_device(h, spec, P, d_params);
}

int spx_sample_params(spx_handle* h, const spx_sweep_spec* spec, int32_t P, double* params) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  int rc = check_spec(spec, P);
  if (rc) return rc;
  const int64_t B = spec->batch;
  if (B == 0 || P == 0) return SPX_OK;
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  void *d = nullptr, *de = nullptr;
  if ((rc = workspace(h, "sp_out", sizeof(double) * B * P, &d))) return rc;
  spx_sweep_spec local = *spec;
  if (spec->generator == SPX_GEN_EXPLICIT) {
    if ((rc = workspace(h, "sp_explicit", sizeof(double) * B * P, &de))) return rc;
    SPX_HIP(hipMemcpyAsync(de, spec->explicit_params, sizeof(double) * B * P,
                           hipMemcpyHostToDevice, h->stream),
            SPX_ERR_TRANSFER);
    local.explicit_params = static_cast<double*>(de);
  }
  if ((rc = sample_device(h, &local, P, static_cast<double*>(d)))) return rc;
  SPX_HIP(hipMemcpyAsync(params, d, sizeof(double) * B * P, hipMemcpyDeviceToHost, h->stream),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  return SPX_OK;
}

void spx_nr_config_default(spx_nr_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->v_abstol = 1e-6;
  c->v_reltol = 1e-3;
  c->i_abstol = 1e-12;
  c->max_iterations = 50;
  c->gmin_stepping = 1;
  c->gmin_initial = 1e-3;
  c->gmin_final = 1e-12;
  c->gmin_factor = 10.0;
  c->gmin_max_attempts = 5;
  c->source_stepping = 1;
  c->src_initial = 0.1;
  c->src_step = 0.1;
  c->src_min_step = 0.01;
  c->src_max_attempts = 5;
  c->precision = SPX_PREC_F64;
  c->limiter = SPX_LIMIT_REFERENCE;
  c->warm_start = SPX_WARM_ZERO;
  c->ir_max_steps = 8;
  c->ir_tol = 0.0;
  c->kcl_reltol = 1e-3;
  c->kcl_abstol = 1e-12;
  c->kcl_vntol = 1e-6;
}

int32_t spx_circuit_size(const spx_circuit* c) {
  if (!c || !c->elements) return c ? c->num_nodes : -1;
  int32_t nv = 0;
  for (int e = 0; e < c->num_elements; ++e)
    nv += has_branch(c->elements[e].kind);
  return c->num_nodes + nv;
}

int spx_sweep_dcop_device(spx_handle* h, const spx_circuit* ckt, const spx_sweep_spec* spec,
                          const spx_nr_config* cfg, const double* d_x0, spx_sweep_result* out) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  return sweep_device_locked(h, ckt, spec, cfg, d_x0, out);
}

int spx_sweep_dcop(spx_handle* h, const spx_circuit* ckt, const spx_sweep_spec* spec,
                   const spx_nr_config* cfg, const double* x0, spx_sweep_result* out) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  if (!out || !out->x) return fail(SPX_ERR_INVALID_ARGUMENT, "result.x is required");
  int rc;
  HostProgram hp;
  if ((rc = compile_circuit(ckt, hp, kSparseMaxN))) return rc;
  if ((rc = check_spec(spec, hp.P))) return rc;
  if ((rc = check_cfg(cfg))) return rc;
  const int64_t B = spec->batch;
  if (B == 0) return SPX_OK;
  const int n = hp.n, P = hp.P;
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  void *dx, *dst, *dit, *dfl, *dres, *dpar, *dstats = nullptr, *dx0 = nullptr, *dexp = nullptr;
  void *dfac = nullptr, *dref = nullptr;
  if ((rc = workspace(h, "sw_x", sizeof(double) * B * n, &dx))) return rc;
  if ((rc = workspace(h, "sw_status", sizeof(int) * B, &dst))) return rc;
  if ((rc = workspace(h, "sw_iter", sizeof(int) * B, &dit))) return rc;
  if ((rc = workspace(h, "sw_flags", sizeof(int) * B, &dfl))) return rc;
  if ((rc = workspace(h, "sw_res", sizeof(double) * B, &dres))) return rc;
  if ((rc = workspace(h, "sw_params", sizeof(double) * std::max<int64_t>(1, B * P), &dpar)))
    return rc;
  if (out->node_stats && (rc = workspace(h, "sw_stats", sizeof(double) * 5 * n, &dstats)))
    return rc;
  if (out->factorizations && (rc = workspace(h, "sw_fac", sizeof(int) * B, &dfac))) return rc;
  if (out->refinements && (rc = workspace(h, "sw_ref", sizeof(int) * B, &dref))) return rc;
  if (cfg->warm_start == SPX_WARM_GIVEN) {
    if (!x0) return fail(SPX_ERR_INVALID_ARGUMENT, "warm start given without x0");
    if ((rc = workspace(h, "sw_x0", sizeof(double) * B * n, &dx0))) return rc;
    SPX_HIP(hipMemcpyAsync(dx0, x0, sizeof(double) * B * n, hipMemcpyHostToDevice, h->stream),
            SPX_ERR_TRANSFER);
  }
  spx_sweep_spec local = *spec;
  if (spec->generator == SPX_GEN_EXPLICIT && P > 0) {
    if ((rc = workspace(h, "sw_explicit", sizeof(double) * B * P, &dexp))) return rc;
    SPX_HIP(hipMemcpyAsync(dexp, spec->explicit_params, sizeof(double) * B * P,
                           hipMemcpyHostToDevice, h->stream),
            SPX_ERR_TRANSFER);
    local.explicit_params = static_cast<double*>(dexp);
  }
  spx_sweep_result dres_s{(double*)dx,   (int*)dst,      (int*)dit, (int*)dfl,
                          (double*)dres, (double*)dpar,  (double*)dstats, (int*)dfac,
                          (int*)dref};
  if ((rc = sweep_device_locked(h, ckt, &local, cfg, (double*)dx0, &dres_s))) return rc;
  auto d2h = [&](void* dst_h, const void* src_d, size_t bytes) -> int {
    if (!dst_h) return SPX_OK;
    SPX_HIP(hipMemcpyAsync(dst_h, src_d, bytes, hipMemcpyDeviceToHost, h->stream),
            SPX_ERR_TRANSFER);
    return SPX_OK;
  };
  if ((rc = d2h(out->x, dx, sizeof(double) * B * n))) return rc;
  if ((rc = d2h(out->status, dst, sizeof(int) * B))) return rc;
  if ((rc = d2h(out->iterations, dit, sizeof(int) * B))) return rc;
  if ((rc = d2h(out->flags, dfl, sizeof(int) * B))) return rc;
  if ((rc = d2h(out->residual, dres, sizeof(double) * B))) return rc;
  if ((rc = d2h(out->params, dpar, sizeof(double) * B * P))) return rc;
  if (out->node_stats && (rc = d2h(out->node_stats, dstats, sizeof(double) * 5 * n))) return rc;
  if ((rc = d2h(out->factorizations, dfac, sizeof(int) * B))) return rc;
  if ((rc = d2h(out->refinements, dref, sizeof(int) * B))) return rc;
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  return SPX_OK;
}

int spx_sparse_solve_batch_f64_device(spx_handle* h, int64_t n, int64_t nnz, const int32_t* rows,
                                      const int32_t* cols, const double* d_values,
                                      const double* d_rhs, int64_t batch, double* d_solutions,
                                      uint8_t* d_singular) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  return sparse_solve_locked(h, n, nnz, rows, cols, d_values, d_rhs, batch, d_solutions,
                             d_singular);
}

int spx_sparse_solve_batch_f64(spx_handle* h, int64_t n, int64_t nnz, const int32_t* rows,
                               const int32_t* cols, const double* values, const double* rhs,
                               int64_t batch, double* solutions, uint8_t* singular,
                               int64_t* num_singular) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  if (num_singular) *num_singular = 0;
  if (batch < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative batch");
  if (batch > 0 && (!values || !rhs || !solutions))
    return fail(SPX_ERR_INVALID_ARGUMENT, "null values, rhs or solutions");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  if (batch == 0) return sparse_solve_locked(h, n, nnz, rows, cols, nullptr, nullptr, 0, nullptr,
                                             nullptr);
  int rc;
  void *dv, *db, *dx, *ds;
  if ((rc = workspace(h, "ls_values", sizeof(double) * batch * nnz, &dv))) return rc;
  if ((rc = workspace(h, "ls_rhs", sizeof(double) * batch * n, &db))) return rc;
  if ((rc = workspace(h, "ls_x", sizeof(double) * batch * n, &dx))) return rc;
  if ((rc = workspace(h, "ls_sing", batch, &ds))) return rc;
  SPX_HIP(hipMemcpyAsync(dv, values, sizeof(double) * batch * nnz, hipMemcpyHostToDevice,
                         h->stream),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipMemcpyAsync(db, rhs, sizeof(double) * batch * n, hipMemcpyHostToDevice, h->stream),
          SPX_ERR_TRANSFER);
  if ((rc = sparse_solve_locked(h, n, nnz, rows, cols, (double*)dv, (double*)db, batch,
                                (double*)dx, (uint8_t*)ds)))
    return rc;
  std::vector<uint8_t> sing(batch);
  SPX_HIP(hipMemcpyAsync(solutions, dx, sizeof(double) * batch * n, hipMemcpyDeviceToHost,
                         h->stream),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipMemcpyAsync(sing.data(), ds, batch, hipMemcpyDeviceToHost, h->stream),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  int64_t ns = 0;
  for (int64_t i = 0; i < batch; ++i) ns += sing[i] != 0;
  if (singular) std::memcpy(singular, sing.data(), batch);
  if (num_singular) *num_singular = ns;
  return SPX_OK;
}

void spx_gmres_config_default(spx_gmres_config* cfg) {
  if (!cfg) return;
  cfg->restart = 30;  // GmresConfig::default, gmres/mod.rs:34-48
  cfg->max_iter = 500;
  cfg->tol = 1e-8;
  cfg->jacobi = 1;
  cfg->precision = SPX_PREC_F64;
}

int spx_gmres_solve_batch_f64_device(spx_handle* h, int64_t n, int64_t nnz, const int32_t* rows,
                                     const int32_t* cols, const double* d_values,
                                     const double* d_rhs, const double* d_x0, int64_t batch,
                                     const spx_gmres_config* cfg, double* d_solutions,
                                     int32_t* d_iterations, double* d_residuals,
                                     uint8_t* d_converged) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  return gmres_locked(h, n, nnz, rows, cols, d_values, d_rhs, d_x0, batch, cfg, d_solutions,
                      d_iterations, d_residuals, d_converged);
}

int spx_gmres_solve_batch_f64(spx_handle* h, int64_t n, int64_t nnz, const int32_t* rows,
                              const int32_t* cols, const double* values, const double* rhs,
                              const double* x0, int64_t batch, const spx_gmres_config* cfg,
                              double* solutions, int32_t* iterations, double* residuals,
                              uint8_t* converged) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  if (batch < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative batch");
  if (batch > 0 && (!values || !rhs || !solutions))
    return fail(SPX_ERR_INVALID_ARGUMENT, "null values, rhs or solutions");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  if (batch == 0)
    return gmres_locked(h, n, nnz, rows, cols, nullptr, nullptr, nullptr, 0, cfg, nullptr,
                        nullptr, nullptr, nullptr);
  int rc;
  void *dv, *db, *dx, *dx0 = nullptr, *dit, *dres, *dconv;
  if ((rc = workspace(h, "gs_values", sizeof(double) * batch * nnz, &dv))) return rc;
  if ((rc = workspace(h, "gs_rhs", sizeof(double) * batch * n, &db))) return rc;
  if ((rc = workspace(h, "gs_x", sizeof(double) * batch * n, &dx))) return rc;
  if ((rc = workspace(h, "gs_it", sizeof(int32_t) * batch, &dit))) return rc;
  if ((rc = workspace(h, "gs_res", sizeof(double) * batch, &dres))) return rc;
  if ((rc = workspace(h, "gs_conv", batch, &dconv))) return rc;
  SPX_HIP(hipMemcpyAsync(dv, values, sizeof(double) * batch * nnz, hipMemcpyHostToDevice,
                         h->stream),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipMemcpyAsync(db, rhs, sizeof(double) * batch * n, hipMemcpyHostToDevice, h->stream),
          SPX_ERR_TRANSFER);
  if (x0) {
    if ((rc = workspace(h, "gs_x0", sizeof(double) * batch * n, &dx0))) return rc;
    SPX_HIP(hipMemcpyAsync(dx0, x0, sizeof(double) * batch * n, hipMemcpyHostToDevice,
                           h->stream),
            SPX_ERR_TRANSFER);
  }
  if ((rc = gmres_locked(h, n, nnz, rows, cols, (double*)dv, (double*)db, (double*)dx0, batch,
                         cfg, (double*)dx, (int32_t*)dit, (double*)dres, (uint8_t*)dconv)))
    return rc;
  auto d2h = [&](void* dst, const void* src, size_t bytes) -> int {
    if (!dst) return SPX_OK;
    SPX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream), SPX_ERR_TRANSFER);
    return SPX_OK;
  };
  if ((rc = d2h(solutions, dx, sizeof(double) * batch * n))) return rc;
  if ((rc = d2h(iterations, dit, sizeof(int32_t) * batch))) return rc;
  if ((rc = d2h(residuals, dres, sizeof(double) * batch))) return rc;
  if ((rc = d2h(converged, dconv, batch))) return rc;
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  return SPX_OK;
}

int spx_transient_sweep_device(spx_handle* h, const spx_circuit* ckt, const spx_sweep_spec* spec,
                               const spx_nr_config* cfg, const spx_tran_config* tran,
                               spx_tran_result* out) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  return transient_device_locked(h, ckt, spec, cfg, tran, out);
}

int spx_transient_sweep(spx_handle* h, const spx_circuit* ckt, const spx_sweep_spec* spec,
                        const spx_nr_config* cfg, const spx_tran_config* tran,
                        spx_tran_result* out) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  if (!out || !out->x) return fail(SPX_ERR_INVALID_ARGUMENT, "result.x is required");
  int rc;
  HostProgram hp;
  if ((rc = compile_circuit(ckt, hp, kSparseMaxN, true))) return rc;
  if ((rc = check_spec(spec, hp.P))) return rc;
  if ((rc = check_tran(tran, hp.n))) return rc;
  const int64_t B = spec->batch;
  if (B == 0) return SPX_OK;
  const int n = hp.n, P = hp.P;
  const int64_t pts = tran_points(tran);
  const int64_t W = B * pts * tran->num_probes;
  const bool adaptive = tran->adaptive != nullptr;
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  void *dx, *dst, *dit, *dfl, *dres, *dpar, *dwav = nullptr, *dfs, *dexp = nullptr;
  if ((rc = workspace(h, "tr_x", sizeof(double) * B * n, &dx))) return rc;
  if ((rc = workspace(h, "tr_status", sizeof(int) * B, &dst))) return rc;
  if ((rc = workspace(h, "tr_iter", sizeof(int) * B, &dit))) return rc;
  if ((rc = workspace(h, "tr_flags", sizeof(int) * B, &dfl))) return rc;
  if ((rc = workspace(h, "tr_res", sizeof(double) * B, &dres))) return rc;
  if ((rc = workspace(h, "tr_params", sizeof(double) * std::max<int64_t>(1, B * P), &dpar)))
    return rc;
  if ((rc = workspace(h, "tr_fail", sizeof(int) * B, &dfs))) return rc;
  if (out->waveforms && W > 0 &&
      (rc = workspace(h, "tr_waves", sizeof(double) * W, &dwav)))
    return rc;
  spx_sweep_spec local = *spec;
  if (spec->generator == SPX_GEN_EXPLICIT && P > 0) {
    if ((rc = workspace(h, "tr_explicit", sizeof(double) * B * P, &dexp))) return rc;
    SPX_HIP(hipMemcpyAsync(dexp, spec->explicit_params, sizeof(double) * B * P,
                           hipMemcpyHostToDevice, h->stream),
            SPX_ERR_TRANSFER);
    local.explicit_params = static_cast<double*>(dexp);
  }
  spx_tran_result d{(double*)dx,   (int*)dst,      (int*)dit, (int*)dfl, (double*)dres,
                    (double*)dpar, (double*)dwav,  (int*)dfs};
  if (adaptive) {
    void *dtm = nullptr, *dnp, *dts, *drj, *dmn, *dmx;
    if (out->times && (rc = workspace(h, "tr_times", sizeof(double) * B * pts, &dtm))) return rc;
    if ((rc = workspace(h, "tr_npoints", sizeof(int) * B, &dnp))) return rc;
    if ((rc = workspace(h, "tr_tried", sizeof(int) * B, &dts))) return rc;
    if ((rc = workspace(h, "tr_rejected", sizeof(int) * B, &drj))) return rc;
    if ((rc = workspace(h, "tr_hmin", sizeof(double) * B, &dmn))) return rc;
    if ((rc = workspace(h, "tr_hmax", sizeof(double) * B, &dmx))) return rc;
    d.times = static_cast<double*>(dtm);
    d.num_points = static_cast<int*>(dnp);
    d.total_steps = static_cast<int*>(dts);
    d.rejected_steps = static_cast<int*>(drj);
    d.min_step = static_cast<double*>(dmn);
    d.max_step = static_cast<double*>(dmx);
  }
  if ((rc = transient_device_locked(h, ckt, &local, cfg, tran, &d))) return rc;
  auto d2h = [&](void* dst_h, const void* src_d, size_t bytes) -> int {
    if (!dst_h || !src_d || bytes == 0) return SPX_OK;
    SPX_HIP(hipMemcpyAsync(dst_h, src_d, bytes, hipMemcpyDeviceToHost, h->stream),
            SPX_ERR_TRANSFER);
    return SPX_OK;
  };
  if ((rc = d2h(out->x, dx, sizeof(double) * B * n))) return rc;
  if ((rc = d2h(out->status, dst, sizeof(int) * B))) return rc;
  if ((rc = d2h(out->iterations, dit, sizeof(int) * B))) return rc;
  if ((rc = d2h(out->flags, dfl, sizeof(int) * B))) return rc;
  if ((rc = d2h(out->residual, dres, sizeof(double) * B))) return rc;
  if ((rc = d2h(out->params, dpar, sizeof(double) * B * P))) return rc;
  if ((rc = d2h(out->waveforms, dwav, sizeof(double) * W))) return rc;
  if ((rc = d2h(out->fail_step, dfs, sizeof(int) * B))) return rc;
  if (adaptive) {
    if ((rc = d2h(out->times, d.times, sizeof(double) * B * pts))) return rc;
    if ((rc = d2h(out->num_points, d.num_points, sizeof(int) * B))) return rc;
    if ((rc = d2h(out->total_steps, d.total_steps, sizeof(int) * B))) return rc;
    if ((rc = d2h(out->rejected_steps, d.rejected_steps, sizeof(int) * B))) return rc;
    if ((rc = d2h(out->min_step, d.min_step, sizeof(double) * B))) return rc;
    if ((rc = d2h(out->max_step, d.max_step, sizeof(double) * B))) return rc;
  }
  SPX_HIP(hipStreamSynchronize(h->stream), SPX_ERR_TRANSFER);
  return SPX_OK;
}

int spx_transient_mixed_device(spx_handle* h, const spx_tran_class* classes, int32_t num_classes,
                               const spx_nr_config* cfg, uint32_t options, int32_t* kernels) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  return mixed_locked(h, classes, num_classes, cfg, options, kernels);
}

int spx_stats_reduce_device(spx_handle* h, const double* d_values, const uint8_t* d_mask,
                            int64_t B, int32_t n, double* d_out) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  if (B < 0 || n < 1) return fail(SPX_ERR_INVALID_DIMENSION, "bad statistics shape");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  void* part = nullptr;
  int rc = workspace(h, "stats_part2", sizeof(double) * 5 * n * stats_num_blocks(B), &part);
  if (rc) return rc;
  SPX_HIP(launch_stats(d_values, d_mask, nullptr, B, n, static_cast<double*>(part), d_out,
                       h->stream),
          SPX_ERR_BACKEND);
  return SPX_OK;
}

}  // extern "C"
