// handle.h — internal: the spx_handle definition and host helpers shared by
// the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spicier_hip.h"
#include "sparse_plan.h"
#include "spectral.h"

struct spx_surrogate;

struct spx_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::map<std::string, std::pair<void*, size_t>> ws;
  spx_surrogate* surrogate = nullptr;  // used by SPX_WARM_SURROGATE
  // circuit programs by slot ("" for single sweeps, one per class of a mixed
  // batch): the program now in workspace "program<slot>" and, for n past the
  // register LU, its shared-symbolic LU plan in "sparse_plan<slot>"
  struct ProgramSlot {
    std::vector<int> blob;
    bool has_sparse = false;
    spx::HostSparsePlan plan;
  };
  std::map<std::string, ProgramSlot> programs;
  // mixed-batch hipGraph: the executable, the description it was captured
  // for, and the device addresses it baked in
  hipGraphExec_t mixed_exec = nullptr;
  std::vector<uint64_t> mixed_key;
  // side streams (default priority) and fork/join events that let the size
  // classes of a mixed batch run concurrently, eagerly or as parallel
  // branches of the captured graph; created on first use
  static constexpr int kSideStreams = 3;
  hipStream_t side[kSideStreams] = {};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev[kSideStreams] = {};
  // autotuned kernel kind per circuit program (transient program blob)
  std::map<std::vector<int>, int> tuned;
  // batched sparse linear solve: the pattern (n, nnz, rows, cols) whose plan
  // is in the "lin_plan" workspace, and that plan's extra offsets
  std::vector<int> lin_key;
  spx::HostSparsePlan lin_plan;
  int lin_num_pos = 0;
  size_t lin_o_off = 0, lin_o_trip = 0;
  // batched GMRES: the pattern whose CSR tables are in the "gm_pattern" workspace
  std::vector<int> gm_key;
  std::vector<int> gm_blob;
  int gm_num_pos = 0;
  size_t gm_o_rowptr = 0, gm_o_col = 0, gm_o_off = 0, gm_o_trip = 0, gm_o_diag = 0;
  // batch spectral analysis: the plan whose tables are in the "sp_tables" workspace
  spx::SpectralPlan spectral;
};

namespace spx {

// Message of the last failing call on this thread (spx_last_error).
std::string& last_error();

inline int fail(int code, const std::string& msg) {
  last_error() = msg;
  return code;
}

#define SPX_HIP(call, code)                                                               \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ::spx::fail(code, std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)

// Grow-only named device buffer of a handle (caller holds the handle's mutex).
inline int workspace(spx_handle* h, const char* name, size_t bytes, void** out) {
  auto& slot = h->ws[name];
  if (slot.second < bytes) {
    if (slot.first) (void)hipFree(slot.first);
    slot = {nullptr, 0};
    size_t want = std::max<size_t>(bytes, 256);
    if (hipMalloc(&slot.first, want) != hipSuccess)
      return fail(SPX_ERR_MEMORY_ALLOC, std::string("hipMalloc of ") + std::to_string(want) +
                                            " bytes for " + name + " failed");
    slot.second = want;
  }
  *out = slot.first;
  return SPX_OK;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    (void)hipGetDevice(&prev);
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (prev >= 0 && cur != prev) (void)hipSetDevice(prev);
  }
};

// Enqueues the surrogate's prediction x0[B][n_x] for params [B][P] on the
// handle's stream (surrogate.hip).
int surrogate_predict_locked(spx_surrogate* s, const double* d_params, int64_t B, int32_t n_x,
                             int32_t P, double* d_x0);

}  // namespace spx
