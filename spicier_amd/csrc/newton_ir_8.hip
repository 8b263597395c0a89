// Newton kernel instantiation: systems of up to 8 unknowns, FP32 LU + FP64 iterative refinement.
#include "newton_kernel.h"

namespace spx {
hipError_t launch_newton_ir_8(const DevProgram& g, const DevNrCfg& cfg, const double* params,
                              const double* x0, int64_t B, const DevOut& out, int escalated_only,
                              hipStream_t st) {
  return launch_newton_impl<8, true, false>(g, cfg, params, x0, B, out, escalated_only, st);
}
}  // namespace spx
