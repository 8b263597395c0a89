// Newton kernel instantiation: systems of up to 64 unknowns, FP64 LU with the
// fast pivot comparison (near-ties go to the exact pass, newton_f64x_64.hip).
#include "newton_kernel.h"

namespace spx {
hipError_t launch_newton_f64_64(const DevProgram& g, const DevNrCfg& cfg, const double* params,
                              const double* x0, int64_t B, const DevOut& out, int escalated_only,
                              hipStream_t st) {
  return launch_newton_impl<64, false, false>(g, cfg, params, x0, B, out, escalated_only, st);
}
}  // namespace spx
