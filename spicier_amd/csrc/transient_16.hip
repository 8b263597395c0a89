// transient_16.hip — instantiation of the Backward-Euler transient kernel for
// systems of up to 16 unknowns.
#include "newton_kernel.h"

namespace spx {
hipError_t launch_transient_16(const DevProgram& g, const DevNrCfg& cfg, const DevTran& tr,
                              const double* params, int64_t B, const DevOut& out,
                              const DevTranOut& tout, hipStream_t st, int phase) {
  return launch_transient_impl<16>(g, cfg, tr, params, B, out, tout, st, phase);
}
}  // namespace spx
