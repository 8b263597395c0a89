// transient_64.hip — dispatch of the dense transient kernels for systems of up
// to 64 unknowns; each variant is compiled in its own transient_64_*.hip.
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_DECL(tran_64_f64_be)
SPX_TRAN_VARIANT_DECL(tran_64_f64_gen)
SPX_TRAN_VARIANT_DECL(tran_64_ir_be)
SPX_TRAN_VARIANT_DECL(tran_64_ir_gen)
SPX_TRAN_VARIANT_DECL(tran_64_exact_be)
SPX_TRAN_VARIANT_DECL(tran_64_exact_gen)

hipError_t launch_transient_64(const DevProgram& g, const DevNrCfg& cfg, const DevTran& tr,
                              const double* params, int64_t B, const DevOut& out,
                              const DevTranOut& tout, hipStream_t st, int phase) {
  static const TranVariantFn v[2][3] = {{tran_64_f64_be, tran_64_ir_be, tran_64_exact_be},
                                       {tran_64_f64_gen, tran_64_ir_gen, tran_64_exact_gen}};
  return launch_transient_variants(v, g, cfg, tr, params, B, out, tout, st, phase);
}
}  // namespace spx
