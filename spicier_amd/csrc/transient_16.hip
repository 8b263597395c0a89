// transient_16.hip — the dense transient kernels for systems of up to 16 unknowns
// (every variant in this file: they compile quickly at this size).
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_FN(tran_16_f64_be, 16, false, false, false)
SPX_TRAN_VARIANT_FN(tran_16_f64_gen, 16, false, true, false)
SPX_TRAN_VARIANT_FN(tran_16_ir_be, 16, true, false, false)
SPX_TRAN_VARIANT_FN(tran_16_ir_gen, 16, true, true, false)
SPX_TRAN_VARIANT_FN(tran_16_exact_be, 16, false, false, true)
SPX_TRAN_VARIANT_FN(tran_16_exact_gen, 16, false, true, true)

hipError_t launch_transient_16(const DevProgram& g, const DevNrCfg& cfg, const DevTran& tr,
                              const double* params, int64_t B, const DevOut& out,
                              const DevTranOut& tout, hipStream_t st, int phase) {
  static const TranVariantFn v[2][3] = {{tran_16_f64_be, tran_16_ir_be, tran_16_exact_be},
                                       {tran_16_f64_gen, tran_16_ir_gen, tran_16_exact_gen}};
  return launch_transient_variants(v, g, cfg, tr, params, B, out, tout, st, phase);
}
}  // namespace spx
