// transient_8.hip — the dense transient kernels for systems of up to 8 unknowns
// (every variant in this file: they compile quickly at this size).
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_FN(tran_8_f64_be, 8, false, false, false)
SPX_TRAN_VARIANT_FN(tran_8_f64_gen, 8, false, true, false)
SPX_TRAN_VARIANT_FN(tran_8_ir_be, 8, true, false, false)
SPX_TRAN_VARIANT_FN(tran_8_ir_gen, 8, true, true, false)
SPX_TRAN_VARIANT_FN(tran_8_exact_be, 8, false, false, true)
SPX_TRAN_VARIANT_FN(tran_8_exact_gen, 8, false, true, true)

hipError_t launch_transient_8(const DevProgram& g, const DevNrCfg& cfg, const DevTran& tr,
                              const double* params, int64_t B, const DevOut& out,
                              const DevTranOut& tout, hipStream_t st, int phase) {
  static const TranVariantFn v[2][3] = {{tran_8_f64_be, tran_8_ir_be, tran_8_exact_be},
                                       {tran_8_f64_gen, tran_8_ir_gen, tran_8_exact_gen}};
  return launch_transient_variants(v, g, cfg, tr, params, B, out, tout, st, phase);
}
}  // namespace spx
