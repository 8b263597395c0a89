// transient_64_ir_gen.hip — dense transient kernel, systems of up to 64 unknowns:
// FP32 LU + FP64 refinement, every method and adaptive stepping.
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_FN(tran_64_ir_gen, 64, true, true, false)
}  // namespace spx
