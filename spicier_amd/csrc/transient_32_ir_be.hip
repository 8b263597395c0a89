// transient_32_ir_be.hip — dense transient kernel, systems of up to 32 unknowns:
// FP32 LU + FP64 refinement, Backward Euler fixed steps.
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_FN(tran_32_ir_be, 32, true, false, false)
}  // namespace spx
