// transient_64_ir_be.hip — dense transient kernel, systems of up to 64 unknowns:
// FP32 LU + FP64 refinement, Backward Euler fixed steps.
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_FN(tran_64_ir_be, 64, true, false, false)
}  // namespace spx
