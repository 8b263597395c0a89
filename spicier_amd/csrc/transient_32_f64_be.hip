// transient_32_f64_be.hip — dense transient kernel, systems of up to 32 unknowns:
// FP64 (fast pivot comparison), Backward Euler fixed steps.
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_FN(tran_32_f64_be, 32, false, false, false)
}  // namespace spx
