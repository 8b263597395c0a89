// transient_64_f64_be.hip — dense transient kernel, systems of up to 64 unknowns:
// FP64 (fast pivot comparison), Backward Euler fixed steps.
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_FN(tran_64_f64_be, 64, false, false, false)
}  // namespace spx
