// transient_32_exact_be.hip — dense transient kernel, systems of up to 32 unknowns:
// FP64 exact-pivot hand-over pass, Backward Euler fixed steps.
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_FN(tran_32_exact_be, 32, false, false, true)
}  // namespace spx
