// transient_64_exact_be.hip — dense transient kernel, systems of up to 64 unknowns:
// FP64 exact-pivot hand-over pass, Backward Euler fixed steps.
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_FN(tran_64_exact_be, 64, false, false, true)
}  // namespace spx
