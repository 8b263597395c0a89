// transient_64_exact_gen.hip — dense transient kernel, systems of up to 64 unknowns:
// FP64 exact-pivot hand-over pass, every method and adaptive stepping.
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_FN(tran_64_exact_gen, 64, false, true, true)
}  // namespace spx
