// transient_64_f64_gen.hip — dense transient kernel, systems of up to 64 unknowns:
// FP64 (fast pivot comparison), every method and adaptive stepping.
#include "newton_kernel.h"

namespace spx {
SPX_TRAN_VARIANT_FN(tran_64_f64_gen, 64, false, true, false)
}  // namespace spx
