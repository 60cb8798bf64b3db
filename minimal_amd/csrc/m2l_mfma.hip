// m2l_mfma.hip -- M2L grouped by translation key on MFMA (in progress).
#include "fmm_internal.h"

namespace xf {
void m2l_fast(Ctx &, const Ctx &) { fail(EXAFMM_ERR_UNSUPPORTED, "MFMA M2L path not built yet"); }
}  // namespace xf
