// Explicit instantiations of the pass launchers of one group (PF_PASS_GROUP
// 0-4, 6 and 7: z-pass operators, 5: column passes) in one precision (PF_PASS_REAL_IS_FLOAT);
// the build compiles this file once per group and precision.
#include "pf_engine.hpp"

namespace pf {
#if PF_PASS_GROUP == 5
#if PF_PASS_REAL_IS_FLOAT
PF_CPASS_DEFINE(float, false)
PF_CPASS_DEFINE(float, true)
#else
PF_CPASS_DEFINE(double, false)
PF_CPASS_DEFINE(double, true)
#endif
#else
#if PF_PASS_REAL_IS_FLOAT
PF_CAT(PF_ZPASS_G, PF_PASS_GROUP)(PF_ZPASS_DEFINE, float)
#else
PF_CAT(PF_ZPASS_G, PF_PASS_GROUP)(PF_ZPASS_DEFINE, double)
#endif
#if PF_PASS_GROUP == 0 || PF_PASS_GROUP == 7
#if PF_PASS_REAL_IS_FLOAT
PF_CAT(PF_ZPASS_TUNED_G, PF_PASS_GROUP)(PF_ZPASS_TUNED_DEFINE, float)
#else
PF_CAT(PF_ZPASS_TUNED_G, PF_PASS_GROUP)(PF_ZPASS_TUNED_DEFINE, double)
#endif
#endif
#endif
}  // namespace pf
