// jfa_fused_prim.hip -- the fused stage sweeps of PRIMITIVE reconstruction, both storage types.
#define JFA_FUSED_INSTANTIATE
#include "jfa_fused.h"

namespace jfa {

JFA_FUSED_SWEEPS(, double, JFA_PRIMITIVE)
JFA_FUSED_SWEEPS(, float, JFA_PRIMITIVE)

}  // namespace jfa
