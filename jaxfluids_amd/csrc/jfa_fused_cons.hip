// jfa_fused_cons.hip -- the fused stage sweeps of CONSERVATIVE reconstruction, both storage types.
#define JFA_FUSED_INSTANTIATE
#include "jfa_fused.h"

namespace jfa {

JFA_FUSED_SWEEPS(, double, JFA_CONSERVATIVE)
JFA_FUSED_SWEEPS(, float, JFA_CONSERVATIVE)

}  // namespace jfa
