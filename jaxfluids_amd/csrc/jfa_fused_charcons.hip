// jfa_fused_charcons.hip -- the fused stage sweeps of CHAR-CONSERVATIVE reconstruction, both storage types.
#define JFA_FUSED_INSTANTIATE
#include "jfa_fused.h"

namespace jfa {

JFA_FUSED_SWEEPS(, double, JFA_CHAR_CONSERVATIVE)
JFA_FUSED_SWEEPS(, float, JFA_CHAR_CONSERVATIVE)

}  // namespace jfa
