// jfa_fused_charprim.hip -- the fused stage sweeps of CHAR-PRIMITIVE reconstruction, both storage types.
#define JFA_FUSED_INSTANTIATE
#include "jfa_fused.h"

namespace jfa {

JFA_FUSED_SWEEPS(, double, JFA_CHAR_PRIMITIVE)
JFA_FUSED_SWEEPS(, float, JFA_CHAR_PRIMITIVE)

}  // namespace jfa
