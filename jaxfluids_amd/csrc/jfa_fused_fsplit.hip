// jfa_fused_fsplit.hip -- the fused stage sweeps of the flux-splitting scheme (the kFluxSplit face
// flux on the primitive window), every stencil and both storage types.
#define JFA_FUSED_INSTANTIATE_FS
#include "jfa_fused.h"

namespace jfa {

JFA_FUSED_SWEEPS_FS(, double, JFA_WENO5_JS)
JFA_FUSED_SWEEPS_FS(, double, JFA_WENO5_Z)
JFA_FUSED_SWEEPS_FS(, double, kStencilExt)
JFA_FUSED_SWEEPS_FS(, float, JFA_WENO5_JS)
JFA_FUSED_SWEEPS_FS(, float, JFA_WENO5_Z)
JFA_FUSED_SWEEPS_FS(, float, kStencilExt)

}  // namespace jfa
