// jfa_adjoint_ext.hip -- reverse-mode transpose sweeps of the stencils selected at run time
// (WENO1, WENO3-JS/-Z, TENO5, TENO5-A), instantiated apart from the WENO5 ones so both
// translation units compile in parallel.
#include "jfa_adjoint.h"

namespace jfa {

void adj_sweeps_ext(int var, int riemann, const double* W, const double* lam, const Geo& G, const Params& prm,
                    double ts, double* gW, hipStream_t st) {
    if (var == JFA_CHAR_PRIMITIVE) adj_sweeps_r<kStencilExt, JFA_CHAR_PRIMITIVE>(riemann, W, lam, G, prm, ts, gW, st);
    else if (var == JFA_CONSERVATIVE) adj_sweeps_r<kStencilExt, JFA_CONSERVATIVE>(riemann, W, lam, G, prm, ts, gW, st);
    else if (var == JFA_CHAR_CONSERVATIVE)
        adj_sweeps_r<kStencilExt, JFA_CHAR_CONSERVATIVE>(riemann, W, lam, G, prm, ts, gW, st);
    else adj_sweeps_r<kStencilExt, JFA_PRIMITIVE>(riemann, W, lam, G, prm, ts, gW, st);
}

}  // namespace jfa
