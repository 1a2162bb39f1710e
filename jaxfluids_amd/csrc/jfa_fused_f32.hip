// jfa_fused_f32.hip -- fused RK3 stage sweeps for float storage (FP64 arithmetic).
#include "jfa_fused.h"

namespace jfa {

int fused_stage_f32(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out, const Geo& G,
                    const jfa_scheme* s, const BC6& bc, int stage, const double* dt_dev, double dt_host, double cfl,
                    double eps, double* dt_next, double* R, double* part, hipStream_t st) {
    return fused_stage_impl<float>(W_in, U_in, U_n, W_out, U_out, G, s, bc, stage, dt_dev, dt_host, cfl, eps, dt_next, R,
                                part, st);
}

int fused_rhs_f32(const void* W, const void* U, const Geo& G, const jfa_scheme* s, double dt, double* R, void* rhs,
                  hipStream_t st) {
    return fused_rhs_impl<float>(W, U, G, s, dt, R, rhs, st);
}

}  // namespace jfa
