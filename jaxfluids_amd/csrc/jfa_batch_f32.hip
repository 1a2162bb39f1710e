// jfa_batch_f32.hip -- 2-D batched stage kernels for float storage (FP64 arithmetic).
#include "jfa_batch.h"

namespace jfa {

int batch2d_stage_f32(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out,
                      const jfa_scheme* s, int stage, const double* dt_dev, double cfl, double eps, double* dt_next,
                      int ncases, int nitems, char* ws, const BatchLayout& L, hipStream_t st) {
    return batch2d_stage_impl<float>(W_in, U_in, U_n, W_out, U_out, s, stage, dt_dev, cfl, eps, dt_next, ncases,
                                     nitems, ws, L, st);
}

}  // namespace jfa
