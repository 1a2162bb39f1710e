// jfa_batch_f64.hip -- 2-D batched stage kernels for double storage and the batch C ABI.
#include <vector>

#include "jfa_batch.h"

namespace jfa {

int batch2d_stage_f32(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out,
                      const jfa_scheme* s, int stage, const double* dt_dev, double cfl, double eps, double* dt_next,
                      int ncases, int nitems, char* ws, const BatchLayout& L, hipStream_t st);

static int check_cases(const jfa_case2d* cases, int ncases, int chunk) {
    if (!cases || ncases < 1 || chunk < 1 || chunk > 64) return JFA_EINVAL;
    for (int c = 0; c < ncases; ++c) {
        const jfa_case2d& k = cases[c];
        if (k.nh < 3 || k.n[0] < 2 || k.n[1] < 2 || !(k.dx[0] > 0.0) || !(k.dx[1] > 0.0) || k.offset < 0)
            return JFA_EINVAL;
        // the kernels address a case block with 32-bit per-lane byte offsets
        if (5.0 * ((double)k.n[0] + 2 * k.nh) * ((double)k.n[1] + 2 * k.nh) * sizeof(double) >= 4294967296.0)
            return JFA_EINVAL;
        for (int f = 0; f < 4; ++f) {
            if (k.bc[f] < 0 || k.bc[f] > 2) return JFA_EUNSUPPORTED;
            if (k.bc[f] != JFA_BC_ZEROGRADIENT && k.nh > k.n[f / 2]) return JFA_EINVAL;
        }
    }
    return JFA_OK;
}

}  // namespace jfa

using namespace jfa;

extern "C" {

size_t jfa_batch2d_workspace_bytes(const jfa_case2d* cases, int32_t ncases, int32_t chunk) {
    if (check_cases(cases, ncases, chunk)) return 0;
    return batch_layout(cases, ncases, chunk).total;
}

int jfa_batch2d_plan(const jfa_case2d* cases, int32_t ncases, int32_t chunk, void* workspace, size_t workspace_bytes,
                     void* stream) {
    int rc = check_cases(cases, ncases, chunk);
    if (rc) return rc;
    const BatchLayout L = batch_layout(cases, ncases, chunk);
    if (!workspace || workspace_bytes < L.total) return JFA_EINVAL;
    std::vector<Case2D> hc((size_t)ncases);
    std::vector<Item2D> hi;
    hi.reserve((size_t)L.nitems);
    for (int c = 0; c < ncases; ++c) {
        const jfa_case2d& k = cases[c];
        Case2D& d = hc[(size_t)c];
        d.n0 = k.n[0];
        d.n1 = k.n[1];
        d.nh = k.nh;
        for (int f = 0; f < 4; ++f) d.bc[f] = k.bc[f];
        d.one_dx0 = 1.0 / k.dx[0];
        d.one_dx1 = 1.0 / k.dx[1];
        d.dx_min = k.dx[0] < k.dx[1] ? k.dx[0] : k.dx[1];
        d.off = k.offset;
        d.s0 = (int64_t)k.n[1] + 2 * k.nh;
        d.fs = ((int64_t)k.n[0] + 2 * k.nh) * d.s0;
        d.item_begin = (int)hi.size();
        for (int x0 = 0; x0 < k.n[0]; x0 += chunk)
            for (int y0 = 0; y0 < k.n[1]; y0 += 64) {
                Item2D it;
                it.c = c;
                it.x0 = x0;
                it.x1 = x0 + chunk < k.n[0] ? x0 + chunk : k.n[0];
                it.y0 = y0;
                hi.push_back(it);
            }
        d.item_end = (int)hi.size();
    }
    hipStream_t st = (hipStream_t)stream;
    char* ws = (char*)workspace;
    if (hipMemcpyAsync(ws + L.cases_off, hc.data(), sizeof(Case2D) * hc.size(), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(ws + L.items_off, hi.data(), sizeof(Item2D) * hi.size(), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        check_launch();  // records the HIP error text for jfa_last_error
        return JFA_ELAUNCH;
    }
    return JFA_OK;
}

int jfa_batch2d_stage(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out,
                      const jfa_case2d* cases, int32_t ncases, int32_t chunk, const jfa_scheme* scheme, int32_t stage,
                      const double* dt_dev, double cfl, double eps, double* dt_next_dev, void* workspace,
                      size_t workspace_bytes, void* stream) {
    int rc = check_cases(cases, ncases, chunk);
    if (rc) return rc;
    rc = check_scheme(scheme);
    if (rc) return rc;
    if (scheme->stencil == JFA_WENO7_JS) return JFA_EUNSUPPORTED;  // the batched stage reads 6-cell windows
    if (scheme->transport && (scheme->transport->viscous || scheme->transport->heat)) return JFA_EUNSUPPORTED;
    if (!W_in || !U_in || !W_out || !U_out || !dt_dev || !workspace || stage < 0 || stage > 2 || (stage > 0 && !U_n))
        return JFA_EINVAL;
    if (W_in == W_out || U_in == U_out || W_out == U_in || U_out == W_in) return JFA_EINVAL;
    if (dt_next_dev && !(cfl > 0.0)) return JFA_EINVAL;
    const BatchLayout L = batch_layout(cases, ncases, chunk);
    if (workspace_bytes < L.total) return JFA_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    if (scheme->dtype == JFA_F32)
        return batch2d_stage_f32(W_in, U_in, U_n, W_out, U_out, scheme, stage, dt_dev, cfl, eps, dt_next_dev, ncases,
                                 L.nitems, (char*)workspace, L, st);
    return batch2d_stage_impl<double>(W_in, U_in, U_n, W_out, U_out, scheme, stage, dt_dev, cfl, eps, dt_next_dev,
                                      ncases, L.nitems, (char*)workspace, L, st);
}

}  // extern "C"
