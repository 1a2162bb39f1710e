// jfa_adjoint.hip -- C ABI of the reverse-mode RK3 stage (FP64).
#include <atomic>

#include "jfa_adjoint.h"

namespace jfa {

static std::atomic<int64_t> g_march_min_waves{2048};

int64_t adjoint_march_min_waves() { return g_march_min_waves.load(std::memory_order_relaxed); }

}  // namespace jfa

using namespace jfa;

extern "C" {

int jfa_adjoint_set_march_min_waves(int64_t waves) {
    if (waves < 0) return JFA_EINVAL;
    g_march_min_waves.store(waves, std::memory_order_relaxed);
    return JFA_OK;
}

static size_t vjp_workspace_bytes(const Geo& G, const jfa_scheme* scheme) {
    size_t n = (size_t)(5 * G.fs);
    if (scheme && scheme->transport && (scheme->transport->viscous || scheme->transport->heat))
        n += (size_t)9 * ((size_t)G.n[0] * G.n[1] * G.n[2]);  // X: transverse-derivative cotangent sums
    return n * sizeof(double);
}

size_t jfa_stage_vjp_workspace_bytes(const jfa_grid* grid, const jfa_scheme* scheme) {
    Geo G;
    if (make_geo(grid, G)) return 0;
    return vjp_workspace_bytes(G, scheme);
}

int jfa_rk3_stage_vjp(const void* W_in, const void* lam_out, void* lam_in, void* lam_n, const jfa_grid* grid,
                      const jfa_scheme* scheme, const int32_t bc[6], int32_t stage, double dt, void* workspace,
                      size_t workspace_bytes, void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    rc = check_scheme(scheme);
    if (rc) return rc;
    if (scheme->dtype != JFA_F64) return JFA_EUNSUPPORTED;
    // PRIMITIVE, CONSERVATIVE, CHAR-PRIMITIVE or CHAR-CONSERVATIVE (ARITHMETIC or ROE frozen state)
    if (scheme->recon_var != JFA_PRIMITIVE && scheme->recon_var != JFA_CONSERVATIVE &&
        scheme->recon_var != JFA_CHAR_PRIMITIVE && scheme->recon_var != JFA_CHAR_CONSERVATIVE)
        return JFA_EUNSUPPORTED;
    if (scheme->interpolation_limiter || scheme->flux_limiter) return JFA_EUNSUPPORTED;
    const bool ext = scheme->stencil == JFA_WENO1 || scheme->stencil == JFA_WENO3_JS ||
                     scheme->stencil == JFA_WENO3_Z || scheme->stencil == JFA_TENO5 || scheme->stencil == JFA_TENO5_A;
    if (scheme->stencil != JFA_WENO5_JS && scheme->stencil != JFA_WENO5_Z && !ext) return JFA_EUNSUPPORTED;
    const bool fsplit = scheme->riemann == JFA_FS_ROE || scheme->riemann == JFA_FS_CLLF ||
                        scheme->riemann == JFA_FS_LLF;
    if (scheme->riemann != JFA_HLLC && scheme->riemann != JFA_HLL && scheme->riemann != JFA_HLLC_LM &&
        scheme->riemann != JFA_AUSMP && scheme->riemann != JFA_RUSANOV && !fsplit)
        return JFA_EUNSUPPORTED;
    if (!W_in || !lam_out || !lam_in || !bc || !workspace || stage < 0 || stage > 2 || (stage > 0 && !lam_n))
        return JFA_EINVAL;
    if (lam_in == lam_out || lam_n == lam_out || (stage > 0 && lam_n == lam_in)) return JFA_EINVAL;
    if (!(dt > 0.0) || workspace_bytes < vjp_workspace_bytes(G, scheme)) return JFA_EINVAL;
    BC6 b6;
    for (int f = 0; f < 6; ++f) {
        const int a = f / 2;
        b6.t[f] = bc[f];
        if (!G.active[a]) continue;
        if (bc[f] < 0 || bc[f] > 2) return JFA_EUNSUPPORTED;
        if (bc[f] != JFA_BC_ZEROGRADIENT && G.nh > G.n[a]) return JFA_EINVAL;
    }
    rc = check_transport(scheme, G, bc);  // dissipative fluxes: periodic boundaries
    if (rc) return rc;
    // stage blend U_out = c0 U_in + c1 U_n + m dt R(U_in) (RK3.py:27-62)
    const double c0 = stage == 0 ? 1.0 : (stage == 1 ? 0.25 : 2.0 / 3.0);
    const double c1 = stage == 0 ? 0.0 : (stage == 1 ? 0.75 : 1.0 / 3.0);
    const double m = stage == 0 ? 1.0 : (stage == 1 ? 0.25 : 2.0 / 3.0);
    const Params prm = make_params(scheme);
    hipStream_t st = (hipStream_t)stream;
    double* gW = static_cast<double*>(workspace);
    const int64_t nz = 5 * G.fs;
    const int64_t zb = (nz + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(adj_zero_kernel<0>, dim3((unsigned)(zb < 4096 ? zb : 4096)), dim3(kBlock), 0, st, gW, nz);
    const double* W = static_cast<const double*>(W_in);
    const double* lam = static_cast<const double*>(lam_out);
    const int var = fsplit ? JFA_PRIMITIVE : scheme->recon_var;  // flux splitting: the primitive window
    if (ext) {
        adj_sweeps_ext(var, scheme->riemann, W, lam, G, prm, m * dt, gW, st);
    } else if (scheme->stencil == JFA_WENO5_JS) {
        if (var == JFA_CHAR_PRIMITIVE)
            adj_sweeps_r<JFA_WENO5_JS, JFA_CHAR_PRIMITIVE>(scheme->riemann, W, lam, G, prm, m * dt, gW, st);
        else if (var == JFA_CONSERVATIVE)
            adj_sweeps_r<JFA_WENO5_JS, JFA_CONSERVATIVE>(scheme->riemann, W, lam, G, prm, m * dt, gW, st);
        else if (var == JFA_CHAR_CONSERVATIVE)
            adj_sweeps_r<JFA_WENO5_JS, JFA_CHAR_CONSERVATIVE>(scheme->riemann, W, lam, G, prm, m * dt, gW, st);
        else adj_sweeps_r<JFA_WENO5_JS, JFA_PRIMITIVE>(scheme->riemann, W, lam, G, prm, m * dt, gW, st);
    } else {
        if (var == JFA_CHAR_PRIMITIVE)
            adj_sweeps_r<JFA_WENO5_Z, JFA_CHAR_PRIMITIVE>(scheme->riemann, W, lam, G, prm, m * dt, gW, st);
        else if (var == JFA_CONSERVATIVE)
            adj_sweeps_r<JFA_WENO5_Z, JFA_CONSERVATIVE>(scheme->riemann, W, lam, G, prm, m * dt, gW, st);
        else if (var == JFA_CHAR_CONSERVATIVE)
            adj_sweeps_r<JFA_WENO5_Z, JFA_CHAR_CONSERVATIVE>(scheme->riemann, W, lam, G, prm, m * dt, gW, st);
        else adj_sweeps_r<JFA_WENO5_Z, JFA_PRIMITIVE>(scheme->riemann, W, lam, G, prm, m * dt, gW, st);
    }
    double* X = nullptr;
    if (prm.dissip) {
        X = gW + nz;  // written whole by the cell kernel: no clearing
        hipLaunchKernelGGL(adj_dissip_cell_kernel<0>, flat_grid((int64_t)G.n[0] * G.n[1] * G.n[2]), dim3(kBlock), 0,
                           st, W, lam, G, prm, m * dt, gW, X);
    }
    const int64_t ncell = (int64_t)G.n[0] * G.n[1] * G.n[2];
    hipLaunchKernelGGL(adj_update_kernel<0>, flat_grid(ncell), dim3(kBlock), 0, st, W, lam, gW, G, prm, b6, c0, c1,
                       static_cast<double*>(lam_in), stage > 0 ? static_cast<double*>(lam_n) : nullptr, X);
    return check_launch();
}

}  // extern "C"
