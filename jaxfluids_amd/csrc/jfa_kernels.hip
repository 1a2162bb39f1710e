// jfa_kernels.hip -- HIP kernels and the extern "C" ABI of jaxfluids_amd (gfx950).
//
// Buffers are structure-of-arrays fields in the JAX-Fluids layout
// (nvar, Nx+2nh, Ny+2nh, Nz+2nh), z contiguous, so a wavefront of 64 lanes walking z
// issues 512-byte coalesced FP64 loads per field.  Storage type T is double or float;
// all arithmetic is FP64.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "jfa_common.h"
#include "jfa_dissipative.h"

namespace jfa {

thread_local char g_last_error[256] = "";


// ------------------------------------------------------------------ reconstruct_xi
template <class T, int STENCIL>
__global__ __launch_bounds__(kBlock) void reconstruct_kernel(const T* __restrict__ buf, int nvar, Geo G, int axis,
                                                             int j, Params prm, FaceIdx F, T* __restrict__ out) {
    int f0, f1, f2;
    if (!flat_index(F.nface, F.fe[1], F.fe[2], f0, f1, f2)) return;
    const int64_t base = (int64_t)(f0 + G.off[0]) * G.s[0] + (int64_t)(f1 + G.off[1]) * G.s[1] + (f2 + G.off[2]);
    const int64_t sa = G.s[axis];
    const int64_t o = ((int64_t)f0 * F.fe[1] + f1) * F.fe[2] + f2;
    for (int v = 0; v < nvar; ++v) {
        const T* b = buf + v * G.fs + base;
        double r;
        if constexpr (STENCIL == kStencilW7) {  // weno7_js.py:31: cells -4..2 and 3..-3
            if (j == 0)
                r = weno7_js(ld(b, -4 * sa), ld(b, -3 * sa), ld(b, -2 * sa), ld(b, -sa), ld(b, 0), ld(b, sa), ld(b, 2 * sa), prm.eps);
            else
                r = weno7_js(ld(b, 3 * sa), ld(b, 2 * sa), ld(b, sa), ld(b, 0), ld(b, -sa), ld(b, -2 * sa), ld(b, -3 * sa), prm.eps);
        } else if (j == 0)  // cells -3..2 (weno6_base.py:27; the 5-point and narrower stencils use -3..1 and inner cells)
            r = recon_side<STENCIL>(ld(b, -3 * sa), ld(b, -2 * sa), ld(b, -sa), ld(b, 0), ld(b, sa), ld(b, 2 * sa), prm, prm.dx2[axis]);
        else  // cells 2..-3
            r = recon_side<STENCIL>(ld(b, 2 * sa), ld(b, sa), ld(b, 0), ld(b, -sa), ld(b, -2 * sa), ld(b, -3 * sa), prm, prm.dx2[axis]);
        out[(int64_t)v * F.nface + o] = static_cast<T>(r);
    }
}

// ------------------------------------------------------------------ Riemann on face arrays
template <class T, int AXIS, int RIEMANN>
__global__ __launch_bounds__(kBlock) void riemann_kernel(const T* __restrict__ WL_, const T* __restrict__ WR_,
                                                         const T* __restrict__ UL_, const T* __restrict__ UR_,
                                                         int64_t n, Params prm, T* __restrict__ flux) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double WL[5], WR[5], UL[5], UR[5], Fo[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        WL[i] = ld(WL_, i * n + t);
        WR[i] = ld(WR_, i * n + t);
        UL[i] = ld(UL_, i * n + t);
        UR[i] = ld(UR_, i * n + t);
    }
    riemann<AXIS, RIEMANN>(WL, WR, UL, UR, prm, Fo);
#pragma unroll
    for (int i = 0; i < 5; ++i) flux[i * n + t] = static_cast<T>(Fo[i]);
}

// ------------------------------------------------------------------ signal speeds (plain functions)
template <class T>
__global__ __launch_bounds__(kBlock) void signal_speed_kernel(int kind, const T* __restrict__ uL,
                                                              const T* __restrict__ uR, const T* __restrict__ aL,
                                                              const T* __restrict__ aR, const T* __restrict__ rL,
                                                              const T* __restrict__ rR, const T* __restrict__ pL,
                                                              const T* __restrict__ pR, const T* __restrict__ HL,
                                                              const T* __restrict__ HR, double gamma, int64_t n,
                                                              T* __restrict__ SL, T* __restrict__ SR) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double sl, sr;
    if (kind == JFA_SS_DAVIS2) {
        signal_speeds_davis2(ld(uL, t), ld(uR, t), ld(rL, t), ld(rR, t), ld(HL, t), ld(HR, t), gamma, sl, sr);
    } else {
        const bool rho = kind == JFA_SS_EINFELDT || kind == JFA_SS_TORO;
        const bool p = kind == JFA_SS_TORO;
        signal_speeds(kind, ld(uL, t), ld(uR, t), ld(aL, t), ld(aR, t), rho ? ld(rL, t) : 0.0,
                      rho ? ld(rR, t) : 0.0, p ? ld(pL, t) : 0.0, p ? ld(pR, t) : 0.0, gamma, sl, sr);
    }
    SL[t] = static_cast<T>(sl);
    SR[t] = static_cast<T>(sr);
}

// ------------------------------------------------------------------ face flux (compute_flux_xi)
// One thread per face.  Loads the 6-cell window of the face along AXIS and evaluates
// reconstruction + Riemann in registers.
template <class T, int AXIS, int VAR, int STENCIL, int RIEMANN>
__device__ __forceinline__ int face_flux_at(const T* __restrict__ W, const T* __restrict__ U, int64_t base,
                                            int64_t sa, int64_t fs, const Params& prm, double Fo[5]) {
    constexpr int NW = Win<STENCIL>::n, H = Win<STENCIL>::h;
    double win[NW][5];
    double wl[5], wr[5];
    constexpr bool cons_win = (VAR == JFA_CONSERVATIVE || VAR == JFA_CHAR_CONSERVATIVE);
    const T* src = cons_win ? U : W;
#pragma unroll
    for (int k = 0; k < NW; ++k)
#pragma unroll
        for (int i = 0; i < 5; ++i) win[k][i] = ld(src, i * fs + base + (k - H) * sa);
    if constexpr (RIEMANN == kFluxSplit) {
        if constexpr (STENCIL != kStencilW7) flux_split_face<AXIS, STENCIL>(win, prm, Fo);
        return 0;
    }
    if constexpr (VAR == JFA_CHAR_CONSERVATIVE) {
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            wl[i] = ld(W, i * fs + base - sa);
            wr[i] = ld(W, i * fs + base);
        }
    }
    double WL[5], WR[5], UL[5], UR[5];
    face_states<AXIS, VAR, STENCIL>(win, wl, wr, prm, WL, WR, UL, UR);
    int nlim = 0;
    if (prm.limiter) {
        double c[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) c[i] = ld(W, i * fs + base - sa);
        nlim = limit_side<VAR>(c, prm, WL, UL);
#pragma unroll
        for (int i = 0; i < 5; ++i) c[i] = ld(W, i * fs + base);
        nlim += limit_side<VAR>(c, prm, WR, UR);
    }
    riemann<AXIS, RIEMANN>(WL, WR, UL, UR, prm, Fo);
    return nlim;
}

template <class T, int AXIS, int VAR, int STENCIL, int RIEMANN>
__global__ __launch_bounds__(kBlock) void flux_xi_kernel(const T* __restrict__ W, const T* __restrict__ U, Geo G,
                                                         FaceIdx F, Params prm, T* __restrict__ flux,
                                                         unsigned long long* __restrict__ nlim_total) {
    int f0, f1, f2;
    if (!flat_index(F.nface, F.fe[1], F.fe[2], f0, f1, f2)) return;
    const int64_t base = (int64_t)(f0 + G.off[0]) * G.s[0] + (int64_t)(f1 + G.off[1]) * G.s[1] + (f2 + G.off[2]);
    double Fo[5];
    const int nlim = face_flux_at<T, AXIS, VAR, STENCIL, RIEMANN>(W, U, base, G.s[AXIS], G.fs, prm, Fo);
    // limiter replacements are rare: one global atomic per face that has any
    if (nlim && nlim_total) atomicAdd(nlim_total, (unsigned long long)nlim);
    const int64_t o = ((int64_t)f0 * F.fe[1] + f1) * F.fe[2] + f2;
#pragma unroll
    for (int i = 0; i < 5; ++i) flux[(int64_t)i * F.nface + o] = static_cast<T>(Fo[i]);
}

// ------------------------------------------------------------------ RK3 stage

template <class T>
__global__ __launch_bounds__(kBlock) void rk3_stage_kernel(T* __restrict__ U, const T* __restrict__ Un,
                                                           const T* __restrict__ rhs, T* __restrict__ W, Geo G,
                                                           int stage, const double* __restrict__ dt_dev,
                                                           double dt_host, Params prm) {
    const int64_t plane = (int64_t)G.n[1] * G.n[2];
    int c0i, c1i, c2i;
    if (!flat_index((int64_t)G.n[0] * plane, G.n[1], G.n[2], c0i, c1i, c2i)) return;
    const int64_t t = (int64_t)c1i * G.n[2] + c2i;
    const int64_t b = (int64_t)(c0i + G.off[0]) * G.s[0] + (int64_t)(c1i + G.off[1]) * G.s[1] + (c2i + G.off[2]);
    const int64_t ncell = (int64_t)G.n[0] * plane;
    const int64_t ro = (int64_t)c0i * plane + t;
    double c0, c1, m;
    rk3_coeffs(stage, c0, c1, m);
    const double dt = dt_dev ? dt_dev[0] : dt_host;
    const double ts = dt * m;  // RK3.py:60
    double u[5], w[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        double x = ld(U, i * G.fs + b);
        if (stage > 0) x = c0 * x + c1 * ld(Un, i * G.fs + b);  // RK3.py:49-50
        x = x + ts * ld(rhs, i * ncell + ro);                    // time_integrator.py:57
        u[i] = x;
    }
    prim_from_cons(u, prm, w);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        U[i * G.fs + b] = static_cast<T>(u[i]);
        W[i * G.fs + b] = static_cast<T>(w[i]);
    }
}

// ------------------------------------------------------------------ halos
// One thread per face-halo cell of one face location.  W halo from the retrieve slice
// (boundary_condition.py:564-595), U halo = cons(W halo) (outer/material.py:248-250).
template <class T>
__global__ __launch_bounds__(kBlock) void halo_kernel(T* __restrict__ W, T* __restrict__ U, Geo G, int axis,
                                                      int high, int bc, Params prm) {
    // halo slab: nh along axis, interior along the others
    int ext[3];
    for (int i = 0; i < 3; ++i) ext[i] = i == axis ? G.nh : G.n[i];
    const int64_t total = (int64_t)ext[0] * ext[1] * ext[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int c[3];
    c[2] = (int)(t % ext[2]);
    c[1] = (int)((t / ext[2]) % ext[1]);
    c[0] = (int)(t / ((int64_t)ext[1] * ext[2]));
    const int m = c[axis];  // 0..nh-1 within the slab
    const int N = G.n[axis], nh = G.nh;
    int dst, src;
    if (high) {
        dst = N + nh + m;
        if (bc == JFA_BC_PERIODIC) src = nh + m;
        else if (bc == JFA_BC_ZEROGRADIENT) src = N + nh - 1;
        else src = N + nh - 1 - m;
    } else {
        dst = m;
        if (bc == JFA_BC_PERIODIC) src = N + m;
        else if (bc == JFA_BC_ZEROGRADIENT) src = nh;
        else src = 2 * nh - 1 - m;
    }
    int64_t bd = 0, bs = 0;
    for (int i = 0; i < 3; ++i) {
        const int pd = i == axis ? dst : c[i] + G.off[i];
        const int ps = i == axis ? src : c[i] + G.off[i];
        bd += (int64_t)pd * G.s[i];
        bs += (int64_t)ps * G.s[i];
    }
    double w[5], u[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) w[i] = ld(W, i * G.fs + bs);
    if (bc == JFA_BC_SYMMETRY) w[axis + 1] = -w[axis + 1];  // boundary_condition.py:721-731
    cons_from_prim(w, prm, u);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        W[i * G.fs + bd] = static_cast<T>(w[i]);
        if (U) U[i * G.fs + bd] = static_cast<T>(u[i]);
    }
}


// grid-stride over the interior cells: at most kCflBlocks atomics per reduction and slot.
// slot[0]: max sum(|u_i| + c); with dissipative fluxes slot[1]: max mu/rho, slot[2]: max kappa/(rho cp).
template <class T>
__global__ __launch_bounds__(kBlock) void cfl_max_kernel(const T* __restrict__ W, Geo G, Params prm,
                                                         unsigned long long* slot) {
    const uint32_t total = (uint32_t)G.n[0] * G.n[1] * G.n[2];
    double v = 0.0, vn = 0.0, va = 0.0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const uint32_t q = t / (uint32_t)G.n[2];
        const int c2i = (int)(t - q * (uint32_t)G.n[2]);
        const int c0i = (int)(q / (uint32_t)G.n[1]);
        const int c1i = (int)(q - (uint32_t)c0i * (uint32_t)G.n[1]);
        const int64_t b =
            (int64_t)(c0i + G.off[0]) * G.s[0] + (int64_t)(c1i + G.off[1]) * G.s[1] + (c2i + G.off[2]);
        double w[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) w[i] = ld(W, i * G.fs + b);
        v = nan_max(v, wave_sum(w, G.active, prm.gamma));
        if (prm.dissip) {
            double nu, al;
            diffusivities(w, prm, nu, al);
            vn = nan_max(vn, nu);
            va = nan_max(va, al);
        }
    }
    block_max_to(slot, v);
    if (prm.dissip) {
        __syncthreads();  // block_max_to's shared partials are reused
        block_max_to(slot + 1, vn);
        __syncthreads();
        block_max_to(slot + 2, va);
    }
}

__global__ void zero_slot_kernel(unsigned long long* slot) {
    slot[0] = 0ull;
    slot[1] = 0ull;
    slot[2] = 0ull;
}

// time_step_size.py:103-135,154-155
__global__ void dt_finalize_kernel(const unsigned long long* slot, double dx_min, double cfl, double eps, int dissip,
                                   double* dt) {
    double d = dx_min / (__longlong_as_double((long long)slot[0]) + eps);
    if (dissip & 1) d = fmin(d, 3.0 / 14.0 * dx_min * dx_min / (__longlong_as_double((long long)slot[1]) + eps));
    if (dissip & 2) d = fmin(d, 0.1 * dx_min * dx_min / (__longlong_as_double((long long)slot[2]) + eps));
    dt[0] = d * cfl;
}

// ------------------------------------------------------------------ whole-buffer conversions
template <class T, bool TO_CONS>
__global__ __launch_bounds__(kBlock) void convert_kernel(const T* __restrict__ in, T* __restrict__ out, int64_t n,
                                                         Params prm) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double a[5], b[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) a[i] = ld(in, i * n + t);
    if (TO_CONS)
        cons_from_prim(a, prm, b);
    else
        prim_from_cons(a, prm, b);
#pragma unroll
    for (int i = 0; i < 5; ++i) out[i * n + t] = static_cast<T>(b[i]);
}

// ------------------------------------------------------------------ fused stage epilogue

// ------------------------------------------------------------------ dispatch tables
using FluxLaunch = void (*)(const void*, const void*, const Geo&, const FaceIdx&, const Params&, void*,
                           unsigned long long*, hipStream_t);

template <class T, int AXIS, int VAR, int STENCIL, int RIEMANN>
static void launch_flux(const void* W, const void* U, const Geo& G, const FaceIdx& F, const Params& prm, void* out,
                        unsigned long long* nlim, hipStream_t st) {
    dim3 grid = flat_grid(F.nface);
    hipLaunchKernelGGL((flux_xi_kernel<T, AXIS, VAR, STENCIL, RIEMANN>), grid, dim3(kBlock), 0, st,
                       static_cast<const T*>(W), static_cast<const T*>(U), G, F, prm, static_cast<T*>(out), nlim);
}

template <class T, int AXIS, int VAR, int STENCIL>
static FluxLaunch pick_r(int riemann) {
    if (riemann == JFA_HLLC) return &launch_flux<T, AXIS, VAR, STENCIL, JFA_HLLC>;
    if (riemann >= JFA_FS_ROE) {  // flux splitting reads the primitive window: run_flux sends it there
        if constexpr (VAR == JFA_PRIMITIVE && STENCIL != kStencilW7) return &launch_flux<T, AXIS, VAR, STENCIL, kFluxSplit>;
        return nullptr;
    }
    if (riemann != JFA_RUSANOV) return &launch_flux<T, AXIS, VAR, STENCIL, kRiemannExt>;
    return &launch_flux<T, AXIS, VAR, STENCIL, JFA_RUSANOV>;
}
template <class T, int AXIS, int VAR>
static FluxLaunch pick_s(int stencil, int riemann) {
    if (stencil == JFA_WENO5_JS) return pick_r<T, AXIS, VAR, JFA_WENO5_JS>(riemann);
    if (stencil == JFA_WENO5_Z) return pick_r<T, AXIS, VAR, JFA_WENO5_Z>(riemann);
    if (stencil == JFA_WENO7_JS) return pick_r<T, AXIS, VAR, kStencilW7>(riemann);
    return pick_r<T, AXIS, VAR, kStencilExt>(riemann);
}
template <class T, int AXIS>
static FluxLaunch pick_v(int var, int stencil, int riemann) {
    switch (var) {
        case JFA_PRIMITIVE: return pick_s<T, AXIS, JFA_PRIMITIVE>(stencil, riemann);
        case JFA_CONSERVATIVE: return pick_s<T, AXIS, JFA_CONSERVATIVE>(stencil, riemann);
        case JFA_CHAR_PRIMITIVE: return pick_s<T, AXIS, JFA_CHAR_PRIMITIVE>(stencil, riemann);
        default: return pick_s<T, AXIS, JFA_CHAR_CONSERVATIVE>(stencil, riemann);
    }
}
template <class T>
static FluxLaunch pick_flux(int axis, int var, int stencil, int riemann) {
    switch (axis) {
        case 0: return pick_v<T, 0>(var, stencil, riemann);
        case 1: return pick_v<T, 1>(var, stencil, riemann);
        default: return pick_v<T, 2>(var, stencil, riemann);
    }
}


static int run_flux(const void* W, const void* U, const Geo& G, const jfa_scheme* s, int axis, void* flux,
                    int64_t* nlim, hipStream_t st) {
    const FaceIdx F = face_extents(G, axis);
    Params prm = make_params(s);
    set_cell_sizes(prm, G);
    const int var = s->riemann >= JFA_FS_ROE ? JFA_PRIMITIVE : s->recon_var;
    FluxLaunch fn = s->dtype == JFA_F32 ? pick_flux<float>(axis, var, s->stencil, s->riemann)
                                        : pick_flux<double>(axis, var, s->stencil, s->riemann);
    fn(W, U, G, F, prm, flux, prm.limiter ? reinterpret_cast<unsigned long long*>(nlim) : nullptr, st);
    return check_launch();
}

}  // namespace jfa

using namespace jfa;

// ==================================================================== C ABI
extern "C" {

const char* jfa_version(void) { return "jaxfluids_amd 0.1.0 gfx950"; }

const char* jfa_last_error(void) { return g_last_error; }

int jfa_weno5_reconstruct_xi(const void* buffer, int32_t nvar, const jfa_grid* grid, int32_t axis, int32_t j,
                             int32_t stencil, double eps, int32_t dtype, void* out, void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    if (!buffer || !out || nvar < 1 || axis < 0 || axis > 2 || !G.active[axis] || (j != 0 && j != 1))
        return JFA_EINVAL;
    if (stencil < JFA_WENO5_JS || stencil > JFA_WENO7_JS) return JFA_EUNSUPPORTED;
    if (stencil == JFA_WENO7_JS && G.nh < 4) return JFA_EINVAL;
    const FaceIdx F = face_extents(G, axis);
    dim3 g = flat_grid(F.nface);
    hipStream_t st = (hipStream_t)stream;
    Params prm = gas_params(1.4);  // only the stencil fields are read
    prm.eps = eps;
    prm.eps_z = 4.0 * eps;
    prm.eps_js = 16.0 * eps;
    prm.stencil = stencil;
    set_cell_sizes(prm, G);
#define JFA_W(T) \
    do { \
        if (stencil == JFA_WENO5_JS) \
            hipLaunchKernelGGL((reconstruct_kernel<T, JFA_WENO5_JS>), g, dim3(kBlock), 0, st, (const T*)buffer, nvar, G, axis, j, prm, F, (T*)out); \
        else if (stencil == JFA_WENO5_Z) \
            hipLaunchKernelGGL((reconstruct_kernel<T, JFA_WENO5_Z>), g, dim3(kBlock), 0, st, (const T*)buffer, nvar, G, axis, j, prm, F, (T*)out); \
        else if (stencil == JFA_WENO7_JS) \
            hipLaunchKernelGGL((reconstruct_kernel<T, kStencilW7>), g, dim3(kBlock), 0, st, (const T*)buffer, nvar, G, axis, j, prm, F, (T*)out); \
        else \
            hipLaunchKernelGGL((reconstruct_kernel<T, kStencilExt>), g, dim3(kBlock), 0, st, (const T*)buffer, nvar, G, axis, j, prm, F, (T*)out); \
    } while (0)
    if (dtype == JFA_F32) JFA_W(float);
    else if (dtype == JFA_F64) JFA_W(double);
    else return JFA_EUNSUPPORTED;
#undef JFA_W
    return check_launch();
}

int jfa_riemann_xi(const void* W_L, const void* W_R, const void* U_L, const void* U_R, int64_t nfaces, int32_t axis,
                   const jfa_scheme* scheme, void* flux, void* stream) {
    int rc = check_scheme(scheme);
    if (rc) return rc;
    if (scheme->riemann >= JFA_FS_ROE) return JFA_EUNSUPPORTED;  // flux splitting has no face states
    if (!W_L || !W_R || !U_L || !U_R || !flux || nfaces < 0 || axis < 0 || axis > 2) return JFA_EINVAL;
    if (nfaces == 0) return JFA_OK;
    const Params prm = make_params(scheme);
    dim3 g((unsigned)((nfaces + kBlock - 1) / kBlock));
    hipStream_t st = (hipStream_t)stream;
#define JFA_R(T, A, R) \
    hipLaunchKernelGGL((riemann_kernel<T, A, R>), g, dim3(kBlock), 0, st, (const T*)W_L, (const T*)W_R, (const T*)U_L, (const T*)U_R, nfaces, prm, (T*)flux)
#define JFA_RA(T, R) \
    do { if (axis == 0) JFA_R(T, 0, R); else if (axis == 1) JFA_R(T, 1, R); else JFA_R(T, 2, R); } while (0)
    if (scheme->dtype == JFA_F32) {
        if (scheme->riemann == JFA_HLLC) JFA_RA(float, JFA_HLLC);
        else if (scheme->riemann != JFA_RUSANOV) JFA_RA(float, kRiemannExt);
        else JFA_RA(float, JFA_RUSANOV);
    } else {
        if (scheme->riemann == JFA_HLLC) JFA_RA(double, JFA_HLLC);
        else if (scheme->riemann != JFA_RUSANOV) JFA_RA(double, kRiemannExt);
        else JFA_RA(double, JFA_RUSANOV);
    }
#undef JFA_RA
#undef JFA_R
    return check_launch();
}

int jfa_compute_flux_xi(const void* W, const void* U, const jfa_grid* grid, const jfa_scheme* scheme, int32_t axis,
                        void* flux, int64_t* limiter_count, void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    rc = check_scheme(scheme);
    if (rc) return rc;
    if (!W || !U || !flux || axis < 0 || axis > 2 || !G.active[axis]) return JFA_EINVAL;
    rc = check_stencil_halos(scheme, G);
    if (rc) return rc;
    return run_flux(W, U, G, scheme, axis, flux, limiter_count, (hipStream_t)stream);
}

size_t jfa_rhs_workspace_bytes(const jfa_grid* grid, int32_t dtype) {
    Geo G;
    if (make_geo(grid, G) || (dtype != JFA_F64 && dtype != JFA_F32)) return 0;
    return 5 * (size_t)G.n[0] * G.n[1] * G.n[2] * sizeof(double) + 256;  // partial RHS + WAVESPEED slots
}

int jfa_compute_rhs(const void* W, const void* U, const jfa_grid* grid, const jfa_scheme* scheme, double dt,
                    void* rhs, void* workspace, size_t workspace_bytes, void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    rc = check_scheme(scheme);
    if (rc) return rc;
    if (!W || !U || !rhs || !workspace) return JFA_EINVAL;
    if (workspace_bytes < jfa_rhs_workspace_bytes(grid, scheme->dtype)) return JFA_EINVAL;
    rc = check_stencil_halos(scheme, G);
    if (rc) return rc;
    rc = check_transport(scheme, G, nullptr);  // transport->halos says whether the caller's halos are read
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (scheme->flux_limiter && !(dt > 0.0)) return JFA_EINVAL;  // the flux limiter needs the step's dt
    if (scheme->dtype == JFA_F32) return fused_rhs_f32(W, U, G, scheme, dt, (double*)workspace, rhs, st);
    return fused_rhs_f64(W, U, G, scheme, dt, (double*)workspace, rhs, st);
}

int jfa_rk3_stage(void* U, const void* U_n, const void* rhs, void* W, const jfa_grid* grid, const jfa_scheme* scheme,
                  int32_t stage, const double* dt_dev, double dt_host, void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    rc = check_scheme(scheme);
    if (rc) return rc;
    if (!U || !rhs || !W || stage < 0 || stage > 2 || (stage > 0 && !U_n)) return JFA_EINVAL;
    const int64_t plane = (int64_t)G.n[1] * G.n[2];
    dim3 g = flat_grid((int64_t)G.n[0] * plane);
    hipStream_t st = (hipStream_t)stream;
    if (scheme->dtype == JFA_F32)
        hipLaunchKernelGGL((rk3_stage_kernel<float>), g, dim3(kBlock), 0, st, (float*)U, (const float*)U_n,
                           (const float*)rhs, (float*)W, G, stage, dt_dev, dt_host, make_params(scheme));
    else
        hipLaunchKernelGGL((rk3_stage_kernel<double>), g, dim3(kBlock), 0, st, (double*)U, (const double*)U_n,
                           (const double*)rhs, (double*)W, G, stage, dt_dev, dt_host, make_params(scheme));
    return check_launch();
}

int jfa_halo_update(void* W, void* U, const jfa_grid* grid, const int32_t bc[6], double gamma, int32_t dtype,
                    void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    if (!W || !bc) return JFA_EINVAL;
    if (dtype != JFA_F64 && dtype != JFA_F32) return JFA_EUNSUPPORTED;
    hipStream_t st = (hipStream_t)stream;
    // halo_slices.py:66-73 face order east, west, north, south, top, bottom
    for (int loc = 0; loc < 6; ++loc) {
        const int axis = loc / 2;
        if (!G.active[axis]) continue;
        const int b = bc[loc];
        if (b < 0 || b > 2) return JFA_EUNSUPPORTED;
        if (G.nh > G.n[axis] && b != JFA_BC_ZEROGRADIENT) return JFA_EINVAL;
        const int high = (loc % 2) == 0;
        int64_t total = G.nh;
        for (int i = 0; i < 3; ++i)
            if (i != axis) total *= G.n[i];
        dim3 g((unsigned)((total + kBlock - 1) / kBlock));
        if (dtype == JFA_F32)
            hipLaunchKernelGGL((halo_kernel<float>), g, dim3(kBlock), 0, st, (float*)W, (float*)U, G, axis, high, b,
                               gas_params(gamma));
        else
            hipLaunchKernelGGL((halo_kernel<double>), g, dim3(kBlock), 0, st, (double*)W, (double*)U, G, axis, high,
                               b, gas_params(gamma));
        rc = check_launch();
        if (rc) return rc;
    }
    return JFA_OK;
}

int jfa_edge_halo_update(void* W, void* U, const jfa_grid* grid, const int32_t bc[6], double gamma,
                         int32_t dtype, void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    if (!W || !bc || !(gamma > 1.0)) return JFA_EINVAL;
    if (dtype != JFA_F64 && dtype != JFA_F32) return JFA_EUNSUPPORTED;
    BC6 b6;
    for (int f = 0; f < 6; ++f) {
        b6.t[f] = bc[f];
        if (!G.active[f / 2]) continue;
        if (bc[f] < 0 || bc[f] > 2) return JFA_EUNSUPPORTED;
        if (G.nh > G.n[f / 2] && bc[f] != JFA_BC_ZEROGRADIENT) return JFA_EINVAL;
    }
    hipStream_t st = (hipStream_t)stream;
    if (dtype == JFA_F32) launch_edge_halos<float>((float*)W, (float*)U, G, b6, gas_params(gamma), st);
    else launch_edge_halos<double>((double*)W, (double*)U, G, b6, gas_params(gamma), st);
    return check_launch();
}

int jfa_compute_dt(const void* W, const jfa_grid* grid, double gamma, double cfl, double eps, int32_t dtype,
                   const jfa_transport* transport, double* dt_dev, void* workspace, void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    if (!W || !dt_dev || !workspace || !(gamma > 1.0)) return JFA_EINVAL;
    if (dtype != JFA_F64 && dtype != JFA_F32) return JFA_EUNSUPPORTED;
    jfa_scheme s = {};
    s.gamma = gamma;
    s.transport = transport;
    rc = check_transport(&s, G, nullptr);
    if (rc) return rc;
    Params prm = gas_params(gamma);
    set_transport(prm, transport, gamma);
    hipStream_t st = (hipStream_t)stream;
    unsigned long long* slot = (unsigned long long*)workspace;
    hipLaunchKernelGGL(zero_slot_kernel, dim3(1), dim3(1), 0, st, slot);
    const int64_t ncell = (int64_t)G.n[0] * G.n[1] * G.n[2];
    const int64_t nblk = (ncell + kBlock - 1) / kBlock;
    dim3 g((unsigned)(nblk < kCflBlocks ? nblk : kCflBlocks));
    if (dtype == JFA_F32)
        hipLaunchKernelGGL((cfl_max_kernel<float>), g, dim3(kBlock), 0, st, (const float*)W, G, prm, slot);
    else
        hipLaunchKernelGGL((cfl_max_kernel<double>), g, dim3(kBlock), 0, st, (const double*)W, G, prm, slot);
    hipLaunchKernelGGL(dt_finalize_kernel, dim3(1), dim3(1), 0, st, slot, G.dx_min, cfl, eps, prm.dissip, dt_dev);
    return check_launch();
}

int jfa_cons_from_prim(const void* W, void* U, int64_t ncells, double gamma, int32_t dtype, void* stream) {
    if (!W || !U || ncells < 0) return JFA_EINVAL;
    if (ncells == 0) return JFA_OK;
    dim3 g((unsigned)((ncells + kBlock - 1) / kBlock));
    hipStream_t st = (hipStream_t)stream;
    if (dtype == JFA_F32)
        hipLaunchKernelGGL((convert_kernel<float, true>), g, dim3(kBlock), 0, st, (const float*)W, (float*)U, ncells, gas_params(gamma));
    else if (dtype == JFA_F64)
        hipLaunchKernelGGL((convert_kernel<double, true>), g, dim3(kBlock), 0, st, (const double*)W, (double*)U, ncells, gas_params(gamma));
    else
        return JFA_EUNSUPPORTED;
    return check_launch();
}

int jfa_prim_from_cons(const void* U, void* W, int64_t ncells, double gamma, int32_t dtype, void* stream) {
    if (!W || !U || ncells < 0) return JFA_EINVAL;
    if (ncells == 0) return JFA_OK;
    dim3 g((unsigned)((ncells + kBlock - 1) / kBlock));
    hipStream_t st = (hipStream_t)stream;
    if (dtype == JFA_F32)
        hipLaunchKernelGGL((convert_kernel<float, false>), g, dim3(kBlock), 0, st, (const float*)U, (float*)W, ncells, gas_params(gamma));
    else if (dtype == JFA_F64)
        hipLaunchKernelGGL((convert_kernel<double, false>), g, dim3(kBlock), 0, st, (const double*)U, (double*)W, ncells, gas_params(gamma));
    else
        return JFA_EUNSUPPORTED;
    return check_launch();
}

int jfa_signal_speeds(int32_t kind, const void* u_L, const void* u_R, const void* a_L, const void* a_R,
                      const void* rho_L, const void* rho_R, const void* p_L, const void* p_R, const void* H_L,
                      const void* H_R, double gamma, int64_t n, int32_t dtype, void* S_L, void* S_R,
                      void* stream) {
    if (kind < JFA_SS_ARITHMETIC || kind > JFA_SS_DAVIS2) return JFA_EUNSUPPORTED;
    if (n < 0 || !S_L || !S_R || S_L == S_R) return JFA_EINVAL;
    if (n == 0) return JFA_OK;
    const bool need_rho = kind == JFA_SS_EINFELDT || kind == JFA_SS_TORO || kind == JFA_SS_DAVIS2;
    if (!u_L || !u_R || (kind != JFA_SS_DAVIS2 && (!a_L || !a_R)) || (need_rho && (!rho_L || !rho_R)) ||
        (kind == JFA_SS_TORO && (!p_L || !p_R)) || (kind == JFA_SS_DAVIS2 && (!H_L || !H_R)))
        return JFA_EINVAL;
    const dim3 g = flat_grid(n);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == JFA_F64)
        hipLaunchKernelGGL(signal_speed_kernel<double>, g, dim3(kBlock), 0, st, kind, (const double*)u_L,
                           (const double*)u_R, (const double*)a_L, (const double*)a_R, (const double*)rho_L,
                           (const double*)rho_R, (const double*)p_L, (const double*)p_R, (const double*)H_L,
                           (const double*)H_R, gamma, n, (double*)S_L, (double*)S_R);
    else if (dtype == JFA_F32)
        hipLaunchKernelGGL(signal_speed_kernel<float>, g, dim3(kBlock), 0, st, kind, (const float*)u_L,
                           (const float*)u_R, (const float*)a_L, (const float*)a_R, (const float*)rho_L,
                           (const float*)rho_R, (const float*)p_L, (const float*)p_R, (const float*)H_L,
                           (const float*)H_R, gamma, n, (float*)S_L, (float*)S_R);
    else
        return JFA_EUNSUPPORTED;
    return check_launch();
}

size_t jfa_stage_workspace_bytes(const jfa_grid* grid, const jfa_scheme* scheme) {
    Geo G;
    if (!scheme || make_geo(grid, G)) return 0;
    // scratch RHS (FP64) + one CFL partial per inner-sweep block
    return 5 * (size_t)G.n[0] * G.n[1] * G.n[2] * sizeof(double) + (size_t)fused_partials(G) * sizeof(double) + 256;
}

}  // extern "C"

// Argument checks and launches of the fused stage with its scratch (FP64 RHS, then the CFL
// partials) at `workspace`.
static int fused_stage_entry(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out,
                             const jfa_grid* grid, const jfa_scheme* scheme, const int32_t bc[6], int32_t stage,
                             const double* dt_dev, double dt_host, double cfl, double eps, double* dt_next_dev,
                             void* workspace, size_t workspace_bytes, void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    rc = check_scheme(scheme);
    if (rc) return rc;
    rc = check_stencil_halos(scheme, G);
    if (rc) return rc;
    if (!W_in || !U_in || !W_out || !U_out || !bc || !workspace || stage < 0 || stage > 2 || (stage > 0 && !U_n))
        return JFA_EINVAL;
    if (W_in == W_out || U_in == U_out || W_out == U_in || U_out == W_in) return JFA_EINVAL;
    if (workspace_bytes < jfa_stage_workspace_bytes(grid, scheme)) return JFA_EINVAL;
    if (dt_next_dev && !(cfl > 0.0)) return JFA_EINVAL;
    BC6 b6;
    for (int i = 0; i < 6; ++i) {
        b6.t[i] = bc[i];
        if (bc[i] < 0 || bc[i] > 2) return JFA_EUNSUPPORTED;
        const int axis = i / 2;
        if (G.active[axis] && G.nh > G.n[axis] && bc[i] != JFA_BC_ZEROGRADIENT) return JFA_EINVAL;
    }
    rc = check_transport(scheme, G, nullptr);
    if (rc) return rc;
    const size_t ncell = (size_t)G.n[0] * G.n[1] * G.n[2];
    double* R = (double*)workspace;
    double* part = R + 5 * ncell;
    hipStream_t st = (hipStream_t)stream;
    if (scheme->dtype == JFA_F32)
        return fused_stage_f32(W_in, U_in, U_n, W_out, U_out, G, scheme, b6, stage, dt_dev, dt_host, cfl, eps,
                               dt_next_dev, R, part, st);
    return fused_stage_f64(W_in, U_in, U_n, W_out, U_out, G, scheme, b6, stage, dt_dev, dt_host, cfl, eps, dt_next_dev,
                           R, part, st);
}

extern "C" {
int jfa_rk3_stage_fused(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out,
                        const jfa_grid* grid, const jfa_scheme* scheme, const int32_t bc[6], int32_t stage,
                        const double* dt_dev, double dt_host, double cfl, double eps, double* dt_next_dev,
                        void* workspace, size_t workspace_bytes, void* stream) {
    return fused_stage_entry(W_in, U_in, U_n, W_out, U_out, grid, scheme, bc, stage, dt_dev, dt_host, cfl, eps,
                             dt_next_dev, workspace, workspace_bytes, stream);
}

}  // extern "C"

// ------------------------------------------------------------------ stage without allocation
// Scratch of jfa_rk3_stage_fused_noalloc (the FP64 RHS and the CFL partials of one stage): a
// file-scope device array, so the entry point needs no allocation.  Every GPU the library runs on
// holds its own copy, from the library's load until the process ends.  Host code never passes
// the array's name to a kernel (on the host it is a stand-in, not a device address): it uses the
// per-device address hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[JFA_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int noalloc_scratch(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        snprintf(g_last_error, sizeof(g_last_error), "%s", hipGetErrorString(e));
        return JFA_ELAUNCH;
    }
    if (dev < 0 || dev >= kMaxDevices) return JFA_EUNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            snprintf(g_last_error, sizeof(g_last_error), "hipGetSymbolAddress(scratch): %s",
                     e != hipSuccess ? hipGetErrorString(e) : "null address");
            return JFA_ELAUNCH;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return JFA_OK;
}
}  // namespace

extern "C" {
int jfa_rk3_stage_fused_noalloc(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out,
                                const jfa_grid* grid, const jfa_scheme* scheme, const int32_t bc[6], int32_t stage,
                                const double* dt_dev, double dt_host, double cfl, double eps, double* dt_next_dev,
                                void* stream) {
    const size_t need = jfa_stage_workspace_bytes(grid, scheme);
    if (need == 0) return JFA_EINVAL;
    if (need > (size_t)JFA_NOALLOC_SCRATCH_BYTES) {
        snprintf(g_last_error, sizeof(g_last_error),
                 "jfa_rk3_stage_fused_noalloc: the grid needs %zu bytes of scratch, the limit is %zu", need,
                 (size_t)JFA_NOALLOC_SCRATCH_BYTES);
        return JFA_EINVAL;
    }
    void* ws = nullptr;
    const int rc = noalloc_scratch(&ws);
    if (rc) return rc;
    // The stage writes the scratch RHS before reading it and one CFL partial per block, so the
    // previous call's contents need no clearing.
    return fused_stage_entry(W_in, U_in, U_n, W_out, U_out, grid, scheme, bc, stage, dt_dev, dt_host, cfl, eps,
                             dt_next_dev, ws, (size_t)JFA_NOALLOC_SCRATCH_BYTES, stream);
}

}  // extern "C"
