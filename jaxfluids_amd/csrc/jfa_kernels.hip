// jfa_kernels.hip -- HIP kernels and the extern "C" ABI of jaxfluids_amd (gfx950).
//
// Buffers are structure-of-arrays fields in the JAX-Fluids layout
// (nvar, Nx+2nh, Ny+2nh, Nz+2nh), z contiguous, so a wavefront of 64 lanes walking z
// issues 512-byte coalesced FP64 loads per field.  Storage type T is double or float;
// all arithmetic is FP64.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "jfa_device.h"

namespace jfa {

// ------------------------------------------------------------------ geometry
struct Geo {
    int n[3];        // interior cells (1 = inactive)
    int nh;
    int off[3];      // nh on active axes, 0 on inactive ones
    int B[3];        // buffer extents
    int64_t s[3];    // buffer strides (elements)
    int64_t fs;      // field stride
    double one_dx[3];
    double dx_min;
    int active[3];
};

static int make_geo(const jfa_grid* g, Geo& G) {
    if (!g) return JFA_EINVAL;
    if (g->nh < 3) return JFA_EINVAL;
    int nact = 0;
    G.dx_min = 1e300;
    for (int i = 0; i < 3; ++i) {
        if (g->n[i] < 1) return JFA_EINVAL;
        G.n[i] = g->n[i];
        G.active[i] = g->n[i] > 1;
        nact += G.active[i];
        G.off[i] = G.active[i] ? g->nh : 0;
        G.B[i] = G.active[i] ? g->n[i] + 2 * g->nh : 1;
        if (G.active[i]) {
            if (!(g->dx[i] > 0.0)) return JFA_EINVAL;
            G.one_dx[i] = 1.0 / g->dx[i];  // domain_information.py:290
            if (g->dx[i] < G.dx_min) G.dx_min = g->dx[i];
        } else {
            G.one_dx[i] = 0.0;
        }
    }
    if (nact == 0) return JFA_EINVAL;
    G.nh = g->nh;
    G.s[2] = 1;
    G.s[1] = G.B[2];
    G.s[0] = (int64_t)G.B[1] * G.B[2];
    G.fs = (int64_t)G.B[0] * G.B[1] * G.B[2];
    return JFA_OK;
}

static thread_local char g_last_error[256] = "";

static int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        snprintf(g_last_error, sizeof(g_last_error), "%s", hipGetErrorString(e));
        return JFA_ELAUNCH;
    }
    return JFA_OK;
}

template <class T>
__device__ __forceinline__ double ld(const T* __restrict__ p, int64_t i) {
    return static_cast<double>(p[i]);
}

// Face extents along axis a: N_a+1 on a, interior extents elsewhere.
struct FaceIdx {
    int fe[3];
    int64_t nface;
};
static FaceIdx face_extents(const Geo& G, int axis) {
    FaceIdx f;
    for (int i = 0; i < 3; ++i) f.fe[i] = G.n[i] + (i == axis ? 1 : 0);
    f.nface = (int64_t)f.fe[0] * f.fe[1] * f.fe[2];
    return f;
}

constexpr int kBlock = 256;

// Flat thread index -> (i0, i1, i2) over extents (e0, e1, e2), i2 fastest (32-bit math;
// every launch here has fewer than 2^31 items).
__device__ __forceinline__ bool flat_index(int64_t total, int e1, int e2, int& i0, int& i1, int& i2) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if ((int64_t)t >= total) return false;
    const uint32_t q = t / (uint32_t)e2;
    i2 = (int)(t - q * (uint32_t)e2);
    i0 = (int)(q / (uint32_t)e1);
    i1 = (int)(q - (uint32_t)i0 * (uint32_t)e1);
    return true;
}

static dim3 flat_grid(int64_t total) { return dim3((unsigned)((total + kBlock - 1) / kBlock)); }

// ------------------------------------------------------------------ WENO5 reconstruct_xi
template <class T, int STENCIL>
__global__ __launch_bounds__(kBlock) void weno5_reconstruct_kernel(const T* __restrict__ buf, int nvar, Geo G,
                                                                   int axis, int j, double eps, FaceIdx F,
                                                                   T* __restrict__ out) {
    int f0, f1, f2;
    if (!flat_index(F.nface, F.fe[1], F.fe[2], f0, f1, f2)) return;
    const int64_t base = (int64_t)(f0 + G.off[0]) * G.s[0] + (int64_t)(f1 + G.off[1]) * G.s[1] + (f2 + G.off[2]);
    const int64_t sa = G.s[axis];
    const int64_t o = ((int64_t)f0 * F.fe[1] + f1) * F.fe[2] + f2;
    for (int v = 0; v < nvar; ++v) {
        const T* b = buf + v * G.fs + base;
        double r;
        if (j == 0)  // cells -3..1 (weno5_base.py:30)
            r = weno5<STENCIL>(ld(b, -3 * sa), ld(b, -2 * sa), ld(b, -sa), ld(b, 0), ld(b, sa), eps);
        else  // cells 2..-2
            r = weno5<STENCIL>(ld(b, 2 * sa), ld(b, sa), ld(b, 0), ld(b, -sa), ld(b, -2 * sa), eps);
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

// ------------------------------------------------------------------ face flux (compute_flux_xi)
// One thread per face.  Loads the 6-cell window of the face along AXIS and evaluates
// reconstruction + Riemann in registers.
template <class T, int AXIS, int VAR, int STENCIL, int RIEMANN>
__device__ __forceinline__ void face_flux_at(const T* __restrict__ W, const T* __restrict__ U, int64_t base,
                                             int64_t sa, int64_t fs, const Params& prm, double Fo[5]) {
    double win[6][5];
    double wl[5], wr[5];
    constexpr bool cons_win = (VAR == JFA_CONSERVATIVE || VAR == JFA_CHAR_CONSERVATIVE);
    const T* src = cons_win ? U : W;
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int i = 0; i < 5; ++i) win[k][i] = ld(src, i * fs + base + (k - 3) * sa);
    if constexpr (VAR == JFA_CHAR_CONSERVATIVE) {
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            wl[i] = ld(W, i * fs + base - sa);
            wr[i] = ld(W, i * fs + base);
        }
    }
    double WL[5], WR[5], UL[5], UR[5];
    face_states<AXIS, VAR, STENCIL>(win, wl, wr, prm, WL, WR, UL, UR);
    riemann<AXIS, RIEMANN>(WL, WR, UL, UR, prm, Fo);
}

template <class T, int AXIS, int VAR, int STENCIL, int RIEMANN>
__global__ __launch_bounds__(kBlock) void flux_xi_kernel(const T* __restrict__ W, const T* __restrict__ U, Geo G,
                                                         FaceIdx F, Params prm, T* __restrict__ flux) {
    int f0, f1, f2;
    if (!flat_index(F.nface, F.fe[1], F.fe[2], f0, f1, f2)) return;
    const int64_t base = (int64_t)(f0 + G.off[0]) * G.s[0] + (int64_t)(f1 + G.off[1]) * G.s[1] + (f2 + G.off[2]);
    double Fo[5];
    face_flux_at<T, AXIS, VAR, STENCIL, RIEMANN>(W, U, base, G.s[AXIS], G.fs, prm, Fo);
    const int64_t o = ((int64_t)f0 * F.fe[1] + f1) * F.fe[2] + f2;
#pragma unroll
    for (int i = 0; i < 5; ++i) flux[(int64_t)i * F.nface + o] = static_cast<T>(Fo[i]);
}

// ------------------------------------------------------------------ flux divergence
// rhs (+)= one_dx * (F[f] - F[f+1]) (space_solver.py:597-599, 314).  rhs: (5, Nx, Ny, Nz).
template <class T>
__global__ __launch_bounds__(kBlock) void divergence_kernel(const T* __restrict__ flux, Geo G, int axis, FaceIdx F,
                                                            int accumulate, T* __restrict__ rhs) {
    const int64_t plane = (int64_t)G.n[1] * G.n[2];
    int c0, c1, c2;
    if (!flat_index((int64_t)G.n[0] * plane, G.n[1], G.n[2], c0, c1, c2)) return;
    const int64_t t = (int64_t)c1 * G.n[2] + c2;
    const int64_t fo = ((int64_t)c0 * F.fe[1] + c1) * F.fe[2] + c2;
    const int64_t fstep = axis == 0 ? (int64_t)F.fe[1] * F.fe[2] : (axis == 1 ? F.fe[2] : 1);
    const int64_t ncell = (int64_t)G.n[0] * plane;
    const int64_t ro = (int64_t)c0 * plane + t;
    const double odx = G.one_dx[axis];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const double d = odx * (ld(flux, i * F.nface + fo) - ld(flux, i * F.nface + fo + fstep));
        const double r = accumulate ? ld(rhs, i * ncell + ro) + d : 0.0 + d;
        rhs[i * ncell + ro] = static_cast<T>(r);
    }
}

// ------------------------------------------------------------------ RK3 stage
// RK3.py:27-29 tables
__device__ __forceinline__ void rk3_coeffs(int stage, double& c0, double& c1, double& m) {
    m = stage == 0 ? 1.0 : (stage == 1 ? 0.25 : 2.0 / 3.0);
    c0 = stage == 1 ? 0.25 : 2.0 / 3.0;
    c1 = stage == 1 ? 0.75 : 1.0 / 3.0;
}

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

// ------------------------------------------------------------------ CFL reduction
// max over interior cells of sum_active(|u_i| + c) (time_step_size.py:103-109).
// Non-negative doubles order like their bit patterns, so a 64-bit atomicMax on the bits
// is an exact, order-independent max (a NaN propagates as a huge value).
__device__ __forceinline__ double wave_sum(const double w[5], const int act[3], double gamma) {
    const double c = sound_speed(w[4], w[0], gamma);
    double s = 0.0;
    if (act[0]) s = s + (fabs(w[1]) + c);
    if (act[1]) s = s + (fabs(w[2]) + c);
    if (act[2]) s = s + (fabs(w[3]) + c);
    return s;
}

// Block-wide max (wavefront shuffles, then LDS across the block's waves) and one
// atomic per block.  Every thread of the block must call it.
__device__ __forceinline__ void block_max_to(unsigned long long* slot, double v) {
    __shared__ double part[kBlock / 64];
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) part[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = part[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, part[i]);
        atomicMax(slot, (unsigned long long)__double_as_longlong(m));
    }
}

constexpr int kCflBlocks = 1024;

// grid-stride over the interior cells: at most kCflBlocks atomics per reduction
template <class T>
__global__ __launch_bounds__(kBlock) void cfl_max_kernel(const T* __restrict__ W, Geo G, double gamma,
                                                         unsigned long long* slot) {
    const uint32_t total = (uint32_t)G.n[0] * G.n[1] * G.n[2];
    double v = 0.0;
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
        v = fmax(v, wave_sum(w, G.active, gamma));
    }
    block_max_to(slot, v);
}

__global__ void zero_slot_kernel(unsigned long long* slot) { slot[0] = 0ull; }

__global__ void dt_finalize_kernel(const unsigned long long* slot, double dx_min, double cfl, double eps,
                                   double* dt) {
    const double mx = __longlong_as_double((long long)slot[0]);
    dt[0] = dx_min / (mx + eps) * cfl;  // time_step_size.py:109,154-155
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
// Divergence of the three face fluxes + RK3 blend/update + prim + face-halo writes of the
// result (+ optional CFL max).  Interior cell c writes W_out/U_out at c and, if c lies in a
// boundary layer, at the halo positions that the face-halo fill copies it to.
struct BC6 {
    int t[6];
};

template <class T>
__device__ __forceinline__ void store_cell(T* __restrict__ W, T* __restrict__ U, int64_t fs, int64_t b,
                                           const double w[5], const double u[5]) {
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        W[i * fs + b] = static_cast<T>(w[i]);
        U[i * fs + b] = static_cast<T>(u[i]);
    }
}

template <class T>
__device__ __forceinline__ void halo_scatter(T* __restrict__ W, T* __restrict__ U, const Geo& G, const BC6& bc,
                                             const int c[3], int64_t b, const double w[5], const double u[5],
                                             const Params& prm) {
    const int nh = G.nh;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!G.active[a]) continue;
        const int N = G.n[a];
        const int ci = c[a];
        const int64_t sa = G.s[a];
        const int64_t bc0 = b - (int64_t)(ci + nh) * sa;  // same cell with axis-a coordinate 0
        // low side (west/south/bottom): bc index 2a+1
        const int tl = bc.t[2 * a + 1], th = bc.t[2 * a];
        // targets of this cell in the low halo
        if (tl == JFA_BC_PERIODIC) {
            if (ci >= N - nh) store_cell(W, U, G.fs, bc0 + (int64_t)(ci - N + nh) * sa, w, u);
        } else if (tl == JFA_BC_ZEROGRADIENT) {
            if (ci == 0)
                for (int m = 0; m < nh; ++m) store_cell(W, U, G.fs, bc0 + (int64_t)m * sa, w, u);
        } else {  // SYMMETRY
            if (ci < nh) {
                double ws[5], us[5];
#pragma unroll
                for (int i = 0; i < 5; ++i) ws[i] = w[i];
                ws[a + 1] = -ws[a + 1];
                cons_from_prim(ws, prm, us);
                store_cell(W, U, G.fs, bc0 + (int64_t)(nh - 1 - ci) * sa, ws, us);
            }
        }
        if (th == JFA_BC_PERIODIC) {
            if (ci < nh) store_cell(W, U, G.fs, bc0 + (int64_t)(N + nh + ci) * sa, w, u);
        } else if (th == JFA_BC_ZEROGRADIENT) {
            if (ci == N - 1)
                for (int m = 0; m < nh; ++m) store_cell(W, U, G.fs, bc0 + (int64_t)(N + nh + m) * sa, w, u);
        } else {
            if (ci >= N - nh) {
                double ws[5], us[5];
#pragma unroll
                for (int i = 0; i < 5; ++i) ws[i] = w[i];
                ws[a + 1] = -ws[a + 1];
                cons_from_prim(ws, prm, us);
                store_cell(W, U, G.fs, bc0 + (int64_t)(N + nh + (N - 1 - ci)) * sa, ws, us);
            }
        }
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void stage_epilogue_kernel(const T* __restrict__ Fx, const T* __restrict__ Fy,
                                                                const T* __restrict__ Fz, const T* __restrict__ Uin,
                                                                const T* __restrict__ Un, T* __restrict__ Wout,
                                                                T* __restrict__ Uout, Geo G, BC6 bc, int stage,
                                                                const double* __restrict__ dt_dev, double dt_host,
                                                                Params prm, unsigned long long* cfl_slot) {
    const int64_t plane = (int64_t)G.n[1] * G.n[2];
    double vmax = 0.0;
    int c[3];
    if (flat_index((int64_t)G.n[0] * plane, G.n[1], G.n[2], c[0], c[1], c[2])) {
        const int64_t b = (int64_t)(c[0] + G.off[0]) * G.s[0] + (int64_t)(c[1] + G.off[1]) * G.s[1] + (c[2] + G.off[2]);
        double rhs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const T* Fs[3] = {Fx, Fy, Fz};
        bool first = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (!G.active[a]) continue;
            int fe[3] = {G.n[0], G.n[1], G.n[2]};
            fe[a] += 1;
            const int64_t nf = (int64_t)fe[0] * fe[1] * fe[2];
            const int64_t fo = ((int64_t)c[0] * fe[1] + c[1]) * fe[2] + c[2];
            const int64_t st = a == 0 ? (int64_t)fe[1] * fe[2] : (a == 1 ? fe[2] : 1);
            const double odx = G.one_dx[a];
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                const double d = odx * (ld(Fs[a], i * nf + fo) - ld(Fs[a], i * nf + fo + st));
                rhs[i] = first ? 0.0 + d : rhs[i] + d;
            }
            first = false;
        }
        double c0, c1, m;
        rk3_coeffs(stage, c0, c1, m);
        const double ts = (dt_dev ? dt_dev[0] : dt_host) * m;
        double u[5], w[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            double x = ld(Uin, i * G.fs + b);
            if (stage > 0) x = c0 * x + c1 * ld(Un, i * G.fs + b);
            u[i] = x + ts * rhs[i];
        }
        prim_from_cons(u, prm, w);
        store_cell(Wout, Uout, G.fs, b, w, u);
        // halo cells: W copied, U recomputed as cons(W) (outer/material.py:248-250)
        double uh[5];
        cons_from_prim(w, prm, uh);
        halo_scatter(Wout, Uout, G, bc, c, b, w, uh, prm);
        if (cfl_slot) vmax = wave_sum(w, G.active, prm.gamma);
    }
    if (cfl_slot) block_max_to(cfl_slot, vmax);
}

// ------------------------------------------------------------------ dispatch tables
using FluxLaunch = void (*)(const void*, const void*, const Geo&, const FaceIdx&, const Params&, void*, hipStream_t);

template <class T, int AXIS, int VAR, int STENCIL, int RIEMANN>
static void launch_flux(const void* W, const void* U, const Geo& G, const FaceIdx& F, const Params& prm, void* out,
                        hipStream_t st) {
    dim3 grid = flat_grid(F.nface);
    hipLaunchKernelGGL((flux_xi_kernel<T, AXIS, VAR, STENCIL, RIEMANN>), grid, dim3(kBlock), 0, st,
                       static_cast<const T*>(W), static_cast<const T*>(U), G, F, prm, static_cast<T*>(out));
}

template <class T, int AXIS, int VAR, int STENCIL>
static FluxLaunch pick_r(int riemann) {
    return riemann == JFA_HLLC ? &launch_flux<T, AXIS, VAR, STENCIL, JFA_HLLC>
                               : &launch_flux<T, AXIS, VAR, STENCIL, JFA_RUSANOV>;
}
template <class T, int AXIS, int VAR>
static FluxLaunch pick_s(int stencil, int riemann) {
    return stencil == JFA_WENO5_JS ? pick_r<T, AXIS, VAR, JFA_WENO5_JS>(riemann)
                                   : pick_r<T, AXIS, VAR, JFA_WENO5_Z>(riemann);
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

static int check_scheme(const jfa_scheme* s) {
    if (!s) return JFA_EINVAL;
    if (s->recon_var < 0 || s->recon_var > 3) return JFA_EUNSUPPORTED;
    if (s->stencil != JFA_WENO5_JS && s->stencil != JFA_WENO5_Z) return JFA_EUNSUPPORTED;
    if (s->riemann != JFA_HLLC && s->riemann != JFA_RUSANOV) return JFA_EUNSUPPORTED;
    if (s->signal_speed < 0 || s->signal_speed > 4) return JFA_EUNSUPPORTED;
    if (s->frozen_state != JFA_FROZEN_ARITHMETIC && s->frozen_state != JFA_FROZEN_ROE) return JFA_EUNSUPPORTED;
    if (s->dtype != JFA_F64 && s->dtype != JFA_F32) return JFA_EUNSUPPORTED;
    if (!(s->gamma > 1.0)) return JFA_EINVAL;
    return JFA_OK;
}

static Params make_params(const jfa_scheme* s) {
    Params p;
    p.gamma = s->gamma;
    p.gm1 = s->gamma - 1.0;
    p.inv_gm1 = 1.0 / (s->gamma - 1.0);
    p.eps = s->eps_stencil;
    p.signal_speed = s->signal_speed;
    p.frozen = s->frozen_state;
    return p;
}

static Params gas_params(double gamma) {
    Params p;
    p.gamma = gamma;
    p.gm1 = gamma - 1.0;
    p.inv_gm1 = 1.0 / (gamma - 1.0);
    p.eps = 1e-30;
    p.signal_speed = JFA_SS_EINFELDT;
    p.frozen = JFA_FROZEN_ARITHMETIC;
    return p;
}

static size_t dsize(int dtype) { return dtype == JFA_F32 ? 4 : 8; }

static int run_flux(const void* W, const void* U, const Geo& G, const jfa_scheme* s, int axis, void* flux,
                    hipStream_t st) {
    const FaceIdx F = face_extents(G, axis);
    const Params prm = make_params(s);
    FluxLaunch fn = s->dtype == JFA_F32 ? pick_flux<float>(axis, s->recon_var, s->stencil, s->riemann)
                                        : pick_flux<double>(axis, s->recon_var, s->stencil, s->riemann);
    fn(W, U, G, F, prm, flux, st);
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
    if (stencil != JFA_WENO5_JS && stencil != JFA_WENO5_Z) return JFA_EUNSUPPORTED;
    const FaceIdx F = face_extents(G, axis);
    dim3 g = flat_grid(F.nface);
    hipStream_t st = (hipStream_t)stream;
#define JFA_W(T, S) \
    hipLaunchKernelGGL((weno5_reconstruct_kernel<T, S>), g, dim3(kBlock), 0, st, (const T*)buffer, nvar, G, axis, j, eps, F, (T*)out)
    if (dtype == JFA_F32) {
        if (stencil == JFA_WENO5_JS) JFA_W(float, JFA_WENO5_JS); else JFA_W(float, JFA_WENO5_Z);
    } else if (dtype == JFA_F64) {
        if (stencil == JFA_WENO5_JS) JFA_W(double, JFA_WENO5_JS); else JFA_W(double, JFA_WENO5_Z);
    } else {
        return JFA_EUNSUPPORTED;
    }
#undef JFA_W
    return check_launch();
}

int jfa_riemann_xi(const void* W_L, const void* W_R, const void* U_L, const void* U_R, int64_t nfaces, int32_t axis,
                   const jfa_scheme* scheme, void* flux, void* stream) {
    int rc = check_scheme(scheme);
    if (rc) return rc;
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
        if (scheme->riemann == JFA_HLLC) JFA_RA(float, JFA_HLLC); else JFA_RA(float, JFA_RUSANOV);
    } else {
        if (scheme->riemann == JFA_HLLC) JFA_RA(double, JFA_HLLC); else JFA_RA(double, JFA_RUSANOV);
    }
#undef JFA_RA
#undef JFA_R
    return check_launch();
}

int jfa_compute_flux_xi(const void* W, const void* U, const jfa_grid* grid, const jfa_scheme* scheme, int32_t axis,
                        void* flux, void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    rc = check_scheme(scheme);
    if (rc) return rc;
    if (!W || !U || !flux || axis < 0 || axis > 2 || !G.active[axis]) return JFA_EINVAL;
    return run_flux(W, U, G, scheme, axis, flux, (hipStream_t)stream);
}

size_t jfa_rhs_workspace_bytes(const jfa_grid* grid, int32_t dtype) {
    Geo G;
    if (make_geo(grid, G)) return 0;
    size_t mx = 0;
    for (int a = 0; a < 3; ++a) {
        if (!G.active[a]) continue;
        const FaceIdx F = face_extents(G, a);
        if ((size_t)F.nface > mx) mx = (size_t)F.nface;
    }
    return 5 * mx * dsize(dtype);
}

int jfa_compute_rhs(const void* W, const void* U, const jfa_grid* grid, const jfa_scheme* scheme, void* rhs,
                    void* workspace, size_t workspace_bytes, void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    rc = check_scheme(scheme);
    if (rc) return rc;
    if (!W || !U || !rhs || !workspace) return JFA_EINVAL;
    if (workspace_bytes < jfa_rhs_workspace_bytes(grid, scheme->dtype)) return JFA_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    bool first = true;
    for (int a = 0; a < 3; ++a) {
        if (!G.active[a]) continue;
        rc = run_flux(W, U, G, scheme, a, workspace, st);
        if (rc) return rc;
        const FaceIdx F = face_extents(G, a);
        const int64_t plane = (int64_t)G.n[1] * G.n[2];
        dim3 g = flat_grid((int64_t)G.n[0] * plane);
        if (scheme->dtype == JFA_F32)
            hipLaunchKernelGGL((divergence_kernel<float>), g, dim3(kBlock), 0, st, (const float*)workspace, G, a, F,
                               first ? 0 : 1, (float*)rhs);
        else
            hipLaunchKernelGGL((divergence_kernel<double>), g, dim3(kBlock), 0, st, (const double*)workspace, G, a,
                               F, first ? 0 : 1, (double*)rhs);
        rc = check_launch();
        if (rc) return rc;
        first = false;
    }
    return JFA_OK;
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

int jfa_compute_dt(const void* W, const jfa_grid* grid, double gamma, double cfl, double eps, int32_t dtype,
                   double* dt_dev, void* workspace, void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    if (!W || !dt_dev || !workspace) return JFA_EINVAL;
    if (dtype != JFA_F64 && dtype != JFA_F32) return JFA_EUNSUPPORTED;
    hipStream_t st = (hipStream_t)stream;
    unsigned long long* slot = (unsigned long long*)workspace;
    hipLaunchKernelGGL(zero_slot_kernel, dim3(1), dim3(1), 0, st, slot);
    const int64_t ncell = (int64_t)G.n[0] * G.n[1] * G.n[2];
    const int64_t nblk = (ncell + kBlock - 1) / kBlock;
    dim3 g((unsigned)(nblk < kCflBlocks ? nblk : kCflBlocks));
    if (dtype == JFA_F32)
        hipLaunchKernelGGL((cfl_max_kernel<float>), g, dim3(kBlock), 0, st, (const float*)W, G, gamma, slot);
    else
        hipLaunchKernelGGL((cfl_max_kernel<double>), g, dim3(kBlock), 0, st, (const double*)W, G, gamma, slot);
    hipLaunchKernelGGL(dt_finalize_kernel, dim3(1), dim3(1), 0, st, slot, G.dx_min, cfl, eps, dt_dev);
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

size_t jfa_stage_workspace_bytes(const jfa_grid* grid, const jfa_scheme* scheme) {
    Geo G;
    if (!scheme || make_geo(grid, G)) return 0;
    size_t tot = 0;
    for (int a = 0; a < 3; ++a) {
        if (!G.active[a]) continue;
        tot += 5 * (size_t)face_extents(G, a).nface * dsize(scheme->dtype);
    }
    return tot + 256;  // + CFL slot
}

int jfa_rk3_stage_fused(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out,
                        const jfa_grid* grid, const jfa_scheme* scheme, const int32_t bc[6], int32_t stage,
                        const double* dt_dev, double dt_host, void* workspace, size_t workspace_bytes,
                        void* stream) {
    Geo G;
    int rc = make_geo(grid, G);
    if (rc) return rc;
    rc = check_scheme(scheme);
    if (rc) return rc;
    if (!W_in || !U_in || !W_out || !U_out || !bc || !workspace || stage < 0 || stage > 2 || (stage > 0 && !U_n))
        return JFA_EINVAL;
    if (W_in == W_out || U_in == U_out) return JFA_EINVAL;
    if (workspace_bytes < jfa_stage_workspace_bytes(grid, scheme)) return JFA_EINVAL;
    BC6 b6;
    for (int i = 0; i < 6; ++i) {
        b6.t[i] = bc[i];
        if (bc[i] < 0 || bc[i] > 2) return JFA_EUNSUPPORTED;
        const int axis = i / 2;
        if (G.active[axis] && G.nh > G.n[axis] && bc[i] != JFA_BC_ZEROGRADIENT) return JFA_EINVAL;
    }
    hipStream_t st = (hipStream_t)stream;
    char* ws = (char*)workspace;
    void* Fp[3] = {nullptr, nullptr, nullptr};
    for (int a = 0; a < 3; ++a) {
        if (!G.active[a]) continue;
        Fp[a] = ws;
        rc = run_flux(W_in, U_in, G, scheme, a, Fp[a], st);
        if (rc) return rc;
        ws += 5 * (size_t)face_extents(G, a).nface * dsize(scheme->dtype);
    }
    const int64_t plane = (int64_t)G.n[1] * G.n[2];
    dim3 g = flat_grid((int64_t)G.n[0] * plane);
    if (scheme->dtype == JFA_F32)
        hipLaunchKernelGGL((stage_epilogue_kernel<float>), g, dim3(kBlock), 0, st, (const float*)Fp[0],
                           (const float*)Fp[1], (const float*)Fp[2], (const float*)U_in, (const float*)U_n,
                           (float*)W_out, (float*)U_out, G, b6, stage, dt_dev, dt_host, make_params(scheme),
                           (unsigned long long*)nullptr);
    else
        hipLaunchKernelGGL((stage_epilogue_kernel<double>), g, dim3(kBlock), 0, st, (const double*)Fp[0],
                           (const double*)Fp[1], (const double*)Fp[2], (const double*)U_in, (const double*)U_n,
                           (double*)W_out, (double*)U_out, G, b6, stage, dt_dev, dt_host, make_params(scheme),
                           (unsigned long long*)nullptr);
    return check_launch();
}

}  // extern "C"
