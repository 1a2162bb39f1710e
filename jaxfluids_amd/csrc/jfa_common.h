// jfa_common.h -- geometry, launch helpers and shared device epilogues of jaxfluids_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>

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
    double dx2[3];   // squared cell sizes (0 on inactive axes)
    double dx_min;
    int active[3];
};

static inline int make_geo(const jfa_grid* g, Geo& G) {
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
            G.dx2[i] = g->dx[i] * g->dx[i];
            if (g->dx[i] < G.dx_min) G.dx_min = g->dx[i];
        } else {
            G.one_dx[i] = 0.0;
            G.dx2[i] = 0.0;
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

extern thread_local char g_last_error[256];

static inline int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        snprintf(g_last_error, sizeof(g_last_error), "%s", hipGetErrorString(e));
        return JFA_ELAUNCH;
    }
    return JFA_OK;
}

// Hooks for a host build of these sources (the test suite's CPU emulation of the kernels):
// JFA_ACCESS(p, bytes) sees every state load and store, JFA_OPAQUE(x) hides a value from the
// optimiser, JFA_WAVES_PER_EU(n) asks the compiler to fit n waves per SIMD.  In the library
// JFA_ACCESS is empty, JFA_OPAQUE is an empty asm statement that keeps per-tap address
// arithmetic from being hoisted out of march loops, and JFA_WAVES_PER_EU the amdgpu attribute.
#ifndef JFA_ACCESS
#define JFA_ACCESS(p, bytes)
#endif
#ifndef JFA_OPAQUE
#define JFA_OPAQUE(x) asm volatile("" : "+v"(x))
#endif
#ifndef JFA_WAVES_PER_EU
#define JFA_WAVES_PER_EU(n) __attribute__((amdgpu_waves_per_eu(n)))
#endif

template <class T>
__device__ __forceinline__ double ld(const T* __restrict__ p, int64_t i) {
    JFA_ACCESS(p + i, sizeof(T));
    return static_cast<double>(p[i]);
}

// Face extents along axis a: N_a+1 on a, interior extents elsewhere.
struct FaceIdx {
    int fe[3];
    int64_t nface;
};
static inline FaceIdx face_extents(const Geo& G, int axis) {
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

// innermost active axis: the one with unit stride among the active axes
static inline int innermost_axis(const Geo& G) { return G.active[2] ? 2 : (G.active[1] ? 1 : 0); }

static inline dim3 flat_grid(int64_t total) { return dim3((unsigned)((total + kBlock - 1) / kBlock)); }

// RK3.py:27-29 tables
__device__ __forceinline__ void rk3_coeffs(int stage, double& c0, double& c1, double& m) {
    m = stage == 0 ? 1.0 : (stage == 1 ? 0.25 : 2.0 / 3.0);
    c0 = stage == 1 ? 0.25 : 2.0 / 3.0;
    c1 = stage == 1 ? 0.75 : 1.0 / 3.0;
}

// ------------------------------------------------------------------ CFL reduction
// max over interior cells of sum_active(|u_i| + c) (time_step_size.py:103-109).
// The reference's jnp.max returns NaN as soon as one cell is NaN, so its dt turns NaN and the
// time loop ends; every maximum of these reductions is therefore taken with nan_max, which
// returns NaN if either operand is NaN (fmax would drop it).  Non-negative doubles and NaN
// (positive or negative) order like their bit patterns read as unsigned 64-bit integers, so a
// 64-bit atomicMax on the bits is an exact, order-independent max that keeps a NaN.
__device__ __forceinline__ double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }

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
    for (int o = 32; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o, 64));
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) part[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = part[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = nan_max(m, part[i]);
        atomicMax(slot, (unsigned long long)__double_as_longlong(m));
    }
}

constexpr int kCflBlocks = 1024;

// lam = dt (1/dx_axis) sigma of the positivity flux limiter (limiter_flux.py:204).  WAVESPEED
// (:717-726): sigma = sum_active(m_a (1/dx_a)) / (m_axis (1/dx_axis)), m_a the maximum of
// |u_a| + c over the stage's input buffer -- the interior maxima in p.wave_max and, on 2-D and 3-D
// grids, the halo corners' constant p.corner_speed.
template <int AXIS>
__device__ __forceinline__ double flux_lambda(const Params& p, double dt) {
    if (p.flux_partition != JFA_PARTITION_WAVESPEED) return (dt * p.flim_odx[AXIS]) * p.flim_sigma[AXIS];
    double m[3], sum = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        m[a] = nan_max(__longlong_as_double((long long)p.wave_max[a]), p.corner_speed);
        if (p.flim_odx[a] != 0.0) sum = sum + m[a] * p.flim_odx[a];
    }
    return (dt * p.flim_odx[AXIS]) * (sum / (m[AXIS] * p.flim_odx[AXIS]));
}

// Zeroes n 64-bit slots (a fill kernel, so streams under graph capture get no memset node).
template <int UNUSED = 0>
__global__ void zero_slots_kernel(unsigned long long* slot, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) slot[i] = 0ull;
}

// Maxima over interior cells of |u_a| + c for a = 0, 1, 2 (0 on inactive axes) into slot[0..2]
// as bit patterns, for the WAVESPEED partition; the face halos hold copies or mirrors of interior
// cells, so they add nothing.  slot must be zeroed first.
template <class T>
__global__ __launch_bounds__(kBlock) void axis_wave_max_kernel(const T* __restrict__ W, Geo G, double gamma,
                                                               unsigned long long* slot) {
    const uint32_t total = (uint32_t)G.n[0] * G.n[1] * G.n[2];
    double v[3] = {0.0, 0.0, 0.0};
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const uint32_t q = t / (uint32_t)G.n[2];
        const int c2i = (int)(t - q * (uint32_t)G.n[2]);
        const int c0i = (int)(q / (uint32_t)G.n[1]);
        const int c1i = (int)(q - (uint32_t)c0i * (uint32_t)G.n[1]);
        const int64_t b =
            (int64_t)(c0i + G.off[0]) * G.s[0] + (int64_t)(c1i + G.off[1]) * G.s[1] + (c2i + G.off[2]);
        const double c = sound_speed(ld(W, 4 * G.fs + b), ld(W, b), gamma);
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (G.active[a]) v[a] = nan_max(v[a], fabs(ld(W, (1 + a) * G.fs + b)) + c);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a) __syncthreads();  // block_max_to's shared partials are reused
        block_max_to(slot + a, v[a]);
    }
}

// Launches the WAVESPEED maxima of W into slot[0..2] on stream st.
template <class T>
static inline void launch_axis_wave_max(const T* W, const Geo& G, double gamma, unsigned long long* slot,
                                        hipStream_t st) {
    hipLaunchKernelGGL(zero_slots_kernel<0>, dim3(1), dim3(64), 0, st, slot, 3);
    const int64_t ncell = (int64_t)G.n[0] * G.n[1] * G.n[2];
    const int64_t nblk = (ncell + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((axis_wave_max_kernel<T>), dim3((unsigned)(nblk < kCflBlocks ? nblk : kCflBlocks)),
                       dim3(kBlock), 0, st, W, G, gamma, slot);
}

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
        JFA_ACCESS(W + i * fs + b, sizeof(T));
        JFA_ACCESS(U + i * fs + b, sizeof(T));
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

// Edge halos of W (and U = cons(W) there) on the active edges, from the face halos already in the
// buffer (BoundaryConditionMaterial.edge_halo_update, halos/outer/material.py:289-380).  blockIdx.y
// is the edge: axis pair e / 4 ((0,1), (1,2), (0,2)), high face of the lower axis a when e & 1, of
// the upper axis b when e & 2; one thread per cell of the nh x nh x n_c block.  The first of the two
// faces in the reference's edge names (the one on a, boundary_condition.py:128-180) that is
// PERIODIC or SYMMETRY decides:
//   PERIODIC on axis x: the cell N_x further in along x (the opposite face's block,
//                       boundary_condition.py:607-654);
//   SYMMETRY on axis x: the mirror across the boundary along x, velocity x negated (:657-760);
//   neither:            0.5 (W[moved nh in along a] + W[moved nh in along b]), both face halos
//                       (halo_slices.py:77-150, the "_10" and "_01" blocks).
// Every source cell is interior along the third axis, so no edge reads another edge.
template <class T>
__global__ __launch_bounds__(kBlock) void edge_halo_kernel(T* __restrict__ W, T* __restrict__ U, Geo G, BC6 bc,
                                                           Params prm) {
    const int e = blockIdx.y;
    const int pair = e >> 2;
    const int a = pair == 1 ? 1 : 0, b = pair == 0 ? 1 : 2, c = 3 - a - b;
    if (!G.active[a] || !G.active[b]) return;
    const int ha = e & 1, hb = (e >> 1) & 1;
    const int nh = G.nh;
    const int64_t total = (int64_t)nh * nh * G.n[c];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int k = (int)(t % G.n[c]);
    const int ib = (int)((t / G.n[c]) % nh);
    const int ia = (int)(t / ((int64_t)G.n[c] * nh));
    const int pa = ha ? G.n[a] + nh + ia : ia, pb = hb ? G.n[b] + nh + ib : ib;  // buffer coordinates
    const int64_t dst = (int64_t)pa * G.s[a] + (int64_t)pb * G.s[b] + (int64_t)(k + G.off[c]) * G.s[c];
    const int ta = bc.t[2 * a + (ha ? 0 : 1)], tb = bc.t[2 * b + (hb ? 0 : 1)];
    const int64_t fs = G.fs;
    double w[5], u[5];
    int x = -1, kind = 0;
    if (ta == JFA_BC_PERIODIC || ta == JFA_BC_SYMMETRY) {
        x = a;
        kind = ta;
    } else if (tb == JFA_BC_PERIODIC || tb == JFA_BC_SYMMETRY) {
        x = b;
        kind = tb;
    }
    if (x < 0) {
        const int64_t s1 = dst + (ha ? -(int64_t)nh : (int64_t)nh) * G.s[a];
        const int64_t s2 = dst + (hb ? -(int64_t)nh : (int64_t)nh) * G.s[b];
#pragma unroll
        for (int i = 0; i < 5; ++i) w[i] = 0.5 * (ld(W, i * fs + s1) + ld(W, i * fs + s2));
    } else {
        const int hi = x == a ? ha : hb, pos = x == a ? pa : pb, N = G.n[x];
        int src;
        if (kind == JFA_BC_PERIODIC) src = hi ? pos - N : pos + N;
        else src = hi ? 2 * (N + nh) - 1 - pos : 2 * nh - 1 - pos;
        const int64_t s = dst + (int64_t)(src - pos) * G.s[x];
#pragma unroll
        for (int i = 0; i < 5; ++i) w[i] = ld(W, i * fs + s);
        if (kind == JFA_BC_SYMMETRY) w[x + 1] = -w[x + 1];
    }
    cons_from_prim(w, prm, u);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        W[i * fs + dst] = static_cast<T>(w[i]);
        if (U) U[i * fs + dst] = static_cast<T>(u[i]);
    }
}

// Whether the dissipative fluxes read halos (a transport with viscous or heat fluxes and a
// non-periodic active face); periodic domains wrap onto the interior instead.
static inline bool dissipative_halos(const Params& p, const Geo& G, const BC6& bc) {
    if (!p.dissip) return false;
    for (int f = 0; f < 6; ++f)
        if (G.active[f / 2] && bc.t[f] != JFA_BC_PERIODIC) return true;
    return false;
}

template <class T>
static void launch_edge_halos(T* W, T* U, const Geo& G, const BC6& bc, const Params& prm, hipStream_t st) {
    int nact = 0, nmax = 1;
    for (int i = 0; i < 3; ++i) {
        nact += G.active[i];
        if (G.n[i] > nmax) nmax = G.n[i];
    }
    if (nact < 2) return;
    const int64_t total = (int64_t)G.nh * G.nh * nmax;
    hipLaunchKernelGGL((edge_halo_kernel<T>), dim3((unsigned)((total + kBlock - 1) / kBlock), 12), dim3(kBlock), 0, st,
                       W, U, G, bc, prm);
}

// Halo cells a stencil's window needs (WENO7-JS: 4; every other stencil reads cells -3..2).
static inline int check_stencil_halos(const jfa_scheme* s, const Geo& G) {
    return (s->stencil == JFA_WENO7_JS && G.nh < 4) ? JFA_EINVAL : JFA_OK;
}

static inline int check_scheme(const jfa_scheme* s) {
    if (!s) return JFA_EINVAL;
    if (s->recon_var < 0 || s->recon_var > 3) return JFA_EUNSUPPORTED;
    if (s->stencil < JFA_WENO5_JS || s->stencil > JFA_WENO7_JS) return JFA_EUNSUPPORTED;
    if (s->stencil == JFA_WENO7_JS && s->riemann >= JFA_FS_ROE) return JFA_EUNSUPPORTED;
    if (s->riemann < JFA_HLLC || s->riemann > JFA_FS_LLF) return JFA_EUNSUPPORTED;
    if (s->signal_speed < 0 || s->signal_speed > 4) return JFA_EUNSUPPORTED;
    if (s->frozen_state != JFA_FROZEN_ARITHMETIC && s->frozen_state != JFA_FROZEN_ROE) return JFA_EUNSUPPORTED;
    if (s->dtype != JFA_F64 && s->dtype != JFA_F32) return JFA_EUNSUPPORTED;
    if (s->interpolation_limiter &&
        (!(s->limiter_eps_density > 0.0) || !(s->limiter_eps_density < INFINITY) ||
         !(s->limiter_eps_pressure > 0.0) || !(s->limiter_eps_pressure < INFINITY)))
        return JFA_EINVAL;
    if (s->flux_limiter < JFA_FLUX_LIMITER_NONE || s->flux_limiter > JFA_FLUX_LIMITER_HAS) return JFA_EUNSUPPORTED;
    if (s->flux_limiter &&
        ((s->flux_partition < JFA_PARTITION_UNIFORM || s->flux_partition > JFA_PARTITION_WAVESPEED) ||
         !(s->flux_limiter_eps_density > 0.0) || !(s->flux_limiter_eps_density < INFINITY) ||
         !(s->flux_limiter_eps_pressure > 0.0) || !(s->flux_limiter_eps_pressure < INFINITY)))
        return JFA_EINVAL;
    if (!(s->gamma > 1.0)) return JFA_EINVAL;
    return JFA_OK;
}

static inline Law make_law(int kind, double ref, double T_ref, double param) {
    Law l;
    l.kind = kind;
    l.ref = ref;
    l.T_ref = T_ref;
    l.inv_T_ref = T_ref > 0.0 ? 1.0 / T_ref : 0.0;
    l.param = param;
    return l;
}

// Dissipative-flux fields of p from a transport (NULL or neither flux: inviscid).
static inline void set_transport(Params& p, const jfa_transport* t, double gamma) {
    p.dissip = 0;
    p.diss_halos = 0;
    p.R = 1.0;
    p.mu = p.kappa = make_law(JFA_LAW_POWER, 0.0, 1.0, 0.0);
    p.bulk = p.kappa_fac = p.inv_cp = 0.0;
    if (!t || !(t->viscous || t->heat)) return;
    p.dissip = (t->viscous ? 1 : 0) | (t->heat ? 2 : 0);
    p.R = t->R;
    p.mu = make_law(t->mu_law, t->mu_ref, t->T_ref, t->mu_param);
    p.kappa = make_law(t->kappa_law, t->kappa_ref, t->kappa_T_ref, t->kappa_param);
    p.bulk = t->bulk_viscosity;
    const double cp = gamma / (gamma - 1.0) * t->R;  // ideal_gas.py:32
    p.kappa_fac = (t->heat && t->kappa_law == JFA_LAW_PRANDTL) ? cp / t->prandtl : 0.0;  // material.py:115-116
    p.inv_cp = 1.0 / cp;
}

static inline Params make_params(const jfa_scheme* s) {
    Params p;
    p.gamma = s->gamma;
    p.gm1 = s->gamma - 1.0;
    p.inv_gm1 = 1.0 / (s->gamma - 1.0);
    p.eps = s->eps_stencil;
    p.eps_z = 4.0 * s->eps_stencil;
    p.eps_js = 16.0 * s->eps_stencil;
    p.stencil = s->stencil;
    p.eps_precision = s->dtype == JFA_F32 ? 1.1920928955078125e-07 : 2.220446049250313e-16;
    p.riemann = s->riemann;
    p.signal_speed = s->signal_speed;
    p.frozen = s->frozen_state;
    set_transport(p, s->transport, s->gamma);
    p.limiter = s->interpolation_limiter ? 1 : 0;
    p.limit_velocity = s->limit_velocity ? 1 : 0;
    p.lim_eps_rho = s->limiter_eps_density;
    p.lim_eps_p = s->limiter_eps_pressure;
    p.flux_limiter = s->flux_limiter;
    p.flux_partition = s->flux_partition;
    p.flim_eps_rho = s->flux_limiter_eps_density;
    p.flim_eps_p = s->flux_limiter_eps_pressure;
    p.flim_sigma[0] = p.flim_sigma[1] = p.flim_sigma[2] = 0.0;
    p.flim_odx[0] = p.flim_odx[1] = p.flim_odx[2] = 0.0;
    p.wave_max = nullptr;
    p.corner_speed = 0.0;
    p.dx2[0] = p.dx2[1] = p.dx2[2] = 0.0;
    p.dt_ptr = nullptr;
    p.dt_val = 0.0;
    return p;
}

// The squared cell sizes that the WENO6-CUM stencils read (reconstruct_xi's dx = cell_sizes[axis]).
static inline void set_cell_sizes(Params& p, const Geo& G) {
    for (int a = 0; a < 3; ++a) p.dx2[a] = G.dx2[a];
}

// 1/sigma of the flux limiter's partition (limiter_flux.py:696-731) for a grid: the number of
// active axes (UNIFORM) or sum_active(1/dx) / (1/dx_axis) (CELLSIZE), summed from 0 in axis order.
// WAVESPEED's sigma depends on the state: flux_lambda evaluates it from p.wave_max, and this sets
// the cell sizes and the halo corners' wave speed it reads.
static inline void set_flux_partition(Params& p, const jfa_scheme* s, const double one_dx[3], const int active[3]) {
    int dim = 0;
    double sum = 0.0;
    for (int a = 0; a < 3; ++a)
        if (active[a]) {
            ++dim;
            sum += one_dx[a];
        }
    // halo corners (2-D, 3-D) keep the initialization's fill, rho = u = p = the precision's eps
    // (initialization/helper_functions.py:49-60): |u_a| + c = eps + sqrt(gamma eps / eps)
    p.corner_speed = dim >= 2 ? p.eps_precision + sqrt(s->gamma) : 0.0;
    for (int a = 0; a < 3; ++a) {
        p.flim_odx[a] = active[a] ? one_dx[a] : 0.0;
        if (!active[a]) {
            p.flim_sigma[a] = 0.0;
            continue;
        }
        p.flim_sigma[a] = s->flux_partition == JFA_PARTITION_CELLSIZE ? sum / one_dx[a] : (double)dim;
    }
}

// POWER or SUTHERLAND law with finite constants, ref >= 0, T_ref > 0 (and T_ref + C > 0).
static inline bool valid_law(int kind, double ref, double T_ref, double param) {
    if (kind != JFA_LAW_POWER && kind != JFA_LAW_SUTHERLAND) return false;
    if (!(ref >= 0.0) || !(ref < INFINITY) || !(T_ref > 0.0) || !(T_ref < INFINITY)) return false;
    if (!(param > -INFINITY) || !(param < INFINITY)) return false;
    return kind == JFA_LAW_POWER || T_ref + param > 0.0;
}

// A transport is usable when its constants are finite and positive where needed.  periodic_bc
// (the adjoint's boundaries, when non-null) must be PERIODIC on every active axis.
static inline int check_transport(const jfa_scheme* s, const Geo& G, const int32_t* periodic_bc) {
    const jfa_transport* t = s->transport;
    if (!t || !(t->viscous || t->heat)) return JFA_OK;
    if (!(t->R > 0.0) || !(t->bulk_viscosity >= 0.0)) return JFA_EINVAL;
    if (!valid_law(t->mu_law, t->mu_ref, t->T_ref, t->mu_param)) return JFA_EINVAL;
    if (t->heat) {
        if (t->kappa_law == JFA_LAW_PRANDTL) {
            if (!(t->prandtl > 0.0) || !(t->prandtl < INFINITY)) return JFA_EINVAL;
        } else if (!valid_law(t->kappa_law, t->kappa_ref, t->kappa_T_ref, t->kappa_param)) {
            return JFA_EINVAL;
        }
    }
    if (periodic_bc)
        for (int f = 0; f < 6; ++f)
            if (G.active[f / 2] && periodic_bc[f] != JFA_BC_PERIODIC) return JFA_EUNSUPPORTED;
    return JFA_OK;
}

static inline Params gas_params(double gamma) {
    Params p;
    p.gamma = gamma;
    p.gm1 = gamma - 1.0;
    p.inv_gm1 = 1.0 / (gamma - 1.0);
    p.eps = 1e-30;
    p.eps_z = 4.0 * 1e-30;
    p.eps_js = 16.0 * 1e-30;
    p.stencil = JFA_WENO5_JS;
    p.eps_precision = 2.220446049250313e-16;
    p.riemann = JFA_HLLC;
    p.signal_speed = JFA_SS_EINFELDT;
    p.frozen = JFA_FROZEN_ARITHMETIC;
    set_transport(p, nullptr, gamma);
    p.limiter = p.limit_velocity = 0;
    p.lim_eps_rho = p.lim_eps_p = 0.0;
    p.flux_limiter = p.flux_partition = 0;
    p.flim_eps_rho = p.flim_eps_p = 0.0;
    p.flim_sigma[0] = p.flim_sigma[1] = p.flim_sigma[2] = 0.0;
    p.dx2[0] = p.dx2[1] = p.dx2[2] = 0.0;
    p.dt_ptr = nullptr;
    p.dt_val = 0.0;
    return p;
}

static inline size_t dsize(int dtype) { return dtype == JFA_F32 ? 4 : 8; }

// Fused sweeps (jfa_fused.h), instantiated per storage type in jfa_fused_f64.hip / _f32.hip.
int fused_stage_f64(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out, const Geo& G,
                    const jfa_scheme* s, const BC6& bc, int stage, const double* dt_dev, double dt_host, double cfl,
                    double eps, double* dt_next, double* R, double* part, hipStream_t st);
int fused_stage_f32(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out, const Geo& G,
                    const jfa_scheme* s, const BC6& bc, int stage, const double* dt_dev, double dt_host, double cfl,
                    double eps, double* dt_next, double* R, double* part, hipStream_t st);
int fused_rhs_f64(const void* W, const void* U, const Geo& G, const jfa_scheme* s, double dt, double* R, void* rhs,
                  hipStream_t st);
int fused_rhs_f32(const void* W, const void* U, const Geo& G, const jfa_scheme* s, double dt, double* R, void* rhs,
                  hipStream_t st);
// number of per-block CFL partials of the inner sweep
int64_t fused_partials(const Geo& G);

}  // namespace jfa
