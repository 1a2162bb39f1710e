// jfa_fused.h -- the fused RK3 stage: face fluxes of every active axis, flux divergence,
// stage blend/update, prim(U), face-halo fill and the CFL partial max, in one pass over
// the state per axis without materialising face-flux arrays.
//
// Let IN be the innermost active axis (unit stride in the buffer) and A the other active
// axes.  For every A (ascending) a MARCH kernel walks its columns along A with a 6-cell
// register window (one new cell loaded per face, lanes along IN so every load is a
// coalesced row), carries the previous face flux in registers and accumulates
// (1/dx)(F_lo - F_hi) into a scratch RHS R.  The last pass is the INNER kernel: each
// wavefront evaluates 64 consecutive faces along IN (63 cells plus the shared upper
// face), takes the upper-face flux of its cell from the neighbouring lane by a wave
// shuffle, adds the IN divergence to R in the reference's axis order (x, y, z) and
// finishes the stage for its cell: U = c0 U + c1 U_n + dt m rhs (RK3.py:27-62),
// W = prim(U), stores both, scatters the cell into the halo positions the face-halo
// fill copies it to (U halo = cons(W halo), outer/material.py:248-250), and reduces the
// CFL wave speed of the new state into one partial per block.
//
// Included by the per-storage-type translation units jfa_fused_f64.hip / _f32.hip.
#pragma once

#include "jfa_common.h"
#include "jfa_dissipative.h"

namespace jfa {

constexpr int kInnerBlock = 256;
constexpr int kMarchBlockMax = 256;

// Face flux from a register window (cell k-h relative to the cell right of the face, k < n; 6 cells,
// 8 for WENO7-JS).
// CHAR-CONSERVATIVE also needs the primitives of the two cells adjacent to the face.
template <class T, int AXIS, int VAR, int STENCIL, int RIEMANN, bool FLIM>
__device__ __forceinline__ void window_flux(const double (&win)[Win<STENCIL>::n][5], const T* __restrict__ W,
                                            const T* __restrict__ U, int64_t b_right, int64_t sa, int64_t fs,
                                            const Params& prm, double F[5]) {
    if constexpr (RIEMANN == kFluxSplit) {
        flux_split_face<AXIS, STENCIL>(win, prm, F);
    } else {
        double wl[5], wr[5];
        if constexpr (VAR == JFA_CHAR_CONSERVATIVE) {
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                wl[i] = ld(W, i * fs + b_right - sa);
                wr[i] = ld(W, i * fs + b_right);
            }
        }
        double WL[5], WR[5], UL[5], UR[5];
        face_states<AXIS, VAR, STENCIL>(win, wl, wr, prm, WL, WR, UL, UR);
        if (prm.limiter) {
            double c[5];
#pragma unroll
            for (int i = 0; i < 5; ++i) c[i] = ld(W, i * fs + b_right - sa);
            limit_side<VAR>(c, prm, WL, UL);
#pragma unroll
            for (int i = 0; i < 5; ++i) c[i] = ld(W, i * fs + b_right);
            limit_side<VAR>(c, prm, WR, UR);
        }
        riemann<AXIS, RIEMANN>(WL, WR, UL, UR, prm, F);
    }
    if constexpr (FLIM) {  // positivity flux limiter, with the step's dt (space_solver.py:530-541)
        double wl_[5], wr_[5], ul_[5], ur_[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            wl_[i] = ld(W, i * fs + b_right - sa);
            wr_[i] = ld(W, i * fs + b_right);
            ul_[i] = ld(U, i * fs + b_right - sa);
            ur_[i] = ld(U, i * fs + b_right);
        }
        const double dt = prm.dt_ptr ? *prm.dt_ptr : prm.dt_val;
        limit_flux<AXIS>(F, wl_, wr_, ul_, ur_, flux_lambda<AXIS>(prm, dt), prm);
    }
}

// ------------------------------------------------------------------ MARCH kernel
// grid: x over IN (lanes), y over the third axis O, z over chunks of L cells along A.
// The plain convective variants are held to 3 waves per SIMD (168 VGPRs; they need 162): one
// register more costs a third of the resident waves of this VALU-bound loop.
template <class T, int A, int IN, int VAR, int STENCIL, int RIEMANN, bool DISS, bool FLIM>
__global__ __launch_bounds__(kMarchBlockMax)
JFA_WAVES_PER_EU((DISS || FLIM || VAR == JFA_CHAR_CONSERVATIVE || RIEMANN == kFluxSplit) ? 1 : 3) void march_kernel(
    const T* __restrict__ W, const T* __restrict__ U, Geo G, Params prm, int L, int accumulate,
    double* __restrict__ R) {
    constexpr int O = 3 - A - IN;
    const int iin = blockIdx.x * blockDim.x + threadIdx.x;
    if (iin >= G.n[IN]) return;  // no barriers in this kernel
    const int io = blockIdx.y;
    const int a0 = blockIdx.z * L;
    const int a1 = min(a0 + L, G.n[A]);
    const int64_t sa = G.s[A];
    const int64_t fs = G.fs;
    // buffer index of the column's cell with interior index 0 along A
    const int64_t bcol = (int64_t)(io + G.off[O]) * G.s[O] + (int64_t)(iin + G.off[IN]) * G.s[IN] +
                         (int64_t)G.off[A] * sa;
    constexpr bool cons_win = (VAR == JFA_CONSERVATIVE || VAR == JFA_CHAR_CONSERVATIVE);
    constexpr int NW = Win<STENCIL>::n, H = Win<STENCIL>::h;
    const T* __restrict__ src = cons_win ? U : W;
    double win[NW][5];
#pragma unroll
    for (int k = 0; k < NW; ++k)
#pragma unroll
        for (int i = 0; i < 5; ++i) win[k][i] = ld(src, i * fs + bcol + (int64_t)(a0 - H + k) * sa);
    double Fp[5];
    window_flux<T, A, VAR, STENCIL, RIEMANN, FLIM>(win, W, U, bcol + (int64_t)a0 * sa, sa, fs, prm, Fp);
    int cr[3];  // right cell of the current face (dissipative fluxes)
    cr[A] = a0;
    cr[O] = io;
    cr[IN] = iin;
    if constexpr (DISS) dissipative_face<T, A>(W, G, prm, cr, Fp);
    // R layout (5, N0, N1, N2)
    int c[3];
    c[A] = a0;
    c[O] = io;
    c[IN] = iin;
    const int64_t ncell = (int64_t)G.n[0] * G.n[1] * G.n[2];
    int64_t r = ((int64_t)c[0] * G.n[1] + c[1]) * G.n[2] + c[2];
    const int64_t rs = A == 0 ? (int64_t)G.n[1] * G.n[2] : (A == 1 ? (int64_t)G.n[2] : 1);
    const double odx = G.one_dx[A];
    // The new cell of the window is loaded at the top of its step, not a step ahead: the five
    // registers of a prefetch push the plain variants past 168 VGPRs into spills, which cost
    // more than the load latency that three resident waves already cover.
    for (int a = a0; a < a1; ++a, r += rs) {
#pragma unroll
        for (int k = 0; k < NW - 1; ++k)
#pragma unroll
            for (int i = 0; i < 5; ++i) win[k][i] = win[k + 1][i];
#pragma unroll
        for (int i = 0; i < 5; ++i) win[NW - 1][i] = ld(src, i * fs + bcol + (int64_t)(a + H) * sa);
        double Fn[5];
        window_flux<T, A, VAR, STENCIL, RIEMANN, FLIM>(win, W, U, bcol + (int64_t)(a + 1) * sa, sa, fs, prm, Fn);
        if constexpr (DISS) {
            cr[A] = a + 1;
            dissipative_face<T, A>(W, G, prm, cr, Fn);
        }
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const double d = odx * (Fp[i] - Fn[i]);  // space_solver.py:597-599
            double* p = R + i * ncell + r;
            *p = accumulate ? *p + d : 0.0 + d;
            Fp[i] = Fn[i];
        }
    }
}

// ------------------------------------------------------------------ INNER kernel
template <class T>
struct InnerArgs {
    const double* R;        // RHS of the previous sweeps (valid when accumulate)
    int accumulate;
    int mode;               // 0: write rhs to rhs_out; 1: finish the RK3 stage
    T* rhs_out;             // (5, N0, N1, N2) when mode 0
    const T* Uin;
    const T* Un;
    T* Wout;
    T* Uout;
    BC6 bc;
    int stage;
    const double* dt_dev;
    double dt_host;
    double* cfl_part;       // per-block partial max of the new state's wave speed, or null
};

// Block max written to part[b] (no atomics); every thread must call it.
__device__ __forceinline__ void block_max_part(double* part, double v, uint32_t b) {
    __shared__ double sm[kInnerBlock / 64];
    for (int o = 32; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sm[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = nan_max(m, sm[i]);
        part[b] = m;
    }
}

// Block maxima of three quantities written to part[b], part[nb + b], part[2 nb + b].
__device__ __forceinline__ void block_max_part3(double* part, double v0, double v1, double v2, uint32_t b,
                                                uint32_t nb) {
    __shared__ double sm[3][kInnerBlock / 64];
    for (int o = 32; o > 0; o >>= 1) {
        v0 = nan_max(v0, __shfl_xor(v0, o, 64));
        v1 = nan_max(v1, __shfl_xor(v1, o, 64));
        v2 = nan_max(v2, __shfl_xor(v2, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        sm[0][threadIdx.x >> 6] = v0;
        sm[1][threadIdx.x >> 6] = v1;
        sm[2][threadIdx.x >> 6] = v2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 3; ++q) {
            double m = sm[q][0];
            for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = nan_max(m, sm[q][i]);
            part[q * nb + b] = m;
        }
    }
}

// Stage finish of interior cell c (buffer index b, interior index rc) from its RHS.  Mode 0
// stores the RHS; mode 1 does U = c0 U + c1 U_n + dt m rhs (RK3.py:27-62), W = prim(U),
// stores both, scatters the cell into the halo positions the face-halo fill copies it to
// (U halo = cons(W halo), outer/material.py:248-250), and raises the running maxima of the
// CFL wave speed (and, with dissipative fluxes, of the diffusivities) of the new state.
template <class T, bool DISS>
__device__ __forceinline__ void finish_cell(const Geo& G, const Params& prm, const InnerArgs<T>& args, const int c[3],
                                            int64_t b, int64_t rc, int64_t ncell, const double rhs[5],
                                            double& vmax, double& nu_max, double& al_max) {
    const int64_t fs = G.fs;
    if (args.mode == 0) {
#pragma unroll
        for (int i = 0; i < 5; ++i) args.rhs_out[i * ncell + rc] = static_cast<T>(rhs[i]);
        return;
    }
    double c0, c1, m;
    rk3_coeffs(args.stage, c0, c1, m);
    const double ts = (args.dt_dev ? args.dt_dev[0] : args.dt_host) * m;  // RK3.py:60
    double u[5], w[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        double x = ld(args.Uin, i * fs + b);
        if (args.stage > 0) x = c0 * x + c1 * ld(args.Un, i * fs + b);  // RK3.py:49-50
        u[i] = x + ts * rhs[i];                                          // time_integrator.py:57
    }
    prim_from_cons(u, prm, w);
    store_cell(args.Wout, args.Uout, fs, b, w, u);
    double uh[5];
    cons_from_prim(w, prm, uh);
    halo_scatter(args.Wout, args.Uout, G, args.bc, c, b, w, uh, prm);
    if (args.cfl_part) {
        vmax = nan_max(vmax, wave_sum(w, G.active, prm.gamma));
        if constexpr (DISS) {
            double nu, al;
            diffusivities(w, prm, nu, al);
            nu_max = nan_max(nu_max, nu);
            al_max = nan_max(al_max, al);
        }
    }
}

template <class T, int IN, int VAR, int STENCIL, int RIEMANN, bool DISS, bool FLIM>
__global__ __launch_bounds__(kInnerBlock) void inner_kernel(const T* __restrict__ W, const T* __restrict__ U,
                                                            Geo G, Params prm, InnerArgs<T> args) {
    constexpr int P = IN == 0 ? 1 : 0;
    constexpr int Q = IN == 2 ? 1 : 2;
    const uint32_t NIN = (uint32_t)G.n[IN];
    const uint32_t NF = NIN + 1;
    const uint32_t nrows = (uint32_t)G.n[P] * (uint32_t)G.n[Q];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t g = (gtid >> 6) * 63u + lane;  // faces [63 w, 63 w + 63] per wavefront
    const uint32_t row = g / NF;
    const uint32_t f = g - row * NF;
    const bool vface = row < nrows;
    const int64_t fs = G.fs;
    double F[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    int cP = 0, cQ = 0;
    int64_t bf = 0;  // buffer index of the cell right of face f (s[IN] == 1)
    if (vface) {
        cP = (int)(row / (uint32_t)G.n[Q]);
        cQ = (int)(row - (uint32_t)cP * (uint32_t)G.n[Q]);
        bf = (int64_t)(cP + G.off[P]) * G.s[P] + (int64_t)(cQ + G.off[Q]) * G.s[Q] + (int64_t)(f + G.off[IN]);
        constexpr bool cons_win = (VAR == JFA_CONSERVATIVE || VAR == JFA_CHAR_CONSERVATIVE);
        const T* __restrict__ src = cons_win ? U : W;
        constexpr int NW = Win<STENCIL>::n, H = Win<STENCIL>::h;
        double win[NW][5];
#pragma unroll
        for (int k = 0; k < NW; ++k)
#pragma unroll
            for (int i = 0; i < 5; ++i) win[k][i] = ld(src, i * fs + bf + (k - H));
        window_flux<T, IN, VAR, STENCIL, RIEMANN, FLIM>(win, W, U, bf, 1, fs, prm, F);
        if constexpr (DISS) {
            int cr[3];
            cr[IN] = (int)f;
            cr[P] = cP;
            cr[Q] = cQ;
            dissipative_face<T, IN>(W, G, prm, cr, F);
        }
    }
    double Fu[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) Fu[i] = __shfl_down(F[i], 1, 64);
    const bool vcell = vface && lane < 63u && f < NIN;
    double vmax = 0.0, nu_max = 0.0, al_max = 0.0;
    if (vcell) {
        int c[3];
        c[IN] = (int)f;
        c[P] = cP;
        c[Q] = cQ;
        const int64_t ncell = (int64_t)G.n[0] * G.n[1] * G.n[2];
        const int64_t rc = ((int64_t)c[0] * G.n[1] + c[1]) * G.n[2] + c[2];
        const double odx = G.one_dx[IN];
        double rhs[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const double d = odx * (F[i] - Fu[i]);
            rhs[i] = args.accumulate ? args.R[i * ncell + rc] + d : 0.0 + d;  // space_solver.py:314
        }
        finish_cell<T, DISS>(G, prm, args, c, bf, rc, ncell, rhs, vmax, nu_max, al_max);
    }
    if constexpr (DISS) {
        if (args.cfl_part) block_max_part3(args.cfl_part, vmax, nu_max, al_max, blockIdx.x, gridDim.x);
    } else {
        if (args.cfl_part) block_max_part(args.cfl_part, vmax, blockIdx.x);
    }
}

// dt from the per-block partials (time_step_size.py:103-135,154-155): dx_min / (max sum(|u|+c) + eps),
// and with dissipative fluxes the minimum with 3/14 dx_min^2 / (max mu/rho + eps) (viscous) and
// 0.1 dx_min^2 / (max kappa/(rho cp) + eps) (heat), times CFL.  part holds 1 or 3 arrays of n.
template <int UNUSED = 0>  // template: one definition shared by both storage-type translation units
__global__ __launch_bounds__(1024) void dt_from_parts_kernel(const double* __restrict__ part, int n, double dx_min,
                                                             double cfl, double eps, int dissip,
                                                             double* __restrict__ dt) {
    __shared__ double sm[3][16];
    const int nq = dissip ? 3 : 1;
    for (int q = 0; q < nq; ++q) {
        double v = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) v = nan_max(v, part[q * n + i]);
        for (int o = 32; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0) sm[q][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < nq; ++q) {
            mx[q] = sm[q][0];
            for (int i = 1; i < (int)(blockDim.x >> 6); ++i) mx[q] = nan_max(mx[q], sm[q][i]);
        }
        double d = dx_min / (mx[0] + eps);
        if (dissip & 1) d = fmin(d, 3.0 / 14.0 * dx_min * dx_min / (mx[1] + eps));
        if (dissip & 2) d = fmin(d, 0.1 * dx_min * dx_min / (mx[2] + eps));
        dt[0] = d * cfl;
    }
}

// ------------------------------------------------------------------ host side
static inline int64_t inner_blocks(const Geo& G) {
    const int IN = innermost_axis(G);
    const int P = IN == 0 ? 1 : 0, Q = IN == 2 ? 1 : 2;
    const int64_t nfaces = (int64_t)G.n[P] * G.n[Q] * (G.n[IN] + 1);
    const int64_t nwaves = (nfaces - 2) / 63 + 1;
    return (nwaves + (kInnerBlock / 64) - 1) / (kInnerBlock / 64);
}

// Chunk length L along A: every wave marches L cells (L+1 faces).  The waves of one launch
// run in ceil(waves / slots) rounds, slots = CUs x resident waves per CU of this kernel, so
// L is chosen to minimise rounds x (L + 1) -- short chunks cost an extra face each, a
// part-filled last round idles the chip.  waves_per_layer: wavefronts of one chunk layer.
static inline int march_chunk(int nA, int64_t waves_per_layer, int64_t slots) {
    int best_L = 4;
    double best = 1e300;
    for (int L = 4; L <= 64; ++L) {
        const int64_t chunks = ((int64_t)nA + L - 1) / L;
        const int64_t waves = chunks * waves_per_layer;
        const int64_t rounds = (waves + slots - 1) / slots;
        const double cost = (double)rounds * (L + 1);
        if (cost < best * 0.999) {
            best = cost;
            best_L = L;
        }
        if (L >= nA) break;
    }
    return best_L;
}

// resident waves of a kernel on the whole device (cached per kernel instance)
static inline int64_t device_wave_slots(const void* kernel, int block) {
    static int num_cu = 0;
    if (num_cu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || num_cu <= 0)
            num_cu = 256;
    }
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, block, 0) != hipSuccess || blocks <= 0)
        blocks = 1;
    return (int64_t)num_cu * blocks * ((block + 63) / 64);
}

template <class T, int A, int IN, int VAR, int STENCIL, int RIEMANN, bool DISS, bool FLIM>
static void launch_march(const void* W, const void* U, const Geo& G, const Params& prm, int accumulate, double* R,
                         hipStream_t st) {
    constexpr int O = 3 - A - IN;
    int bx = ((G.n[IN] + 63) / 64) * 64;
    if (bx > kMarchBlockMax) bx = kMarchBlockMax;
    static int64_t slots[kMarchBlockMax / 64 + 1] = {0};
    int64_t& sl = slots[bx / 64];
    if (sl == 0) sl = device_wave_slots((const void*)march_kernel<T, A, IN, VAR, STENCIL, RIEMANN, DISS, FLIM>, bx);
    const int L = march_chunk(G.n[A], ((int64_t)G.n[IN] + 63) / 64 * G.n[O], sl);
    dim3 grid((unsigned)((G.n[IN] + bx - 1) / bx), (unsigned)G.n[O], (unsigned)((G.n[A] + L - 1) / L));
    hipLaunchKernelGGL((march_kernel<T, A, IN, VAR, STENCIL, RIEMANN, DISS, FLIM>), grid, dim3(bx), 0, st,
                       static_cast<const T*>(W), static_cast<const T*>(U), G, prm, L, accumulate, R);
}

template <class T, int IN, int VAR, int STENCIL, int RIEMANN, bool DISS, bool FLIM>
static int64_t launch_inner(const void* W, const void* U, const Geo& G, const Params& prm, const InnerArgs<T>& a,
                            hipStream_t st) {
    dim3 grid((unsigned)inner_blocks(G));
    hipLaunchKernelGGL((inner_kernel<T, IN, VAR, STENCIL, RIEMANN, DISS, FLIM>), grid, dim3(kInnerBlock), 0, st,
                       static_cast<const T*>(W), static_cast<const T*>(U), G, prm, a);
    return grid.x;
}

// All sweeps of one stage (or of one RHS evaluation) for fixed scheme templates: a MARCH
// kernel per active axis below IN, then the INNER kernel.  Returns the number of blocks of the
// INNER launch (its CFL partials per quantity).
template <class T, int VAR, int STENCIL, int RIEMANN, bool DISS, bool FLIM>
static int64_t sweeps(const void* W, const void* U, const Geo& G, const Params& prm, double* R, InnerArgs<T> ia,
                      hipStream_t st) {
    const int IN = innermost_axis(G);
    int acc = 0;
    for (int A = 0; A < IN; ++A) {
        if (!G.active[A]) continue;
        if (A == 0 && IN == 1) launch_march<T, 0, 1, VAR, STENCIL, RIEMANN, DISS, FLIM>(W, U, G, prm, acc, R, st);
        else if (A == 0) launch_march<T, 0, 2, VAR, STENCIL, RIEMANN, DISS, FLIM>(W, U, G, prm, acc, R, st);
        else launch_march<T, 1, 2, VAR, STENCIL, RIEMANN, DISS, FLIM>(W, U, G, prm, acc, R, st);
        acc = 1;
    }
    ia.R = R;
    ia.accumulate = acc;
    if (IN == 0) return launch_inner<T, 0, VAR, STENCIL, RIEMANN, DISS, FLIM>(W, U, G, prm, ia, st);
    if (IN == 1) return launch_inner<T, 1, VAR, STENCIL, RIEMANN, DISS, FLIM>(W, U, G, prm, ia, st);
    return launch_inner<T, 2, VAR, STENCIL, RIEMANN, DISS, FLIM>(W, U, G, prm, ia, st);
}

template <class T, int VAR, int STENCIL, int RIEMANN>
static int64_t sweeps_f(const void* W, const void* U, const Geo& G, const Params& prm, double* R,
                        const InnerArgs<T>& ia, hipStream_t st) {
    // dissipative fluxes and the positivity flux limiter are compile-time variants, so the plain
    // convective sweeps keep their register budget
    if (prm.dissip) {
        if (prm.flux_limiter) return sweeps<T, VAR, STENCIL, RIEMANN, true, true>(W, U, G, prm, R, ia, st);
        return sweeps<T, VAR, STENCIL, RIEMANN, true, false>(W, U, G, prm, R, ia, st);
    }
    if (prm.flux_limiter) return sweeps<T, VAR, STENCIL, RIEMANN, false, true>(W, U, G, prm, R, ia, st);
    return sweeps<T, VAR, STENCIL, RIEMANN, false, false>(W, U, G, prm, R, ia, st);
}
// The flux-splitting sweeps (the kFluxSplit face flux on the primitive window), instantiated in
// jfa_fused_fsplit.hip.
template <class T, int STENCIL>
int64_t sweeps_fs(const void* W, const void* U, const Geo& G, const Params& prm, double* R, const InnerArgs<T>& ia,
                  hipStream_t st) {
    return sweeps_f<T, JFA_PRIMITIVE, STENCIL, kFluxSplit>(W, U, G, prm, R, ia, st);
}
#define JFA_FUSED_SWEEPS_FS(PREFIX, T, STENCIL)                                                                \
    PREFIX template int64_t sweeps_fs<T, STENCIL>(const void*, const void*, const Geo&, const Params&, double*, \
                                                  const InnerArgs<T>&, hipStream_t);
#ifndef JFA_FUSED_INSTANTIATE_FS
JFA_FUSED_SWEEPS_FS(extern, double, JFA_WENO5_JS)
JFA_FUSED_SWEEPS_FS(extern, double, JFA_WENO5_Z)
JFA_FUSED_SWEEPS_FS(extern, double, kStencilExt)
JFA_FUSED_SWEEPS_FS(extern, float, JFA_WENO5_JS)
JFA_FUSED_SWEEPS_FS(extern, float, JFA_WENO5_Z)
JFA_FUSED_SWEEPS_FS(extern, float, kStencilExt)
#endif

template <class T, int VAR, int STENCIL>
static int64_t sweeps_r(int riemann, const void* W, const void* U, const Geo& G, const Params& prm, double* R,
                        const InnerArgs<T>& ia, hipStream_t st) {
    if (riemann == JFA_HLLC) return sweeps_f<T, VAR, STENCIL, JFA_HLLC>(W, U, G, prm, R, ia, st);
    if (riemann >= JFA_FS_ROE) {  // flux splitting reads the primitive window: run_sweeps sends it there
        if constexpr (VAR == JFA_PRIMITIVE && STENCIL != kStencilW7) return sweeps_fs<T, STENCIL>(W, U, G, prm, R, ia, st);
        return -1;  // not built (check_stencil refuses it)
    }
    if (riemann != JFA_RUSANOV) return sweeps_f<T, VAR, STENCIL, kRiemannExt>(W, U, G, prm, R, ia, st);
    return sweeps_f<T, VAR, STENCIL, JFA_RUSANOV>(W, U, G, prm, R, ia, st);
}
template <class T, int VAR>
int64_t sweeps_s(int stencil, int riemann, const void* W, const void* U, const Geo& G, const Params& prm, double* R,
                 const InnerArgs<T>& ia, hipStream_t st) {
    if (stencil == JFA_WENO5_JS) return sweeps_r<T, VAR, JFA_WENO5_JS>(riemann, W, U, G, prm, R, ia, st);
    if (stencil == JFA_WENO5_Z) return sweeps_r<T, VAR, JFA_WENO5_Z>(riemann, W, U, G, prm, R, ia, st);
    if (stencil == JFA_WENO7_JS) return sweeps_r<T, VAR, kStencilW7>(riemann, W, U, G, prm, R, ia, st);
    return sweeps_r<T, VAR, kStencilExt>(riemann, W, U, G, prm, R, ia, st);
}

// The sweeps of each reconstruction variable (both storage types) are instantiated in a
// translation unit of their own, jfa_fused_<variable>.hip, so the kernels compile in parallel.
#define JFA_FUSED_SWEEPS(PREFIX, T, VAR)                                                                    \
    PREFIX template int64_t sweeps_s<T, VAR>(int, int, const void*, const void*, const Geo&, const Params&, \
                                             double*, const InnerArgs<T>&, hipStream_t);
#ifndef JFA_FUSED_INSTANTIATE
JFA_FUSED_SWEEPS(extern, double, JFA_PRIMITIVE)
JFA_FUSED_SWEEPS(extern, float, JFA_PRIMITIVE)
JFA_FUSED_SWEEPS(extern, double, JFA_CONSERVATIVE)
JFA_FUSED_SWEEPS(extern, float, JFA_CONSERVATIVE)
JFA_FUSED_SWEEPS(extern, double, JFA_CHAR_PRIMITIVE)
JFA_FUSED_SWEEPS(extern, float, JFA_CHAR_PRIMITIVE)
JFA_FUSED_SWEEPS(extern, double, JFA_CHAR_CONSERVATIVE)
JFA_FUSED_SWEEPS(extern, float, JFA_CHAR_CONSERVATIVE)
#endif

// Launches the sweeps; *nparts receives the CFL partials per quantity of the finishing launch.
// wave_slots: 3 device slots for the WAVESPEED partition's maxima of W (used when the scheme's flux
// limiter has that partition).
template <class T>
static int run_sweeps(const jfa_scheme* s, const void* W, const void* U, const Geo& G, double* R,
                      const InnerArgs<T>& ia, const double* dt_dev, double dt_host, hipStream_t st,
                      int64_t* nparts, unsigned long long* wave_slots, int diss_halos) {
    Params prm = make_params(s);
    prm.diss_halos = diss_halos;
    set_flux_partition(prm, s, G.one_dx, G.active);
    set_cell_sizes(prm, G);
    prm.dt_ptr = dt_dev;  // the flux limiter's step dt
    prm.dt_val = dt_host;
    if (s->flux_limiter && s->flux_partition == JFA_PARTITION_WAVESPEED) {
        launch_axis_wave_max<T>(static_cast<const T*>(W), G, s->gamma, wave_slots, st);
        prm.wave_max = wave_slots;
    }
    int64_t nb = 0;
    switch (s->riemann >= JFA_FS_ROE ? JFA_PRIMITIVE : s->recon_var) {  // flux splitting: primitive window
        case JFA_PRIMITIVE: nb = sweeps_s<T, JFA_PRIMITIVE>(s->stencil, s->riemann, W, U, G, prm, R, ia, st); break;
        case JFA_CONSERVATIVE: nb = sweeps_s<T, JFA_CONSERVATIVE>(s->stencil, s->riemann, W, U, G, prm, R, ia, st); break;
        case JFA_CHAR_PRIMITIVE: nb = sweeps_s<T, JFA_CHAR_PRIMITIVE>(s->stencil, s->riemann, W, U, G, prm, R, ia, st); break;
        default: nb = sweeps_s<T, JFA_CHAR_CONSERVATIVE>(s->stencil, s->riemann, W, U, G, prm, R, ia, st); break;
    }
    if (nparts) *nparts = nb;
    return check_launch();
}

template <class T>
int fused_stage_impl(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out, const Geo& G,
                     const jfa_scheme* s, const BC6& bc, int stage, const double* dt_dev, double dt_host, double cfl,
                     double eps, double* dt_next, double* R, double* part, hipStream_t st) {
    InnerArgs<T> ia;
    ia.R = nullptr;
    ia.accumulate = 0;
    ia.mode = 1;
    ia.rhs_out = nullptr;
    ia.Uin = static_cast<const T*>(U_in);
    ia.Un = static_cast<const T*>(U_n);
    ia.Wout = static_cast<T*>(W_out);
    ia.Uout = static_cast<T*>(U_out);
    ia.bc = bc;
    ia.stage = stage;
    ia.dt_dev = dt_dev;
    ia.dt_host = dt_host;
    ia.cfl_part = dt_next ? part : nullptr;
    int64_t nparts = 0;
    // the WAVESPEED slots follow the CFL partials (inside the workspace's 256-byte tail)
    unsigned long long* slots = reinterpret_cast<unsigned long long*>(part + fused_partials(G));
    // Non-periodic dissipative fluxes read the halos (periodic ones wrap onto the interior, so no
    // stencil reads an edge halo there).  The inner kernel's scatter fills the face halos of the
    // new state; on non-periodic domains the edge halos follow (fill_edge_halos_material,
    // halo_manager.py:125-129,183-186).
    const Params base = make_params(s);
    const int diss_halos = dissipative_halos(base, G, bc) ? 1 : 0;
    int rc = run_sweeps<T>(s, W_in, U_in, G, R, ia, dt_dev, dt_host, st, &nparts, slots, diss_halos);
    if (rc) return rc;
    if (diss_halos) {
        launch_edge_halos<T>(static_cast<T*>(W_out), static_cast<T*>(U_out), G, bc, base, st);
        rc = check_launch();
        if (rc) return rc;
    }
    if (dt_next) {
        hipLaunchKernelGGL(dt_from_parts_kernel<0>, dim3(1), dim3(1024), 0, st, part, (int)nparts, G.dx_min,
                           cfl, eps, base.dissip, dt_next);
        rc = check_launch();
    }
    return rc;
}

template <class T>
int fused_rhs_impl(const void* W, const void* U, const Geo& G, const jfa_scheme* s, double dt, double* R,
                   void* rhs, hipStream_t st) {
    InnerArgs<T> ia = {};
    ia.mode = 0;
    ia.rhs_out = static_cast<T*>(rhs);
    // the WAVESPEED slots follow the partial RHS (inside the workspace's 256-byte tail)
    unsigned long long* slots = reinterpret_cast<unsigned long long*>(R + 5 * (int64_t)G.n[0] * G.n[1] * G.n[2]);
    const int halos = s->transport && s->transport->halos ? 1 : 0;
    return run_sweeps<T>(s, W, U, G, R, ia, nullptr, dt, st, nullptr, slots, halos);
}

}  // namespace jfa
