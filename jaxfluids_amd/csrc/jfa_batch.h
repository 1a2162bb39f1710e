// jfa_batch.h -- one fused launch per RK3 stage for 2-D domains (z inactive), for a single
// domain or an ensemble of independent domains of uneven size.
//
// Work item = (case, x-chunk [x0, x1) of <= 64 planes, y-tile [y0, y0+64)).  One wavefront
// owns one item: lane l holds the column y = y0 + l and marches x through the chunk,
// carrying the flux of the x-face below the current plane.  At each plane it evaluates the y-face
// below its cell from a y-window, takes the face above from lane l+1 (wave shuffle) and,
// for lane 63, from the extra upper face precomputed for the whole chunk (lane k holds the
// extra face of plane x0+k).  The stage is then finished in registers: rhs = (0 + rhs_x) +
// rhs_y (space_solver.py:266,314), RK3 blend/update, W = prim(U), stores, halo scatter,
// CFL partial.  No scratch RHS, no LDS, no cross-wave synchronisation.
//
// The grid holds one wavefront per item of every case; the hardware hands blocks to compute
// units as they free up, which balances small and large cases without a software work queue
// (and keeps every wave-level operation out of loops with lane-divergent exits).  Per-case
// time steps live in device memory.
#pragma once

#include "jfa_common.h"

namespace jfa {

struct Case2D {
    int n0, n1;          // interior cells along x, y
    int nh;
    int bc[4];           // east, west, north, south (JFA_BC_*)
    double one_dx0, one_dx1, dx_min;
    int64_t off;         // element offset of field 0 of this case in the pooled buffers
    int64_t fs;          // field stride (B0 * B1)
    int64_t s0;          // x stride (= B1); y has unit stride
    int item_begin, item_end;  // this case's items in the work table
};

struct Item2D {
    int c;       // case index
    int x0, x1;  // planes [x0, x1), x1 - x0 <= 64
    int y0;      // first cell of the 64-cell y tile
};

constexpr int kBatchBlock = 256;

template <class T>
struct BatchArgs {
    const T* Win;
    const T* Uin;
    const T* Un;
    T* Wout;
    T* Uout;
    const Case2D* cases;
    const Item2D* items;
    int nitems;
    int stage;
    const double* dt;        // per case
    double* part;            // per item CFL partial (nullable)
    const unsigned long long* wave_max;  // per case, 3 slots: WAVESPEED maxima of W_in (nullable)
};

// Wave-uniform copies in scalar registers.  The records are read with vector loads (the
// kernel's stores could alias them as far as the compiler knows); readfirstlane moves each
// field into an SGPR, so the per-case constants do not occupy vector registers.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni(int64_t v) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double uni(double v) { return __longlong_as_double(uni((int64_t)__double_as_longlong(v))); }

__device__ __forceinline__ Case2D load_case(const Case2D* __restrict__ p) {
    const Case2D r = *p;
    Case2D c;
    c.n0 = uni(r.n0);
    c.n1 = uni(r.n1);
    c.nh = uni(r.nh);
    for (int f = 0; f < 4; ++f) c.bc[f] = uni(r.bc[f]);
    c.one_dx0 = uni(r.one_dx0);
    c.one_dx1 = uni(r.one_dx1);
    c.dx_min = uni(r.dx_min);
    c.off = uni(r.off);
    c.fs = uni(r.fs);
    c.s0 = uni(r.s0);
    c.item_begin = uni(r.item_begin);
    c.item_end = uni(r.item_end);
    return c;
}

__device__ __forceinline__ Item2D load_item(const Item2D* __restrict__ p) {
    const Item2D r = *p;
    Item2D it;
    it.c = uni(r.c);
    it.x0 = uni(r.x0);
    it.x1 = uni(r.x1);
    it.y0 = uni(r.y0);
    return it;
}

// Geo of one 2-D case, for the shared epilogue helpers.
__device__ __forceinline__ Geo geo_of(const Case2D& cs) {
    Geo G;
    G.n[0] = cs.n0;
    G.n[1] = cs.n1;
    G.n[2] = 1;
    G.nh = cs.nh;
    G.off[0] = cs.nh;
    G.off[1] = cs.nh;
    G.off[2] = 0;
    G.B[0] = cs.n0 + 2 * cs.nh;
    G.B[1] = cs.n1 + 2 * cs.nh;
    G.B[2] = 1;
    G.s[0] = cs.s0;
    G.s[1] = 1;
    G.s[2] = 1;
    G.fs = cs.fs;
    G.one_dx[0] = cs.one_dx0;
    G.one_dx[1] = cs.one_dx1;
    G.one_dx[2] = 0.0;
    G.dx_min = cs.dx_min;
    G.active[0] = 1;
    G.active[1] = 1;
    G.active[2] = 0;
    return G;
}

// Loads and stores at a wave-uniform base pointer plus a per-lane byte offset.  The base
// lives in scalar registers (it carries the plane, field and stencil-tap offsets), the
// lane part is one 32-bit vector register (the column), so every access of a face window
// compiles to the scalar-base form and the march keeps no per-tap vector addresses.
template <class T>
__device__ __forceinline__ double ldo(const T* __restrict__ base, uint32_t voff) {
    const T* p = reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + voff);
    JFA_ACCESS(p, sizeof(T));
    return static_cast<double>(*p);
}
template <class T>
__device__ __forceinline__ void sto(T* __restrict__ base, uint32_t voff, double v) {
    T* p = reinterpret_cast<T*>(reinterpret_cast<char*>(base) + voff);
    JFA_ACCESS(p, sizeof(T));
    *p = static_cast<T>(v);
}

// Flux of one face along AXIS (stride st): window cells -3..2 relative to the cell right of
// the face, which sits at element rb (wave-uniform) + voff bytes (per lane) of the case block.
template <class T, int AXIS, int VAR, int STENCIL, int RIEMANN, bool FLIM>
__device__ __forceinline__ void face_flux(const T* __restrict__ W, const T* __restrict__ U, int64_t rb,
                                          uint32_t voff, int64_t st, int64_t fs, const Params& prm, double F[5]) {
    constexpr bool cons_win = (VAR == JFA_CONSERVATIVE || VAR == JFA_CHAR_CONSERVATIVE);
    const T* __restrict__ src = cons_win ? U : W;
    double win[6][5];
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int i = 0; i < 5; ++i) win[k][i] = ldo(src + (rb + i * fs + (int64_t)(k - 3) * st), voff);
    if constexpr (RIEMANN == kFluxSplit) {
        flux_split_face<AXIS, STENCIL>(win, prm, F);
    } else {
        double wl[5], wr[5];
        if constexpr (VAR == JFA_CHAR_CONSERVATIVE) {
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                wl[i] = ldo(W + (rb + i * fs - st), voff);
                wr[i] = ldo(W + (rb + i * fs), voff);
            }
        }
        double WL[5], WR[5], UL[5], UR[5];
        face_states<AXIS, VAR, STENCIL>(win, wl, wr, prm, WL, WR, UL, UR);
        if (prm.limiter) {
            double c[5];
#pragma unroll
            for (int i = 0; i < 5; ++i) c[i] = ldo(W + (rb + i * fs - st), voff);
            limit_side<VAR>(c, prm, WL, UL);
#pragma unroll
            for (int i = 0; i < 5; ++i) c[i] = ldo(W + (rb + i * fs), voff);
            limit_side<VAR>(c, prm, WR, UR);
        }
        riemann<AXIS, RIEMANN>(WL, WR, UL, UR, prm, F);
    }
    if constexpr (FLIM) {  // positivity flux limiter with the case's dt (set per item)
        double wl_[5], wr_[5], ul_[5], ur_[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            wl_[i] = ldo(W + (rb + i * fs - st), voff);
            wr_[i] = ldo(W + (rb + i * fs), voff);
            ul_[i] = ldo(U + (rb + i * fs - st), voff);
            ur_[i] = ldo(U + (rb + i * fs), voff);
        }
        limit_flux<AXIS>(F, wl_, wr_, ul_, ur_, flux_lambda<AXIS>(prm, prm.dt_val), prm);
    }
}

// One work item.  Only flux terms are carried from plane to plane (the x-face below the
// current plane and, in lane k, the extra y-face of plane x0+k); each face reloads its
// six-cell window, which the previous plane's loads leave in cache.  This keeps the
// register footprint near that of a single face evaluation.
template <class T, int VAR, int STENCIL, int RIEMANN, bool FLIM>
__device__ __forceinline__ void run_item(const BatchArgs<T>& a, const Params& prm_in, const Item2D it) {
    const Case2D cs = load_case(a.cases + it.c);
    // the flux limiter's lam = dt (1/dx) sigma uses this case's dt and cell sizes
    Params prm = prm_in;
    const double dx0 = 1.0 / cs.one_dx0, dx1 = 1.0 / cs.one_dx1;  // this case's cell sizes
    prm.dx2[0] = dx0 * dx0;
    prm.dx2[1] = dx1 * dx1;
    if constexpr (FLIM) {
        prm.dt_val = uni(a.dt[it.c]);
        const double sum = cs.one_dx0 + cs.one_dx1;
        prm.flim_sigma[0] = prm.flux_partition == JFA_PARTITION_CELLSIZE ? sum / cs.one_dx0 : 2.0;
        prm.flim_sigma[1] = prm.flux_partition == JFA_PARTITION_CELLSIZE ? sum / cs.one_dx1 : 2.0;
        prm.flim_odx[0] = cs.one_dx0;
        prm.flim_odx[1] = cs.one_dx1;
        prm.flim_odx[2] = 0.0;
        if (a.wave_max) prm.wave_max = a.wave_max + 3 * it.c;
    }
    const int lane = threadIdx.x & 63;
    const int y = it.y0 + lane;               // this lane's cell / lower face along y
    const int64_t fs = cs.fs;
    const int64_t sx = cs.s0;
    const T* __restrict__ W = a.Win + cs.off;
    const T* __restrict__ U = a.Uin + cs.off;
    const bool has_face = y <= cs.n1;         // faces 0..n1
    const bool has_cell = y < cs.n1;
    const int yc = has_face ? y : cs.n1;      // clamp so idle lanes address valid memory
    // byte offset of this lane's column (cell y = yc at x-buffer index 0); case blocks are
    // below 4 GiB (checked on the host)
    const uint32_t voff = (uint32_t)(yc + cs.nh) * (uint32_t)sizeof(T);
    // extra upper y-face (y0 + 64) of plane x0 + lane, needed by lane 63
    double Fext[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (it.y0 + 64 <= cs.n1 && lane < it.x1 - it.x0) {
        const uint32_t vext = ((uint32_t)lane * (uint32_t)sx + (uint32_t)(it.y0 + 64 + cs.nh)) * (uint32_t)sizeof(T);
        face_flux<T, 1, VAR, STENCIL, RIEMANN, FLIM>(W, U, (int64_t)(it.x0 + cs.nh) * sx, vext, 1, fs, prm, Fext);
    }
    // x-face x0 - 1/2
    double Fxp[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (has_cell) face_flux<T, 0, VAR, STENCIL, RIEMANN, FLIM>(W, U, (int64_t)(it.x0 + cs.nh) * sx, voff, sx, fs, prm, Fxp);
    double c0, c1, m;
    rk3_coeffs(a.stage, c0, c1, m);
    const double ts = uni(a.dt[it.c]) * m;
    const Geo G = geo_of(cs);
    double vmax = 0.0;
    for (int x = it.x0; x < it.x1; ++x) {
        const int64_t rb = (int64_t)(x + cs.nh) * sx;  // this plane, uniform part
        // opaque per plane: keeps the compiler from hoisting per-tap addresses out of the loop
        uint32_t vo = voff;
        JFA_OPAQUE(vo);
        // y-direction term: face below this lane's cell, face above from lane + 1 (lane 63:
        // the extra face held by lane x - x0)
        double dyt[5];
        {
            double Fy[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            if (has_face) face_flux<T, 1, VAR, STENCIL, RIEMANN, FLIM>(W, U, rb, vo, 1, fs, prm, Fy);
            const int src_lane = x - it.x0;
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                double up = __shfl_down(Fy[i], 1, 64);
                const double e = __shfl(Fext[i], src_lane, 64);
                if (lane == 63) up = e;
                dyt[i] = cs.one_dx1 * (Fy[i] - up);
            }
        }
        if (has_cell) {
            double Fxn[5];
            face_flux<T, 0, VAR, STENCIL, RIEMANN, FLIM>(W, U, rb + sx, vo, sx, fs, prm, Fxn);
            double u[5], w[5];
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                // rhs = (0 + x-term) + y-term, the reference's accumulation order
                const double rhs = (0.0 + cs.one_dx0 * (Fxp[i] - Fxn[i])) + dyt[i];
                Fxp[i] = Fxn[i];
                double xv = ldo(a.Uin + (cs.off + rb + i * fs), vo);
                if (a.stage > 0) xv = c0 * xv + c1 * ldo(a.Un + (cs.off + rb + i * fs), vo);
                u[i] = xv + ts * rhs;
            }
            prim_from_cons(u, prm, w);
            T* Wo = a.Wout + cs.off;
            T* Uo = a.Uout + cs.off;
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                sto(Wo + (rb + i * fs), vo, w[i]);
                sto(Uo + (rb + i * fs), vo, u[i]);
            }
            const int64_t bcell = rb + (int64_t)(yc + cs.nh);
            double uh[5];
            cons_from_prim(w, prm, uh);
            BC6 bc;
            bc.t[0] = cs.bc[0];
            bc.t[1] = cs.bc[1];
            bc.t[2] = cs.bc[2];
            bc.t[3] = cs.bc[3];
            bc.t[4] = JFA_BC_PERIODIC;
            bc.t[5] = JFA_BC_PERIODIC;
            const int c3[3] = {x, y, 0};
            halo_scatter(Wo, Uo, G, bc, c3, bcell, w, uh, prm);
            if (a.part) vmax = nan_max(vmax, wave_sum(w, G.active, prm.gamma));
        }
    }
    if (a.part) {  // a.part points at this item's slot
        for (int o = 32; o > 0; o >>= 1) vmax = nan_max(vmax, __shfl_xor(vmax, o, 64));
        if (lane == 0) a.part[0] = vmax;
    }
}

template <class T, int VAR, int STENCIL, int RIEMANN, bool FLIM>
// Three waves per SIMD (at most 168 VGPRs) for the variants that fit without spilling; the
// dense 5x5 projections of CHAR-CONSERVATIVE need more registers and keep two.
__global__ __launch_bounds__(kBatchBlock)
JFA_WAVES_PER_EU((VAR == JFA_CHAR_CONSERVATIVE || RIEMANN == kFluxSplit) ? 1 : 3)
void batch2d_stage_kernel(BatchArgs<T> a, Params prm) {
    const unsigned idx = blockIdx.x * (kBatchBlock / 64) + (threadIdx.x >> 6);  // one item per wavefront
    if (idx >= (unsigned)a.nitems) return;  // whole wavefronts only; no barriers follow
    const Item2D it = load_item(a.items + idx);
    BatchArgs<T> aa = a;
    if (a.part) aa.part = a.part + idx;
    run_item<T, VAR, STENCIL, RIEMANN, FLIM>(aa, prm, it);
}

// WAVESPEED maxima of every case (max over its interior of |u| + c and |v| + c) into
// slot[3 c], slot[3 c + 1]; grid.y runs over the cases.  slot must be zeroed first.
template <class T>
__global__ __launch_bounds__(kBlock) void batch2d_wave_max_kernel(const T* __restrict__ W,
                                                                  const Case2D* __restrict__ cases, double gamma,
                                                                  unsigned long long* slot) {
    const Case2D cs = cases[blockIdx.y];
    const T* __restrict__ Wc = W + cs.off;
    const uint32_t total = (uint32_t)cs.n0 * (uint32_t)cs.n1;
    double v0 = 0.0, v1 = 0.0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const uint32_t i = t / (uint32_t)cs.n1;
        const uint32_t j = t - i * (uint32_t)cs.n1;
        const int64_t b = (int64_t)(i + cs.nh) * cs.s0 + (int64_t)(j + cs.nh);
        const double c = sound_speed(ld(Wc, 4 * cs.fs + b), ld(Wc, b), gamma);
        v0 = nan_max(v0, fabs(ld(Wc, cs.fs + b)) + c);
        v1 = nan_max(v1, fabs(ld(Wc, 2 * cs.fs + b)) + c);
    }
    block_max_to(slot + 3 * blockIdx.y, v0);
    __syncthreads();  // block_max_to's shared partials are reused
    block_max_to(slot + 3 * blockIdx.y + 1, v1);
}

// Per-case CFL time step from the per-item partials (time_step_size.py:109,154-155).
template <int UNUSED = 0>  // template: shared by both storage-type translation units
__global__ __launch_bounds__(256) void batch2d_dt_kernel(const Case2D* __restrict__ cases,
                                                         const double* __restrict__ part, double cfl, double eps,
                                                         double* __restrict__ dt) {
    __shared__ double sm[4];
    const Case2D cs = cases[blockIdx.x];
    double v = 0.0;
    for (int i = cs.item_begin + threadIdx.x; i < cs.item_end; i += blockDim.x) v = nan_max(v, part[i]);
    for (int o = 32; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double mx = nan_max(nan_max(sm[0], sm[1]), nan_max(sm[2], sm[3]));
        dt[blockIdx.x] = cs.dx_min / (mx + eps) * cfl;
    }
}

// ------------------------------------------------------------------ host side
struct BatchLayout {
    size_t cases_off, items_off, parts_off, waves_off, total;
    int nitems;
};

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static inline int count_items(const jfa_case2d* cs, int n, int chunk) {
    int64_t tot = 0;
    for (int c = 0; c < n; ++c) tot += (int64_t)((cs[c].n[0] + chunk - 1) / chunk) * ((cs[c].n[1] + 63) / 64);
    return (int)tot;
}

static inline BatchLayout batch_layout(const jfa_case2d* cs, int n, int chunk) {
    BatchLayout L;
    L.nitems = count_items(cs, n, chunk);
    L.cases_off = 0;
    L.items_off = L.cases_off + align256(sizeof(Case2D) * (size_t)n);
    L.parts_off = L.items_off + align256(sizeof(Item2D) * (size_t)L.nitems);
    L.waves_off = L.parts_off + align256(sizeof(double) * (size_t)L.nitems);
    L.total = L.waves_off + align256(3 * sizeof(unsigned long long) * (size_t)n);
    return L;
}

template <class T, int VAR, int STENCIL, int RIEMANN>
static void launch_batch(const BatchArgs<T>& a, const Params& prm, hipStream_t st) {
    const int waves_per_block = kBatchBlock / 64;
    const int blocks = (a.nitems + waves_per_block - 1) / waves_per_block;
    // the positivity flux limiter is a compile-time variant, so the plain kernel keeps its registers
    if (prm.flux_limiter)
        hipLaunchKernelGGL((batch2d_stage_kernel<T, VAR, STENCIL, RIEMANN, true>), dim3(blocks), dim3(kBatchBlock), 0,
                           st, a, prm);
    else
        hipLaunchKernelGGL((batch2d_stage_kernel<T, VAR, STENCIL, RIEMANN, false>), dim3(blocks), dim3(kBatchBlock),
                           0, st, a, prm);
}

template <class T, int VAR, int STENCIL>
static void launch_batch_r(int riemann, const BatchArgs<T>& a, const Params& prm, hipStream_t st) {
    if (riemann == JFA_HLLC) launch_batch<T, VAR, STENCIL, JFA_HLLC>(a, prm, st);
    else if (riemann >= JFA_FS_ROE) {  // flux splitting reads the primitive window
        if constexpr (VAR == JFA_PRIMITIVE) launch_batch<T, VAR, STENCIL, kFluxSplit>(a, prm, st);
    } else if (riemann != JFA_RUSANOV) launch_batch<T, VAR, STENCIL, kRiemannExt>(a, prm, st);
    else launch_batch<T, VAR, STENCIL, JFA_RUSANOV>(a, prm, st);
}
template <class T, int VAR>
static void launch_batch_s(int stencil, int riemann, const BatchArgs<T>& a, const Params& prm, hipStream_t st) {
    if (stencil == JFA_WENO5_JS) launch_batch_r<T, VAR, JFA_WENO5_JS>(riemann, a, prm, st);
    else if (stencil == JFA_WENO5_Z) launch_batch_r<T, VAR, JFA_WENO5_Z>(riemann, a, prm, st);
    else launch_batch_r<T, VAR, kStencilExt>(riemann, a, prm, st);
}

template <class T>
int batch2d_stage_impl(const void* W_in, const void* U_in, const void* U_n, void* W_out, void* U_out,
                       const jfa_scheme* s, int stage, const double* dt_dev, double cfl, double eps, double* dt_next,
                       int ncases, int nitems, char* ws, const BatchLayout& L, hipStream_t st) {
    BatchArgs<T> a;
    a.Win = static_cast<const T*>(W_in);
    a.Uin = static_cast<const T*>(U_in);
    a.Un = static_cast<const T*>(U_n);
    a.Wout = static_cast<T*>(W_out);
    a.Uout = static_cast<T*>(U_out);
    a.cases = reinterpret_cast<const Case2D*>(ws + L.cases_off);
    a.items = reinterpret_cast<const Item2D*>(ws + L.items_off);
    a.nitems = nitems;
    a.stage = stage;
    a.dt = dt_dev;
    a.part = dt_next ? reinterpret_cast<double*>(ws + L.parts_off) : nullptr;
    a.wave_max = nullptr;
    Params prm = make_params(s);
    if (s->flux_limiter && s->flux_partition == JFA_PARTITION_WAVESPEED) {
        unsigned long long* slots = reinterpret_cast<unsigned long long*>(ws + L.waves_off);
        hipLaunchKernelGGL(zero_slots_kernel<0>, dim3(1), dim3(256), 0, st, slots, 3 * ncases);
        hipLaunchKernelGGL((batch2d_wave_max_kernel<T>), dim3(64, (unsigned)ncases), dim3(kBlock), 0, st, a.Win,
                           a.cases, s->gamma, slots);
        a.wave_max = slots;
        prm.corner_speed = prm.eps_precision + sqrt(s->gamma);  // 2-D cases: the halo corners' eps state
    }
    switch (s->riemann >= JFA_FS_ROE ? JFA_PRIMITIVE : s->recon_var) {
        case JFA_PRIMITIVE: launch_batch_s<T, JFA_PRIMITIVE>(s->stencil, s->riemann, a, prm, st); break;
        case JFA_CONSERVATIVE: launch_batch_s<T, JFA_CONSERVATIVE>(s->stencil, s->riemann, a, prm, st); break;
        case JFA_CHAR_PRIMITIVE: launch_batch_s<T, JFA_CHAR_PRIMITIVE>(s->stencil, s->riemann, a, prm, st); break;
        default: launch_batch_s<T, JFA_CHAR_CONSERVATIVE>(s->stencil, s->riemann, a, prm, st); break;
    }
    int rc = check_launch();
    if (rc) return rc;
    if (dt_next) {
        hipLaunchKernelGGL(batch2d_dt_kernel<0>, dim3(ncases), dim3(256), 0, st, a.cases, a.part, cfl, eps, dt_next);
        rc = check_launch();
    }
    return rc;
}

}  // namespace jfa
