// jfa_device.h -- device-side arithmetic of the convective-flux path (gfx950).
//
// Each function restates the JAX-Fluids expression it cites (src/jaxfluids/...),
// evaluated per face in registers.  The FP64 path is VALU bound, so the algebra
// is rearranged to cut divisions and square roots while keeping every value equal
// to the reference's up to rounding:
//  * reciprocals and inverse square roots use the hardware estimates
//    (v_rcp_f64 / v_rsq_f64) refined by two Newton-Raphson steps (relative error
//    ~1 ulp for the normal, positive operands of this path);
//  * WENO weights are normalised with one reciprocal: omega_k = a_k / sum(a), with
//    a_k multiplied through by the product of the other two denominators;
//  * E = rho|u|^2/2 + p/(gamma-1) (no division), the characteristic projection
//    uses c^2 rho = gamma p and a single rsq, the HLLC star state one reciprocal.
// Velocity components are addressed through compile-time axis tables so the
// three sweep directions compile to straight-line register code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/jaxfluids_amd.h"

namespace jfa {

// equation_information.py:110 velocity_minor_axes, per normal axis
template <int AXIS>
struct Ax {
    static constexpr int v = AXIS + 1;
    static constexpr int m0 = AXIS == 0 ? 2 : (AXIS == 1 ? 3 : 1);
    static constexpr int m1 = AXIS == 0 ? 3 : (AXIS == 1 ? 1 : 2);
};

// Temperature law of one transport property (jfa_transport's JFA_LAW_*):
// POWER ref (T / T_ref)^param, SUTHERLAND ref ((T_ref + param) / (T + param)) (T / T_ref)^(3/2).
struct Law {
    int kind;
    double ref;
    double T_ref;
    double inv_T_ref;
    double param;
};

struct Params {
    double gamma;
    double gm1;        // gamma - 1
    double inv_gm1;    // 1 / (gamma - 1)
    double eps;        // WENO epsilon (config/precision.py:39,53)
    double eps_z;      // 4 eps: epsilon of the scaled WENO-Z smoothness indicators (4 beta)
    double eps_js;     // 16 eps: epsilon of the scaled WENO-JS indicators (squared)
    int stencil;       // JFA_WENO5_JS ... JFA_TENO5_A (read by the kStencilExt kernels)
    double eps_precision;  // machine epsilon of the working precision (config/precision.py:36,50)
    int riemann;           // JFA_HLL / JFA_HLLC_LM / JFA_AUSMP (read by the kRiemannExt kernels)
    int signal_speed;  // JFA_SS_*
    int frozen;        // JFA_FROZEN_*
    // dissipative fluxes (jfa_transport); dissip = 0: inviscid
    int dissip;        // bit 0: viscous stresses, bit 1: heat flux
    int diss_halos;    // 1: the dissipative stencils read face and edge halos; 0: they wrap periodically
    double R;          // specific gas constant
    Law mu;            // dynamic viscosity mu(T)
    Law kappa;         // thermal conductivity kappa(T); kind JFA_LAW_PRANDTL: kappa_fac mu
    double bulk;       // bulk viscosity
    double kappa_fac;  // cp / Pr (JFA_LAW_PRANDTL)
    double inv_cp;     // 1 / cp
    // positivity limiter of the reconstructed face states (jfa_scheme.interpolation_limiter)
    int limiter;
    int limit_velocity;  // replace all five primitives, not only density and pressure
    double lim_eps_rho;
    double lim_eps_p;
    // positivity flux limiter (jfa_scheme.flux_limiter): JFA_FLUX_LIMITER_SIMPLE / _NASA, 0 off
    int flux_limiter;
    int flux_partition;     // JFA_PARTITION_*
    double flim_eps_rho;
    double flim_eps_p;
    double flim_sigma[3];   // 1/sigma of the flux partition per axis (UNIFORM, CELLSIZE)
    double flim_odx[3];     // 1/dx of the active axes, 0 on inactive ones (WAVESPEED partition)
    const unsigned long long* wave_max;  // WAVESPEED: bits of max |u_a| + c over the interior, per axis
    double corner_speed;    // WAVESPEED: |u_a| + c of the halo corners' eps state (0 on 1-D grids)
    double dx2[3];          // squared cell size per axis (the WENO6-CUM stencils read it)
    const double* dt_ptr;   // the step's time step in device memory, or null: dt_val
    double dt_val;
};

// ---------------------------------------------------------------- refined hardware estimates
// v_rcp_f64 / v_rsq_f64 return estimates with relative error below 2^-24 (measured on the
// device over 1e-30..1e30); each Newton-Raphson step squares the error.  One step (~2^-48)
// is used inside the face flux, two steps (full precision) where a value is stored in the
// state; a / b and sqrt end with a residual correction, which restores full precision.
__device__ __forceinline__ double rcp1(double x) {  // 1/x, one Newton step
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double rcp(double x) {  // 1/x, two Newton steps
    const double r = rcp1(x);
    return fma(r, fma(-x, r, 1.0), r);
}
// a / b with a final residual correction q <- q + r(a - b q).
__device__ __forceinline__ double fdiv(double a, double b) {
    const double r = rcp1(b);
    const double q = a * r;
    return fma(r, fma(-b, q, a), q);
}
// 1/sqrt(x), one Newton step y <- y + y(1/2 - (x/2) y^2)
__device__ __forceinline__ double rsq1(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    return fma(y, fma(-(0.5 * x) * y, y, 0.5), y);
}
// sqrt(x) = x rsq(x) with one residual correction g <- g + (y/2)(x - g^2).
__device__ __forceinline__ double fsqrt(double x) {
    const double y = rsq1(x);
    const double g = x * y;
    return fma(0.5 * y, fma(-g, g, x), g);
}

// ---------------------------------------------------------------- WENO5
// weno5_base.py:34-50 smoothness / candidate polynomials; weno5_js.py:39-58; weno5_z.py:39-49.
//
// Written in the first differences d_j = u_{j+1} - u_j of the stencil u_imm..u_ipp (c = u_i):
//   beta_k = 13/12 D_k^2 + 1/4 S_k^2 with the second differences D = (d1-d0, d2-d1, d3-d2) and
//   the one-sided slopes S = (3 d1 - d0, d1 + d2, d3 - 3 d2), evaluated as
//   4 beta_k = (13/3) D_k^2 + S_k^2; the weights are unchanged if the epsilon is scaled alike
//   (WENO-Z: tau5/(beta_k + eps) -> 4 eps; WENO-JS: 1/(beta_k^2 + eps) -> 16 eps).
//   Candidate k is p_k = c + delta_k, delta = ((5 d1 - 2 d0)/6, (d1 + 2 d2)/6, (4 d2 - d3)/6),
//   so the state is c + sum_k omega_k delta_k; the linear weights (1/10, 6/10, 3/10) are folded
//   into g_k = d_k delta_k.
constexpr double kQ = 13.0 / 3.0;

// c + sum_k omega_k delta_k from the scaled indicators b0..b2 and g0..g2.  With q_k = b_k + eps
// (WENO-Z) the weights are omega_k ~ d_k (1 + tau5/q_k) ~ d_k e_k, e_k = P + tau5 prod_{j != k} q_j
// and P = q0 q1 q2; with q_k = b_k^2 + eps (WENO-JS) omega_k ~ d_k / q_k ~ d_k prod_{j != k} q_j.
template <int STENCIL>
__device__ __forceinline__ double weno5_combine(double b0, double b1, double b2, double g0, double g1, double g2,
                                                double c, const double eps_z, const double eps_js) {
    double e0, e1, e2;
    if constexpr (STENCIL == JFA_WENO5_JS) {
        const double q0 = fma(b0, b0, eps_js), q1 = fma(b1, b1, eps_js), q2 = fma(b2, b2, eps_js);
        e0 = q1 * q2;
        e1 = q0 * q2;
        e2 = q0 * q1;
    } else {
        const double tau5 = fabs(b0 - b2);
        const double q0 = b0 + eps_z, q1 = b1 + eps_z, q2 = b2 + eps_z;
        const double p12 = q1 * q2, p02 = q0 * q2, p01 = q0 * q1;
        const double P = p01 * q2;
        e0 = fma(tau5, p12, P);
        e1 = fma(tau5, p02, P);
        e2 = fma(tau5, p01, P);
    }
    const double num = fma(e2, g2, fma(e1, g1, e0 * g0));
    const double den = fma(0.3, e2, fma(0.6, e1, 0.1 * e0));
    return fma(num, rcp1(den), c);
}

// One side: (u_imm, u_im, u_i, u_ip, u_ipp).
template <int STENCIL>
__device__ __forceinline__ double weno5(double a, double b, double c, double d, double e, const double eps_z,
                                        const double eps_js) {
    const double d0 = b - a, d1 = c - b, d2 = d - c, d3 = e - d;
    const double D0 = d1 - d0, D1 = d2 - d1, D2 = d3 - d2;
    const double S0 = fma(3.0, d1, -d0), S1 = d1 + d2, S2 = fma(-3.0, d2, d3);
    const double b0 = fma(S0, S0, (kQ * D0) * D0), b1 = fma(S1, S1, (kQ * D1) * D1), b2 = fma(S2, S2, (kQ * D2) * D2);
    const double g0 = fma(1.0 / 12.0, d1, (-1.0 / 30.0) * d0);
    const double g1 = fma(0.2, d2, 0.1 * d1);
    const double g2 = fma(0.2, d2, -0.05 * d3);
    return weno5_combine<STENCIL>(b0, b1, b2, g0, g1, g2, c, eps_z, eps_js);
}

// Left (j = 0: cells -3..1) and right (j = 1: cells 2..-2) states of one face from its
// 6-cell window v[0..5] = cells -3..2 (weno5_base.py:30-31).  The first differences, the
// second differences of cells -1 and 0 and their scaled squares are shared by the two sides;
// the right side's stencil is the window reversed, whose first differences are -d4..-d1.
template <int STENCIL>
__device__ __forceinline__ void weno5_pair(const double v[6], const double eps_z, const double eps_js, double& L,
                                           double& R) {
    const double d0 = v[1] - v[0], d1 = v[2] - v[1], d2 = v[3] - v[2], d3 = v[4] - v[3], d4 = v[5] - v[4];
    const double E0 = d1 - d0, E1 = d2 - d1, E2 = d3 - d2, E3 = d4 - d3;
    const double Q1 = (kQ * E1) * E1, Q2 = (kQ * E2) * E2;
    {  // left: stencil v0..v4, centre v2
        const double S0 = fma(3.0, d1, -d0), S1 = d1 + d2, S2 = fma(-3.0, d2, d3);
        const double b0 = fma(S0, S0, (kQ * E0) * E0), b1 = fma(S1, S1, Q1), b2 = fma(S2, S2, Q2);
        const double g0 = fma(1.0 / 12.0, d1, (-1.0 / 30.0) * d0);
        const double g1 = fma(0.2, d2, 0.1 * d1);
        const double g2 = fma(0.2, d2, -0.05 * d3);
        L = weno5_combine<STENCIL>(b0, b1, b2, g0, g1, g2, v[2], eps_z, eps_js);
    }
    {  // right: stencil v5..v1, centre v3; its differences are (-d4, -d3, -d2, -d1)
        const double S0 = fma(-3.0, d3, d4), S1 = d3 + d2, S2 = fma(3.0, d2, -d1);
        const double b0 = fma(S0, S0, (kQ * E3) * E3), b1 = fma(S1, S1, Q2), b2 = fma(S2, S2, Q1);
        const double g0 = fma(-1.0 / 12.0, d3, (1.0 / 30.0) * d4);
        const double g1 = fma(-0.2, d2, -0.1 * d3);
        const double g2 = fma(-0.2, d2, 0.05 * d1);
        R = weno5_combine<STENCIL>(b0, b1, b2, g0, g1, g2, v[3], eps_z, eps_js);
    }
}

// ---------------------------------------------------------------- further stencils
// Stencils that read a sub-window of the same 6 cells, selected at run time (prm.stencil) in
// the kStencilExt instantiation of the kernels, so the WENO5 kernels keep their code.
//   WENO1 (weno/weno1_js.py:25-30): the adjacent cell.
//   WENO3-JS / WENO3-Z (weno3_base.py:20-45, weno/weno3_js.py:15-42, weno/weno3_z.py:21-41):
//     beta = ((u_i - u_im)^2, (u_ip - u_i)^2), d = (1/3, 2/3), p = (u_i + dm/2, u_i + dp/2).
//   TENO5 (teno/teno5.py:32-73): WENO5 indicators and candidates, gamma_k = (1 + tau5/(beta_k + eps))^6,
//     sharp cutoff gamma_k / sum gamma < CT = 1e-5, d = (0.05, 0.55, 0.40).
//   TENO5-A (teno/teno5_a.py:42-99): the same with the adaptive cutoff CT = 10^-(ceil(10 - 5 (1 - g)) - 1),
//     g from the smoothness ratio eta of neighbouring differences, and WENO5's d = (1/10, 6/10, 3/10).
constexpr int kStencilExt = 100;
// WENO7-JS reads an 8-cell window (cells -4..3 of a face), so it has instantiations of its own.
constexpr int kStencilW7 = 101;
// Face window of a stencil instantiation: n cells, -h..h-1 relative to the cell right of the face.
template <int STENCIL>
struct Win {
    static constexpr int n = STENCIL == kStencilW7 ? 8 : 6;
    static constexpr int h = n / 2;
};

// One WENO3 side: stencil (u_im, u_i, u_ip); Z selects WENO3-Z.
template <bool Z>
__device__ __forceinline__ double weno3(double um, double ui, double up, double eps) {
    const double dm = ui - um, dp = up - ui;
    const double b0 = dm * dm, b1 = dp * dp;
    double e0, e1;  // d_k alpha_k q0 q1
    if constexpr (Z) {
        const double tau3 = fabs(b0 - b1);
        const double q0 = b0 + eps, q1 = b1 + eps;
        e0 = (1.0 / 3.0) * ((q0 + tau3) * q1);
        e1 = (2.0 / 3.0) * ((q1 + tau3) * q0);
    } else {
        e0 = (1.0 / 3.0) * fma(b1, b1, eps);
        e1 = (2.0 / 3.0) * fma(b0, b0, eps);
    }
    return fma(0.5 * fma(e1, dp, e0 * dm), rcp1(e0 + e1), ui);
}

// One TENO5 side: stencil (u_imm, u_im, u_i, u_ip, u_ipp); A selects TENO5-A.
// The TENO5 weights w_k = d_k or 0 (the sharp cutoff) from the first differences of the stencil.
template <bool A>
__device__ __forceinline__ void teno5_weights(double d0, double d1, double d2, double d3, const Params& prm,
                                              double& w0, double& w1, double& w2) {
    const double D0 = d1 - d0, D1 = d2 - d1, D2 = d3 - d2;
    const double S0 = fma(3.0, d1, -d0), S1 = d1 + d2, S2 = fma(-3.0, d2, d3);
    // 4 beta_k and 4 eps: tau5 / (beta_k + eps) is unchanged
    const double b0 = fma(S0, S0, (kQ * D0) * D0), b1 = fma(S1, S1, (kQ * D1) * D1), b2 = fma(S2, S2, (kQ * D2) * D2);
    const double tau5 = fabs(b0 - b2);
    double x0 = fma(tau5, rcp(b0 + prm.eps_z), 1.0), x1 = fma(tau5, rcp(b1 + prm.eps_z), 1.0),
           x2 = fma(tau5, rcp(b2 + prm.eps_z), 1.0);
    x0 = x0 * x0;
    x1 = x1 * x1;
    x2 = x2 * x2;
    const double g0 = x0 * x0 * x0, g1 = x1 * x1 * x1, g2 = x2 * x2 * x2;
    double CT = 1e-5;
    if constexpr (A) {
        const double ed = 0.9 * 0.24 / (1.0 - 0.24) * 1e-6;  // eps_dissipation = 0.9 Cr/(1-Cr) xi^2
        const double eta0 = fdiv(fabs(2.0 * d1 * d0) + ed, fma(d1, d1, d0 * d0) + ed);
        const double eta1 = fdiv(fabs(2.0 * d2 * d1) + ed, fma(d2, d2, d1 * d1) + ed);
        const double eta2 = fdiv(fabs(2.0 * d3 * d2) + ed, fma(d3, d3, d2 * d2) + ed);
        const double eta = fmin(eta0, fmin(eta1, eta2));
        const double m = 1.0 - fmin(1.0, fdiv(eta, 0.24));
        const double om = 1.0 - m, om2 = om * om;
        const double g = om2 * om2 * fma(4.0, m, 1.0);
        const int bb = (int)ceil(10.0 - 5.0 * (1.0 - g)) - 1;  // 4..9
        constexpr double kPow10[10] = {1.0, 1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9};
        CT = kPow10[bb < 0 ? 0 : (bb > 9 ? 9 : bb)];
    }
    const double thr = CT * (g0 + g1 + g2);
    constexpr double w_[3] = {A ? 0.1 : 0.05, A ? 0.6 : 0.55, A ? 0.3 : 0.40};
    w0 = g0 < thr ? 0.0 : w_[0];
    w1 = g1 < thr ? 0.0 : w_[1];
    w2 = g2 < thr ? 0.0 : w_[2];
}

template <bool A>
__device__ __forceinline__ double teno5(double a, double b, double c, double d, double e, const Params& prm) {
    const double d0 = b - a, d1 = c - b, d2 = d - c, d3 = e - d;
    double w0, w1, w2;
    teno5_weights<A>(d0, d1, d2, d3, prm, w0, w1, w2);
    const double del0 = (5.0 * d1 - 2.0 * d0) * (1.0 / 6.0), del1 = fma(2.0, d2, d1) * (1.0 / 6.0),
                 del2 = fma(4.0, d2, -d3) * (1.0 / 6.0);
    const double num = fma(w2, del2, fma(w1, del1, w0 * del0));
    return fma(num, rcp(w0 + w1 + w2 + prm.eps), c);
}

// Six-point stencils (weno6_base.py:20-59, teno6_base.py:20-64): the 6-point smoothness indicator is
// a quadratic form with 8-digit integer coefficients whose terms cancel, so these functions keep
// the reference's operation order, without contraction into FMAs and with IEEE divisions.  They are
// called, not inlined, so the many kernel instantiations that can reach them compile quickly.
__device__ __forceinline__ double beta_six_point(double a, double b, double c, double d, double e, double f) {
#pragma clang fp contract(off)
    return 1.0 / 10080 / 12 *
           (271779 * a * a + a * (-2380800 * b + 4086352 * c - 3462252 * d + 1458762 * e - 245620 * f) +
            b * (5653317 * b - 20427884 * c + 17905032 * d - 7727988 * e + 1325006 * f) +
            c * (19510972 * c - 35817664 * d + 15929912 * e - 2792660 * f) +
            d * (17195652 * d - 15880404 * e + 2863984 * f) + e * (3824847 * e - 1429976 * f) + 139633 * f * f);
}

__device__ __forceinline__ void beta_three_point(double a, double b, double c, double d, double e, double& b0,
                                                 double& b1, double& b2) {
#pragma clang fp contract(off)
    const double s0 = a - 2 * b + c, t0 = a - 4 * b + 3 * c;
    const double s1 = b - 2 * c + d, t1 = b - d;
    const double s2 = c - 2 * d + e, t2 = 3 * c - 4 * d + e;
    b0 = 13.0 / 12.0 * (s0 * s0) + 1.0 / 4.0 * (t0 * t0);
    b1 = 13.0 / 12.0 * (s1 * s1) + 1.0 / 4.0 * (t1 * t1);
    b2 = 13.0 / 12.0 * (s2 * s2) + 1.0 / 4.0 * (t2 * t2);
}

// WENO6-CU (weno/weno6_cu.py:24-61; Hu, Wang & Adams 2010): d = (1/20, 9/20, 9/20, 1/20), C = 20,
// tau_6 = beta_3 - (beta_0 + 4 beta_1 + beta_2)/6, alpha_k = d_k (C + tau_6/(beta_k + eps)).
__device__ __noinline__ double weno6_cu(double a, double b, double c, double d, double e, double f, double eps) {
#pragma clang fp contract(off)
    double b0, b1, b2;
    beta_three_point(a, b, c, d, e, b0, b1, b2);
    const double b3 = beta_six_point(a, b, c, d, e, f);
    const double tau6 = b3 - 1.0 / 6 * (b0 + 4 * b1 + b2);
    const double a0 = 1.0 / 20 * (20 + tau6 / (b0 + eps)), a1 = 9.0 / 20 * (20 + tau6 / (b1 + eps));
    const double a2 = 9.0 / 20 * (20 + tau6 / (b2 + eps)), a3 = 1.0 / 20 * (20 + tau6 / (b3 + eps));
    const double one_alpha = 1.0 / (a0 + a1 + a2 + a3);
    const double p0 = 1.0 / 3 * a + -7.0 / 6 * b + 11.0 / 6 * c;
    const double p1 = -1.0 / 6 * b + 5.0 / 6 * c + 1.0 / 3 * d;
    const double p2 = 1.0 / 3 * c + 5.0 / 6 * d + -1.0 / 6 * e;
    const double p3 = 11.0 / 6 * d + -7.0 / 6 * e + 1.0 / 3 * f;
    return a0 * one_alpha * p0 + a1 * one_alpha * p1 + a2 * one_alpha * p2 + a3 * one_alpha * p3;
}

// WENO6-CUM1 / WENO6-CUM2 (weno/weno6_cum1.py:28-67, weno6_cum2.py:29-72; Hu & Adams 2011): the
// WENO6-CU candidates and d with C_q = 1000, q = 4 (the fourth power by squaring, as JAX's
// integer_pow).  CUM1: alpha_k = d_k (C_q + tau_6/(beta_k + eps dx^2))^4.  CUM2: eps_ = 1e-8,
// chi = 1/eps_, beta_ave = (beta_0 + beta_2 + 4 beta_1)/6, tau_6 = beta_3 - beta_ave and
// alpha_k = d_k (C_q + tau_6/(beta_k + eps_ dx^2) (beta_ave + chi dx^2)/(beta_k + chi dx^2))^4.
// dx2: the squared cell size along the axis.  Reference operation order, no contraction.
template <bool CUM2>
__device__ __noinline__ double weno6_cum(double a, double b, double c, double d, double e, double f, double eps,
                                         double dx2) {
#pragma clang fp contract(off)
    double b0, b1, b2;
    beta_three_point(a, b, c, d, e, b0, b1, b2);
    const double b3 = beta_six_point(a, b, c, d, e, f);
    const double bk[4] = {b0, b1, b2, b3};
    const double dr[4] = {1.0 / 20, 9.0 / 20, 9.0 / 20, 1.0 / 20};
    double al[4];
    if constexpr (!CUM2) {
        const double tau6 = b3 - 1.0 / 6 * (b0 + 4 * b1 + b2);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double x = 1000.0 + tau6 / (bk[k] + eps * dx2);
            const double x2 = x * x;
            al[k] = dr[k] * (x2 * x2);
        }
    } else {
        const double eps_ = 1e-8, chi = 1.0 / 1e-8;
        const double bave = 1.0 / 6 * (b0 + b2 + 4 * b1);
        const double tau6 = b3 - bave;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double x = 1000.0 + tau6 / (bk[k] + eps_ * dx2) * (bave + chi * dx2) / (bk[k] + chi * dx2);
            const double x2 = x * x;
            al[k] = dr[k] * (x2 * x2);
        }
    }
    const double one_alpha = 1.0 / (al[0] + al[1] + al[2] + al[3]);
    const double p0 = 1.0 / 3 * a + -7.0 / 6 * b + 11.0 / 6 * c;
    const double p1 = -1.0 / 6 * b + 5.0 / 6 * c + 1.0 / 3 * d;
    const double p2 = 1.0 / 3 * c + 5.0 / 6 * d + -1.0 / 6 * e;
    const double p3 = 11.0 / 6 * d + -7.0 / 6 * e + 1.0 / 3 * f;
    return al[0] * one_alpha * p0 + al[1] * one_alpha * p1 + al[2] * one_alpha * p2 + al[3] * one_alpha * p3;
}

// x^6 as the reference's x ** q with the Python int q = 6 evaluates it (lax.integer_pow, binary
// powering): x^2 (x^2 x^2).
__device__ __forceinline__ double ipow6(double x) {
    const double x2 = x * x;
    return x2 * (x2 * x2);
}

// TENO6 (teno/teno6.py:27-67, teno6_base.py:20-64): three 3-point candidates and a 4-point one
// (3/12, 13/12, -5/12, 1/12) with its own indicator, tau_6 = |beta_6 - (beta_0 + 4 beta_1 + beta_2)/6|,
// gamma_k = (1 + tau_6/(beta_k + eps))^6, sharp cutoff 1e-7, d = (0.05, 0.45, 0.30, 0.20).
__device__ __noinline__ double teno6(double a, double b, double c, double d, double e, double f, double eps) {
#pragma clang fp contract(off)
    double b0, b1, b2;
    beta_three_point(a, b, c, d, e, b0, b1, b2);
    const double b3 = 1.0 / 240.0 *
                      (c * (2107 * c - 9402 * d + 7042 * e - 1854 * f) + d * (11003 * d - 17246 * e + 4642 * f) +
                       e * (7043 * e - 3882 * f) + 547 * f * f);
    const double b6 = beta_six_point(a, b, c, d, e, f);
    const double tau6 = fabs(b6 - 1.0 / 6 * (b0 + 4 * b1 + b2));
    const double g0 = ipow6(1.0 + tau6 / (b0 + eps)), g1 = ipow6(1.0 + tau6 / (b1 + eps));
    const double g2 = ipow6(1.0 + tau6 / (b2 + eps)), g3 = ipow6(1.0 + tau6 / (b3 + eps));
    const double one_gamma_sum = 1.0 / (g0 + g1 + g2 + g3);
    const double w0 = g0 * one_gamma_sum < 1e-7 ? 0.0 : 0.050, w1 = g1 * one_gamma_sum < 1e-7 ? 0.0 : 0.450;
    const double w2 = g2 * one_gamma_sum < 1e-7 ? 0.0 : 0.300, w3 = g3 * one_gamma_sum < 1e-7 ? 0.0 : 0.200;
    const double one_dk = 1.0 / (w0 + w1 + w2 + w3 + eps);
    const double p0 = 1.0 / 3 * a + -7.0 / 6 * b + 11.0 / 6 * c;
    const double p1 = -1.0 / 6 * b + 5.0 / 6 * c + 1.0 / 3 * d;
    const double p2 = 1.0 / 3 * c + 5.0 / 6 * d + -1.0 / 6 * e;
    const double p3 = 3.0 / 12 * c + 13.0 / 12 * d + -5.0 / 12 * e + 1.0 / 12 * f;
    return w0 * one_dk * p0 + w1 * one_dk * p1 + w2 * one_dk * p2 + w3 * one_dk * p3;
}

// TENO6-A (teno/teno6_a.py:20-131): TENO6's candidates with |beta_3| and |beta_6|, gamma_k =
// (1 + tau_6/(beta_k + eps))^6, the spectrally optimised d and the adaptive cutoff
// CT = 10^-(ceil(10.5 - 4.5 (1 - g)) - 1) (5..10) from eta, the minimum over the four neighbouring
// difference pairs, with xi = 1e-3 and Cr = 0.17; no eps in the weight normalisation.
__device__ __noinline__ double teno6_a(double a, double b, double c, double d, double e, double f, double eps) {
#pragma clang fp contract(off)
    double b0, b1, b2;
    beta_three_point(a, b, c, d, e, b0, b1, b2);
    const double b3 = fabs(1.0 / 240.0 *
                           (c * (2107 * c - 9402 * d + 7042 * e - 1854 * f) + d * (11003 * d - 17246 * e + 4642 * f) +
                            e * (7043 * e - 3882 * f) + 547 * f * f));
    const double b6 = fabs(beta_six_point(a, b, c, d, e, f));
    const double tau6 = fabs(b6 - 1.0 / 6 * (b0 + 4 * b1 + b2));
    const double g0 = ipow6(1.0 + tau6 / (b0 + eps)), g1 = ipow6(1.0 + tau6 / (b1 + eps));
    const double g2 = ipow6(1.0 + tau6 / (b2 + eps)), g3 = ipow6(1.0 + tau6 / (b3 + eps));
    const double one_gamma_sum = 1.0 / (g0 + g1 + g2 + g3);
    const double ed = 0.9 * 0.17 / (1 - 0.17) * (1e-3 * 1e-3);  // eps_dissipation = 0.9 Cr/(1-Cr) xi^2
    const double dd[5] = {b - a, c - b, d - c, e - d, f - e};
    double eta = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double ek = (fabs(2 * dd[k + 1] * dd[k]) + ed) / (dd[k + 1] * dd[k + 1] + dd[k] * dd[k] + ed);
        eta = k == 0 ? ek : fmin(eta, ek);
    }
    const double m = 1 - fmin(1.0, eta / 0.17);
    const double om = 1 - m, om2 = om * om;
    const double g = om2 * om2 * (1 + 4 * m);
    const int bb = (int)ceil(10.5 - 4.5 * (1 - g)) - 1;  // 5..10
    constexpr double kPow10[11] = {1.0, 1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10};
    const double CT = kPow10[bb < 0 ? 0 : (bb > 10 ? 10 : bb)];
    const double w0 = g0 * one_gamma_sum < CT ? 0.0 : 0.0855682281039113;
    const double w1 = g1 * one_gamma_sum < CT ? 0.0 : 0.4294317718960898;
    const double w2 = g2 * one_gamma_sum < CT ? 0.0 : 0.1727270875843552;
    const double w3 = g3 * one_gamma_sum < CT ? 0.0 : 0.3122729124156450;
    const double one_dk = 1.0 / (w0 + w1 + w2 + w3);
    const double p0 = 1.0 / 3 * a + -7.0 / 6 * b + 11.0 / 6 * c;
    const double p1 = -1.0 / 6 * b + 5.0 / 6 * c + 1.0 / 3 * d;
    const double p2 = 1.0 / 3 * c + 5.0 / 6 * d + -1.0 / 6 * e;
    const double p3 = 3.0 / 12 * c + 13.0 / 12 * d + -5.0 / 12 * e + 1.0 / 12 * f;
    return w0 * one_dk * p0 + w1 * one_dk * p1 + w2 * one_dk * p2 + w3 * one_dk * p3;
}

// WENO7-JS (weno/weno7_js.py:10-90; Balsara & Shu 2000), one side from cells (u_-3 .. u_3) around
// the upwind cell u_0: four 4-point candidates, indicators with integer coefficients whose terms
// cancel, alpha_k = d_k / (beta_k^2 + eps) with d = (1, 12, 18, 4)/35.  The reference's operation
// order, no contraction, IEEE divisions, as the six-point stencils.
__device__ __noinline__ double weno7_js(double a, double b, double c, double d, double e, double f, double g,
                                        double eps) {
#pragma clang fp contract(off)
    const double b0 = a * (547 * a - 3882 * b + 4642 * c - 1854 * d) + b * (7043 * b - 17246 * c + 7042 * d) +
                      c * (11003 * c - 9402 * d) + d * (2107 * d);
    const double b1 = b * (267 * b - 1642 * c + 1602 * d - 494 * e) + c * (2843 * c - 5966 * d + 1922 * e) +
                      d * (3443 * d - 2522 * e) + e * (547 * e);
    const double b2 = c * (547 * c - 2522 * d + 1922 * e - 494 * f) + d * (3443 * d - 5966 * e + 1602 * f) +
                      e * (2843 * e - 1642 * f) + f * (267 * f);
    const double b3 = d * (2107 * d - 9402 * e + 7042 * f - 1854 * g) + e * (11003 * e - 17246 * f + 4642 * g) +
                      f * (7043 * f - 3882 * g) + g * (547 * g);
    const double o0 = 1.0 / (b0 * b0 + eps), o1 = 1.0 / (b1 * b1 + eps);
    const double o2 = 1.0 / (b2 * b2 + eps), o3 = 1.0 / (b3 * b3 + eps);
    const double a0 = (1.0 / 35) * o0, a1 = (12.0 / 35) * o1, a2 = (18.0 / 35) * o2, a3 = (4.0 / 35) * o3;
    const double one_alpha = 1.0 / (a0 + a1 + a2 + a3);
    const double w0 = a0 * one_alpha, w1 = a1 * one_alpha, w2 = a2 * one_alpha, w3 = a3 * one_alpha;
    const double p0 = (-1.0 / 4) * a + (13.0 / 12) * b + (-23.0 / 12) * c + (25.0 / 12) * d;
    const double p1 = (1.0 / 12) * b + (-5.0 / 12) * c + (13.0 / 12) * d + (1.0 / 4) * e;
    const double p2 = (-1.0 / 12) * c + (7.0 / 12) * d + (7.0 / 12) * e + (-1.0 / 12) * f;
    const double p3 = (1.0 / 4) * d + (13.0 / 12) * e + (-5.0 / 12) * f + (1.0 / 12) * g;
    return w0 * p0 + w1 * p1 + w2 * p2 + w3 * p3;
}

// Left and right states of one face from its window (cells -3..2; WENO7-JS: -4..3).
template <int STENCIL>
__device__ __forceinline__ void recon_pair(const double v[Win<STENCIL>::n], const Params& prm, double dx2, double& L,
                                           double& R) {
    if constexpr (STENCIL == kStencilW7) {  // weno7_js.py:31: cells -4..2 and 3..-3
        L = weno7_js(v[0], v[1], v[2], v[3], v[4], v[5], v[6], prm.eps);
        R = weno7_js(v[7], v[6], v[5], v[4], v[3], v[2], v[1], prm.eps);
    } else if constexpr (STENCIL != kStencilExt) {
        weno5_pair<STENCIL>(v, prm.eps_z, prm.eps_js, L, R);
    } else {
        switch (prm.stencil) {
            case JFA_WENO1:
                L = v[2];
                R = v[3];
                break;
            case JFA_WENO3_JS:  // weno3_base.py:29: cells -2..0 and 1..-1
                L = weno3<false>(v[1], v[2], v[3], prm.eps);
                R = weno3<false>(v[4], v[3], v[2], prm.eps);
                break;
            case JFA_WENO3_Z:
                L = weno3<true>(v[1], v[2], v[3], prm.eps);
                R = weno3<true>(v[4], v[3], v[2], prm.eps);
                break;
            case JFA_TENO5:
                L = teno5<false>(v[0], v[1], v[2], v[3], v[4], prm);
                R = teno5<false>(v[5], v[4], v[3], v[2], v[1], prm);
                break;
            case JFA_WENO6_CU:  // weno6_base.py:27: cells -3..2 and 2..-3
                L = weno6_cu(v[0], v[1], v[2], v[3], v[4], v[5], prm.eps);
                R = weno6_cu(v[5], v[4], v[3], v[2], v[1], v[0], prm.eps);
                break;
            case JFA_TENO6:
                L = teno6(v[0], v[1], v[2], v[3], v[4], v[5], prm.eps);
                R = teno6(v[5], v[4], v[3], v[2], v[1], v[0], prm.eps);
                break;
            case JFA_TENO6_A:
                L = teno6_a(v[0], v[1], v[2], v[3], v[4], v[5], prm.eps);
                R = teno6_a(v[5], v[4], v[3], v[2], v[1], v[0], prm.eps);
                break;
            case JFA_WENO6_CUM1:
                L = weno6_cum<false>(v[0], v[1], v[2], v[3], v[4], v[5], prm.eps, dx2);
                R = weno6_cum<false>(v[5], v[4], v[3], v[2], v[1], v[0], prm.eps, dx2);
                break;
            case JFA_WENO6_CUM2:
                L = weno6_cum<true>(v[0], v[1], v[2], v[3], v[4], v[5], prm.eps, dx2);
                R = weno6_cum<true>(v[5], v[4], v[3], v[2], v[1], v[0], prm.eps, dx2);
                break;
            default:  // JFA_TENO5_A
                L = teno5<true>(v[0], v[1], v[2], v[3], v[4], prm);
                R = teno5<true>(v[5], v[4], v[3], v[2], v[1], prm);
                break;
        }
    }
}

// One side from the 6-point stencil (u_imm, u_im, u_i, u_ip, u_ipp, u_ippp) of WENO6Base's slices
// (the 5-point stencils read the first five).
template <int STENCIL>
__device__ __forceinline__ double recon_side(double a, double b, double c, double d, double e, double f,
                                             const Params& prm, double dx2) {
    if constexpr (STENCIL != kStencilExt) {
        return weno5<STENCIL>(a, b, c, d, e, prm.eps_z, prm.eps_js);
    } else {
        switch (prm.stencil) {
            case JFA_WENO1: return c;
            case JFA_WENO3_JS: return weno3<false>(b, c, d, prm.eps);
            case JFA_WENO3_Z: return weno3<true>(b, c, d, prm.eps);
            case JFA_TENO5: return teno5<false>(a, b, c, d, e, prm);
            case JFA_WENO6_CU: return weno6_cu(a, b, c, d, e, f, prm.eps);
            case JFA_TENO6: return teno6(a, b, c, d, e, f, prm.eps);
            case JFA_TENO6_A: return teno6_a(a, b, c, d, e, f, prm.eps);
            case JFA_WENO6_CUM1: return weno6_cum<false>(a, b, c, d, e, f, prm.eps, dx2);
            case JFA_WENO6_CUM2: return weno6_cum<true>(a, b, c, d, e, f, prm.eps, dx2);
            default: return teno5<true>(a, b, c, d, e, prm);
        }
    }
}

// ---------------------------------------------------------------- ideal gas
// materials/single_materials/ideal_gas.py:69-88, equation_manager.py:93-101,164-171
// c = sqrt(gamma p / rho) = gamma p / sqrt(gamma p rho)
__device__ __forceinline__ double sound_speed(double p, double rho, double gamma) {
    const double gp = gamma * p;
    return gp * rsq1(gp * rho);
}

// E = rho (|u|^2/2 + p/(rho (gamma-1))) = rho |u|^2 / 2 + p / (gamma - 1)
__device__ __forceinline__ void cons_from_prim(const double W[5], const Params& prm, double U[5]) {
    const double rho = W[0];
    U[0] = rho;
    U[1] = rho * W[1];
    U[2] = rho * W[2];
    U[3] = rho * W[3];
    U[4] = (0.5 * rho) * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]) + W[4] * prm.inv_gm1;
}

__device__ __forceinline__ void prim_from_cons(const double U[5], const Params& prm, double W[5]) {
    const double rho = U[0];
    const double one_rho = rcp(rho);
    const double u = U[1] * one_rho, v = U[2] * one_rho, w = U[3] * one_rho;
    const double e = U[4] * one_rho - 0.5 * (u * u + v * v + w * w);
    W[0] = rho;
    W[1] = u;
    W[2] = v;
    W[3] = w;
    W[4] = prm.gm1 * e * rho;
}

// equation_manager.py:237-252
template <int AXIS>
__device__ __forceinline__ void physical_flux(const double W[5], const double U[5], double F[5]) {
    constexpr int v = Ax<AXIS>::v;
    const double m = U[v];
    F[0] = m;
    F[1] = m * W[1];
    F[2] = m * W[2];
    F[3] = m * W[3];
    F[v] += W[4];
    F[4] = W[v] * (U[4] + W[4]);
}

// ---------------------------------------------------------------- frozen state
// eigendecomposition.py:139-281 (single phase).
struct Frozen {
    double P[5];            // primes_ave
    double H, g, c, cc, q2; // enthalpy, grueneisen, c, c^2, velocity_square
    double inv_c, rcc;      // 1/c and c^2 * rho_ave
};

__device__ __forceinline__ double total_enthalpy(const double W[5], const Params& prm) {
    // ideal_gas.py:90-110: (p/(gamma-1) + rho|u|^2/2 + p) / rho
    const double E = W[4] * prm.inv_gm1 + 0.5 * W[0] * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
    return fdiv(E + W[4], W[0]);
}

template <bool FULL>
__device__ __forceinline__ void frozen_state(const double WL[5], const double WR[5], const Params& prm,
                                             Frozen& f) {
    if (prm.frozen == JFA_FROZEN_ARITHMETIC) {
#pragma unroll
        for (int i = 0; i < 5; ++i) f.P[i] = 0.5 * (WL[i] + WR[i]);  // :148
        // c = sqrt(gamma p / rho) (:223-225); c^2 rho = gamma p; 1/c = rho / sqrt(gamma p rho)
        const double gp = prm.gamma * f.P[4];
        const double r = rsq1(gp * f.P[0]);
        f.c = gp * r;
        f.inv_c = f.P[0] * r;
        f.cc = f.c * f.c;       // :230
        f.rcc = gp;
        if (FULL) {
            f.g = prm.gm1;                                             // :216-218
            f.H = total_enthalpy(f.P, prm);                            // :219-222
            f.q2 = f.P[1] * f.P[1] + f.P[2] * f.P[2] + f.P[3] * f.P[3];  // :231
        }
    } else {  // ROE :233-275
        const double sL = fsqrt(WL[0]), sR = fsqrt(WR[0]);
        const double rho_div = rcp1(sL + sR);
#pragma unroll
        for (int i = 1; i < 5; ++i) f.P[i] = (sL * WL[i] + sR * WR[i]) * rho_div;  // :293
        f.P[0] = fsqrt(WL[0] * WR[0]);                                             // :237
        const double pr_L = fdiv(WL[4], WL[0]), pr_R = fdiv(WR[4], WR[0]);         // psi = p/rho
        const double psi_ave = (sL * pr_L + sR * pr_R) * rho_div;                  // :258
        const double g = prm.gm1;
        f.g = (sL * g + sR * g) * rho_div;                                         // :260-262
        const double du = WR[1] - WL[1], dv = WR[2] - WL[2], dw = WR[3] - WL[3];
        const double sq_dvel = du * du + dv * dv + dw * dw;                        // :265
        const double p_over_rho = psi_ave + 0.5 * f.P[0] * rho_div * rho_div * sq_dvel;  // :267-268
        f.cc = psi_ave + f.g * p_over_rho;                                         // :274
        f.inv_c = rsq1(f.cc);
        f.c = f.cc * f.inv_c;                                                      // :275
        f.rcc = f.cc * f.P[0];
        if (FULL) {
            f.H = (sL * total_enthalpy(WL, prm) + sR * total_enthalpy(WR, prm)) * rho_div;  // :250
            f.q2 = f.P[1] * f.P[1] + f.P[2] * f.P[2] + f.P[3] * f.P[3];                    // :271
        }
    }
}

// Right and left eigenvectors of the conservative flux Jacobian at the frozen state f (single
// phase; eigendecomposition_conservatives, eigendecomposition.py:595-658), indexed [row][col].
template <int AXIS>
__device__ __forceinline__ void conservative_eigenvectors(const Frozen& f, double L[5][5], double R[5][5]) {
    constexpr int v = Ax<AXIS>::v, m0 = Ax<AXIS>::m0, m1 = Ax<AXIS>::m1;
    const double one_cc = f.inv_c * f.inv_c;
    const double one_rho = rcp1(f.P[0]);
    const double* P = f.P;
    const double g = f.g, H = f.H, c = f.c, q2 = f.q2;
#pragma unroll
    for (int a = 0; a < 5; ++a)
#pragma unroll
        for (int b = 0; b < 5; ++b) {
            L[a][b] = 0.0;
            R[a][b] = 0.0;
        }
    R[0][0] = 1.0;
    R[v][0] = P[v] - c;
    R[m0][0] = P[m0];
    R[m1][0] = P[m1];
    R[4][0] = H - P[v] * c;
    R[0][v] = g;
    R[1][v] = g * P[1];
    R[2][v] = g * P[2];
    R[3][v] = g * P[3];
    R[4][v] = g * H - f.cc;
    R[m0][m0] = -P[0];
    R[4][m0] = -P[0] * P[m0];
    R[m1][m1] = P[0];
    R[4][m1] = P[0] * P[m1];
    R[0][4] = 1.0;
    R[v][4] = P[v] + c;
    R[m0][4] = P[m0];
    R[m1][4] = P[m1];
    R[4][4] = H + P[v] * c;
    L[0][0] = 0.5 * one_cc * (g * q2 - g * H + (P[v] + c) * c);
    L[0][v] = 0.5 * one_cc * (-P[v] * g - c);
    L[0][m0] = 0.5 * one_cc * (-P[m0] * g);
    L[0][m1] = 0.5 * one_cc * (-P[m1] * g);
    L[0][4] = 0.5 * one_cc * g;
    L[v][0] = one_cc * (H - q2);
    L[v][1] = P[1] * one_cc;
    L[v][2] = P[2] * one_cc;
    L[v][3] = P[3] * one_cc;
    L[v][4] = -one_cc;
    L[m0][0] = P[m0] * one_rho;
    L[m0][m0] = -one_rho;
    L[m1][0] = -P[m1] * one_rho;
    L[m1][m1] = one_rho;
    L[4][0] = 0.5 * one_cc * (g * q2 - g * H - (P[v] - c) * c);
    L[4][v] = 0.5 * one_cc * (-P[v] * g + c);
    L[4][m0] = 0.5 * one_cc * (-P[m0] * g);
    L[4][m1] = 0.5 * one_cc * (-P[m1] * g);
    L[4][4] = 0.5 * one_cc * g;
}

// ---------------------------------------------------------------- signal speeds
// solvers/riemann_solvers/signal_speeds.py
__device__ __forceinline__ void signal_speeds(int kind, double uL, double uR, double aL, double aR, double rL,
                                              double rR, double pL, double pR, double gamma, double& SL,
                                              double& SR) {
    switch (kind) {
        case JFA_SS_ARITHMETIC: {  // :25-28
            const double um = 0.5 * (uL + uR), am = 0.5 * (aL + aR);
            SL = fmin(um - am, uL - aL);
            SR = fmax(um + am, uR + aR);
        } break;
        case JFA_SS_RUSANOV: {  // :46-48
            const double sp = fmax(fabs(uL) + aL, fabs(uR) + aR);
            SL = -sp;
            SR = sp;
        } break;
        case JFA_SS_DAVIS:  // :67-68
            SL = fmin(uL - aL, uR - aR);
            SR = fmax(uL + aL, uR + aR);
            break;
        case JFA_SS_TORO: {  // :151-156, estimate_pressure :227-230
            const double rho_bar = 0.5 * (rL + rR), a_bar = 0.5 * (aL + aR);
            const double p_pvrs = 0.5 * (pL + pR) - 0.5 * (uR - uL) * rho_bar * a_bar;
            const double ps = fmax(0.0, p_pvrs);
            const double g_ = (gamma + 1.0) * 0.5 / gamma;
            const double qL = ps <= pL ? 1.0 : fsqrt(1.0 + g_ * (fdiv(ps, pL) - 1.0));
            const double qR = ps <= pR ? 1.0 : fsqrt(1.0 + g_ * (fdiv(ps, pR) - 1.0));
            SL = uL - aL * qL;
            SR = uR + aR * qR;
        } break;
        default: {  // EINFELDT :125-132
            const double yL = rsq1(rL), yR = rsq1(rR);
            const double sL = rL * yL, sR = rR * yR;  // sqrt(rho)
            const double one_dens = rcp1(sL + sR);
            const double eta2 = 0.5 * sL * sR * one_dens * one_dens;
            const double u_bar = (sL * uL + sR * uR) * one_dens;
            const double du = uR - uL;
            const double d_bar = fsqrt((sL * aL * aL + sR * aR * aR) * one_dens + eta2 * (du * du));
            SL = fmin(u_bar - d_bar, uL - aL);
            SR = fmax(u_bar + d_bar, uR + aR);
        } break;
    }
}

// signal_speed_Davis_2 (signal_speeds.py:99-107): u_Roe -/+ a_Roe with the sqrt(rho)-weighted
// averages of u and of the total enthalpy H.  Only the plain-function entry point calls it: the
// reference's Riemann solvers pass p_L/p_R, not H_L/H_R, so it cannot run inside them.
__device__ __forceinline__ void signal_speeds_davis2(double uL, double uR, double rL, double rR, double HL,
                                                     double HR, double gamma, double& SL, double& SR) {
    const double sL = rL * rsq1(rL), sR = rR * rsq1(rR);  // sqrt(rho)
    const double one_dens = rcp1(sL + sR);
    const double u_roe = (sL * uL + sR * uR) * one_dens;
    const double H_roe = (sL * HL + sR * HR) * one_dens;
    const double a_roe = fsqrt((gamma - 1.0) * (H_roe - 0.5 * u_roe * u_roe));
    SL = u_roe - a_roe;
    SR = u_roe + a_roe;
}

// ---------------------------------------------------------------- Riemann solvers
// HLLC._compute_flux_star_K_single_phase (HLLC.py:41-78) for one side K with wave speed SK,
// its clipped value S_clip (min(.,0) for L, max(.,0) for R) and contact speed Ss.
// With d = SK - u and inv = 1/(SK - Ss):
//   pre = rho d inv                                  (:52)
//   U*_4 = pre (E/rho + (Ss - u)(Ss + p/(rho d)))   (:59)  = inv (d E + (Ss - u)(rho d Ss + p))
template <int AXIS>
__device__ __forceinline__ void hllc_star_flux(const double W[5], const double U[5], double SK, double S_clip,
                                               double Ss, double F[5]) {
    constexpr int v = Ax<AXIS>::v, m0 = Ax<AXIS>::m0, m1 = Ax<AXIS>::m1;
    const double uK = W[v];
    const double d = SK - uK;
    const double inv = rcp1(SK - Ss);
    const double rd = W[0] * d;
    const double pre = rd * inv;
    double us[5];
    us[0] = pre;
    us[v] = pre * Ss;                                                   // :61
    us[m0] = pre * W[m0];                                               // :62
    us[m1] = pre * W[m1];                                               // :63
    us[4] = inv * (d * U[4] + (Ss - uK) * (rd * Ss + W[4]));
    physical_flux<AXIS>(W, U, F);                                       // :66
#pragma unroll
    for (int i = 0; i < 5; ++i) F[i] = F[i] + S_clip * (us[i] - U[i]);  // :76
}

// HLLC._solve_riemann_problem_xi_single_phase (HLLC.py:80-126).  The blend
// 0.5(1+sgn S*) F*_L + 0.5(1-sgn S*) F*_R equals F*_L for S*>0 and F*_R for S*<0 exactly,
// so only the needed side is evaluated; S* == 0 averages both (:123-124).
template <int AXIS>
__device__ __forceinline__ void hllc(const double WL[5], const double WR[5], const double UL[5],
                                     const double UR[5], const Params& prm, double F[5]) {
    constexpr int v = Ax<AXIS>::v;
    const double cL = sound_speed(WL[4], WL[0], prm.gamma);
    const double cR = sound_speed(WR[4], WR[0], prm.gamma);
    double SL, SR;
    signal_speeds(prm.signal_speed, WL[v], WR[v], cL, cR, WL[0], WR[0], WL[4], WR[4], prm.gamma, SL, SR);
    // compute_sstar (signal_speeds.py:194-197)
    const double rdL = WL[0] * (SL - WL[v]);
    const double rdR = WR[0] * (SR - WR[v]);
    const double Ss = fdiv((WR[4] - WL[4]) + (WL[v] * rdL - WR[v] * rdR), rdL - rdR);
    const bool left = Ss >= 0.0;
    double Wk[5], Uk[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        Wk[i] = left ? WL[i] : WR[i];
        Uk[i] = left ? UL[i] : UR[i];
    }
    const double Sk = left ? SL : SR;
    const double Sclip = left ? fmin(SL, 0.0) : fmax(SR, 0.0);
    hllc_star_flux<AXIS>(Wk, Uk, Sk, Sclip, Ss, F);
    if (Ss == 0.0) {
        double FR[5];
        hllc_star_flux<AXIS>(WR, UR, SR, fmax(SR, 0.0), Ss, FR);
#pragma unroll
        for (int i = 0; i < 5; ++i) F[i] = 0.5 * F[i] + 0.5 * FR[i];
    }
}

// Rusanov._solve_riemann_problem_xi_single_phase (Rusanov.py:35-45)
template <int AXIS>
__device__ __forceinline__ void rusanov(const double WL[5], const double WR[5], const double UL[5],
                                        const double UR[5], const Params& prm, double F[5]) {
    constexpr int v = Ax<AXIS>::v;
    const double cL = sound_speed(WL[4], WL[0], prm.gamma);
    const double cR = sound_speed(WR[4], WR[0], prm.gamma);
    const double alpha = fmax(fabs(WL[v]) + cL, fabs(WR[v]) + cR);
    double FL[5], FR[5];
    physical_flux<AXIS>(WL, UL, FL);
    physical_flux<AXIS>(WR, UR, FR);
#pragma unroll
    for (int i = 0; i < 5; ++i) F[i] = 0.5 * (FL[i] + FR[i]) - 0.5 * alpha * (UR[i] - UL[i]);
}

// HLL._solve_riemann_problem_xi_single_phase (HLL.py:29-61): with S_L^- = min(S_L, 0) and
// S_R^+ = max(S_R, 0) of the configured signal speeds,
//   F = (S_R^+ F_L - S_L^- F_R + S_L^- S_R^+ (U_R - U_L)) / (S_R^+ - S_L^- + eps),
// eps the working precision's machine epsilon (riemann_solver.py:28).
template <int AXIS>
__device__ __forceinline__ void hll(const double WL[5], const double WR[5], const double UL[5],
                                    const double UR[5], const Params& prm, double F[5]) {
    constexpr int v = Ax<AXIS>::v;
    const double cL = sound_speed(WL[4], WL[0], prm.gamma);
    const double cR = sound_speed(WR[4], WR[0], prm.gamma);
    double SL, SR;
    signal_speeds(prm.signal_speed, WL[v], WR[v], cL, cR, WL[0], WR[0], WL[4], WR[4], prm.gamma, SL, SR);
    const double sl = fmin(SL, 0.0), sr = fmax(SR, 0.0);
    const double r = rcp(sr - sl + prm.eps_precision);
    const double slsr = sl * sr;
    double FL[5], FR[5];
    physical_flux<AXIS>(WL, UL, FL);
    physical_flux<AXIS>(WR, UR, FR);
#pragma unroll
    for (int i = 0; i < 5; ++i) F[i] = fma(slsr, UR[i] - UL[i], fma(sr, FL[i], -sl * FR[i])) * r;
}

// HLLC star state of side K (HLLCLM.py:95-108, the same as HLLC.py:52-63).
template <int AXIS>
__device__ __forceinline__ void hllc_star_state(const double W[5], const double U[5], double SK, double Ss,
                                                double us[5]) {
    constexpr int v = Ax<AXIS>::v, m0 = Ax<AXIS>::m0, m1 = Ax<AXIS>::m1;
    const double d = SK - W[v];
    const double inv = rcp1(SK - Ss);
    const double rd = W[0] * d;
    const double pre = rd * inv;
    us[0] = pre;
    us[v] = pre * Ss;
    us[m0] = pre * W[m0];
    us[m1] = pre * W[m1];
    us[4] = inv * (d * U[4] + (Ss - W[v]) * (rd * Ss + W[4]));
}

__device__ __forceinline__ double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// HLLC-LM (HLLCLM.py:59-136; Fleischmann et al. 2020): the HLLC star states, outer wave speeds
// scaled by phi = sin(min(1, Ma / 0.1) pi / 2), Ma = max(|u_L / c_L|, |u_R / c_R|) (:37-57),
// the central star flux
//   F* = (F_L + F_R)/2 + (phi S_L (u*_L - U_L) + |S*| (u*_L - u*_R) + phi S_R (u*_R - U_R))/2
// and the blend (1 + sgn S_L)/2 F_L + (1 - sgn S_R)/2 F_R + (1 - sgn S_L)(1 + sgn S_R)/4 F*.
template <int AXIS>
__device__ __forceinline__ void hllc_lm(const double WL[5], const double WR[5], const double UL[5],
                                        const double UR[5], const Params& prm, double F[5]) {
    constexpr int v = Ax<AXIS>::v;
    const double cL = sound_speed(WL[4], WL[0], prm.gamma);
    const double cR = sound_speed(WR[4], WR[0], prm.gamma);
    double SL, SR;
    signal_speeds(prm.signal_speed, WL[v], WR[v], cL, cR, WL[0], WR[0], WL[4], WR[4], prm.gamma, SL, SR);
    const double rdL = WL[0] * (SL - WL[v]);
    const double rdR = WR[0] * (SR - WR[v]);
    const double Ss = fdiv((WR[4] - WL[4]) + (WL[v] * rdL - WR[v] * rdR), rdL - rdR);
    double usL[5], usR[5];
    hllc_star_state<AXIS>(WL, UL, SL, Ss, usL);
    hllc_star_state<AXIS>(WR, UR, SR, Ss, usR);
    const double Ma = fmax(fabs(fdiv(WL[v], cL)), fabs(fdiv(WR[v], cR)));
    const double phi = sin(fmin(1.0, Ma / 0.1) * (0.5 * 3.141592653589793));
    const double wL = phi * SL, wR = phi * SR, aS = fabs(Ss);
    const double bL = 0.5 * (1.0 + sgn(SL)), bR = 0.5 * (1.0 - sgn(SR));
    const double bS = 0.25 * (1.0 - sgn(SL)) * (1.0 + sgn(SR));
    double FL[5], FR[5];
    physical_flux<AXIS>(WL, UL, FL);
    physical_flux<AXIS>(WR, UR, FR);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const double Fs = 0.5 * (FL[i] + FR[i]) +
                          0.5 * (wL * (usL[i] - UL[i]) + aS * (usL[i] - usR[i]) + wR * (usR[i] - UR[i]));
        F[i] = bL * FL[i] + bR * FR[i] + bS * Fs;
    }
}

// AUSM+ (AUSMP.py:30-103; Liou 1996) with the arithmetic interface speed of sound
// a = (c_L + c_R)/2 (:26,54-55), alpha = 3/16, beta = 1/8 (:27-28): split Mach numbers
// M+(M_L), M-(M_R) and pressures P+(M_L), P-(M_R) (:65-75), and
//   F = a (max(M, 0) phi_L + min(M, 0) phi_R) + (P+ p_L + P- p_R) e_n,  M = M+ + M-,
// phi = (rho, rho u, rho v, rho w, E + p) (:85-103).
template <int AXIS>
__device__ __forceinline__ void ausmp(const double WL[5], const double WR[5], const double UL[5],
                                      const double UR[5], const Params& prm, double F[5]) {
    constexpr int v = Ax<AXIS>::v;
    constexpr double alpha = 3.0 / 16.0, beta = 1.0 / 8.0;
    const double a = 0.5 * (sound_speed(WL[4], WL[0], prm.gamma) + sound_speed(WR[4], WR[0], prm.gamma));
    const double Ml = fdiv(WL[v], a), Mr = fdiv(WR[v], a);
    const double ql = Ml * Ml - 1.0, qr = Mr * Mr - 1.0;
    const double lp = Ml + 1.0, rm = Mr - 1.0;
    const bool sl = fabs(Ml) >= 1.0, sr = fabs(Mr) >= 1.0;
    const double Mp = sl ? 0.5 * (Ml + fabs(Ml)) : 0.25 * (lp * lp) + beta * (ql * ql);
    const double Mm = sr ? 0.5 * (Mr - fabs(Mr)) : -0.25 * (rm * rm) - beta * (qr * qr);
    const double M = Mp + Mm;
    const double Map = 0.5 * (M + fabs(M)), Mam = 0.5 * (M - fabs(M));
    const double Pp = sl ? 0.5 * (1.0 + sgn(Ml)) : 0.25 * (lp * lp) * (2.0 - Ml) + alpha * Ml * (ql * ql);
    const double Pm = sr ? 0.5 * (1.0 - sgn(Mr)) : 0.25 * (rm * rm) * (2.0 + Mr) - alpha * Mr * (qr * qr);
    const double pa = Pp * WL[4] + Pm * WR[4];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const double phl = i == 4 ? UL[4] + WL[4] : UL[i];
        const double phr = i == 4 ? UR[4] + WR[4] : UR[i];
        F[i] = a * (Map * phl + Mam * phr);
    }
    F[v] += pa;
}

// The further Riemann solvers, selected at run time (prm.riemann) in one instantiation of the
// face kernels, so the HLLC and Rusanov kernels keep their code.
constexpr int kRiemannExt = 100;

template <int AXIS, int RIEMANN>
__device__ __forceinline__ void riemann(const double WL[5], const double WR[5], const double UL[5],
                                        const double UR[5], const Params& prm, double F[5]) {
    if constexpr (RIEMANN == JFA_HLLC) {
        hllc<AXIS>(WL, WR, UL, UR, prm, F);
    } else if constexpr (RIEMANN == kRiemannExt) {
        if (prm.riemann == JFA_HLL) hll<AXIS>(WL, WR, UL, UR, prm, F);
        else if (prm.riemann == JFA_HLLC_LM) hllc_lm<AXIS>(WL, WR, UL, UR, prm, F);
        else ausmp<AXIS>(WL, WR, UL, UR, prm, F);
    } else {
        rusanov<AXIS>(WL, WR, UL, UR, prm, F);
    }
}

// Positivity flux limiter of one face (limiter_flux.py:147-316,452-640,733-774; SIMPLE / NASA,
// single phase).  wl, wr, ul, ur: primitives and conservatives of the face's left and right
// cells; lam = dt (1/dx) sigma.  The pseudo-updated states are U_minus = ur + 2 lam (F - Fs_r)
// and U_plus = ul - 2 lam (F - Fs_l), Fs the cells' physical flux (NASA) or 0 (SIMPLE); where the
// smaller density is below eps the flux becomes the first-order one (HLLC with Einfeldt speeds
// on the cells' primitives, the reference's WENO1 Godunov flux, :57-98), then, with the updated
// flux, likewise where the smaller pressure is.  Returns the number of switches.
// Pressure of prim(U) in the reference's operation order (equation_manager.py:164-171,
// ideal_gas.py:73-75): 1/rho by IEEE division, e = E/rho - 0.5 ((u^2 + v^2) + w^2), p = ((gamma-1) e) rho.
__device__ __forceinline__ double prim_pressure(const double U[5], const Params& prm) {
#pragma clang fp contract(off)
    const double one_rho = 1.0 / U[0];
    const double u = U[1] * one_rho, v = U[2] * one_rho, w = U[3] * one_rho;
    const double e = U[4] * one_rho - 0.5 * ((u * u + v * v) + w * w);
    return prm.gm1 * e * U[0];
}

template <int AXIS>
__device__ __forceinline__ int limit_flux_has(double F[5], const double wl[5], const double wr[5], const double ul[5],
                                              const double ur[5], double lam, const Params& prm) {
    // Hu, Adams & Shu (limiter_flux.py:200-251): theta = (eps - q_LF) / (q - q_LF) on a side whose
    // density (then pressure) q is below eps, q_LF the same side's value with the first-order
    // flux; F <- (1 - theta) F_pos + theta F with theta the smaller of the two sides'.  IEEE
    // divisions, as the reference's jnp.where evaluates them (a zero denominator gives +-inf/NaN).
    // theta divides by the difference of two nearly equal pseudo-updated values wherever the
    // first-order flux is close to the high-order one, so this follows the reference's operation
    // order: no contraction, IEEE divisions (prim_pressure).
#pragma clang fp contract(off)
    Params pe = prm;
    pe.signal_speed = JFA_SS_EINFELDT;
    double cl[5], cr[5], Fp[5];
    cons_from_prim(wl, pe, cl);
    cons_from_prim(wr, pe, cr);
    riemann<AXIS, JFA_HLLC>(wl, wr, cl, cr, pe, Fp);
    const double l2 = 2.0 * lam;
    double um[5], up[5], umf[5], upf[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        umf[i] = ur[i] + l2 * Fp[i];
        upf[i] = ul[i] - l2 * Fp[i];
    }
    // density (:220-234)
    const double rm = ur[0] + l2 * F[0], rp = ul[0] - l2 * F[0];
    const double er = prm.flim_eps_rho, ep = prm.flim_eps_p;
    const double tm = rm < er ? (er - umf[0]) / (rm - umf[0]) : 1.0;
    const double tp = rp < er ? (er - upf[0]) / (rp - upf[0]) : 1.0;
    const double t_rho = (tm < tp || tm != tm) ? tm : tp;  // jnp.minimum: a NaN operand gives NaN
#pragma unroll
    for (int i = 0; i < 5; ++i) F[i] = (1.0 - t_rho) * Fp[i] + t_rho * F[i];
    // pressure of the updated flux's states and of the first-order states (:237-251)
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        um[i] = ur[i] + l2 * F[i];
        up[i] = ul[i] - l2 * F[i];
    }
    const double pm = prim_pressure(um, prm), pp = prim_pressure(up, prm);
    const double pmf = prim_pressure(umf, prm), ppf = prim_pressure(upf, prm);
    const double sm = pm < ep ? (ep - pmf) / (pm - pmf) : 1.0;
    const double sp = pp < ep ? (ep - ppf) / (pp - ppf) : 1.0;
    const double t_p = (sm < sp || sm != sm) ? sm : sp;
#pragma unroll
    for (int i = 0; i < 5; ++i) F[i] = (1.0 - t_p) * Fp[i] + t_p * F[i];
    return (t_rho < 1.0 || t_p < 1.0) ? 1 : 0;
}

template <int AXIS>
__device__ __forceinline__ int limit_flux(double F[5], const double wl[5], const double wr[5], const double ul[5],
                                          const double ur[5], double lam, const Params& prm) {
    if (prm.flux_limiter == JFA_FLUX_LIMITER_HAS) return limit_flux_has<AXIS>(F, wl, wr, ul, ur, lam, prm);
    double fsl[5], fsr[5];
    if (prm.flux_limiter == JFA_FLUX_LIMITER_NASA) {
        physical_flux<AXIS>(wl, ul, fsl);
        physical_flux<AXIS>(wr, ur, fsr);
    } else {
#pragma unroll
        for (int i = 0; i < 5; ++i) fsl[i] = fsr[i] = 0.0;
    }
    const double l2 = 2.0 * lam;
    int n = 0;
    bool pos = false;
    double Fp[5];
    auto first_order = [&]() {
        Params pe = prm;
        pe.signal_speed = JFA_SS_EINFELDT;
        double cl[5], cr[5];
        cons_from_prim(wl, pe, cl);
        cons_from_prim(wr, pe, cr);
        riemann<AXIS, JFA_HLLC>(wl, wr, cl, cr, pe, Fp);
        pos = true;
    };
    // density of the pseudo-updated states (:570-575)
    const double rm = ur[0] + l2 * (F[0] - fsr[0]);
    const double rp = ul[0] - l2 * (F[0] - fsl[0]);
    if (fmin(rm, rp) < prm.flim_eps_rho) {
        first_order();
#pragma unroll
        for (int i = 0; i < 5; ++i) F[i] = Fp[i];
        ++n;
    }
    // pressure, with the flux after the density check (:619-627)
    double um[5], up[5], wm[5], wp[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        um[i] = ur[i] + l2 * (F[i] - fsr[i]);
        up[i] = ul[i] - l2 * (F[i] - fsl[i]);
    }
    prim_from_cons(um, prm, wm);
    prim_from_cons(up, prm, wp);
    if (fmin(wm[4], wp[4]) < prm.flim_eps_p) {
        if (!pos) first_order();
#pragma unroll
        for (int i = 0; i < 5; ++i) F[i] = Fp[i];
        ++n;
    }
    return n;
}

// ---------------------------------------------------------------- reconstruction
// HighOrderGodunov.reconstruct_xi (high_order_godunov.py:233-419) for one face.
// win[k][i]: variable i of cell k-3 relative to the cell right of the face (k = 0..5),
// i.e. the 6-cell stencil window of eigendecomposition.py:75-95.  For VAR PRIMITIVE and
// CHAR-PRIMITIVE `win` holds primitives; for CONSERVATIVE and CHAR-CONSERVATIVE it holds
// conservatives and `wl`, `wr` hold the primitives of the two cells adjacent to the face.
// `win` is read only: the marching sweeps keep it in registers across faces.
// Positivity limiter of one reconstructed face state (limiter_interpolation.py:86-208, single
// phase): density below eps takes the first-order (upwind cell) value q1 in the limited fields
// (density and pressure; all five with limit_velocity), then, on the updated state, pressure
// below eps likewise (pb = 0 for an ideal gas).  Returns the number of replacements.
__device__ __forceinline__ int limit_face_state(double Wf[5], const double q1[5], const Params& prm) {
    int n = 0;
#pragma unroll
    for (int check = 0; check < 2; ++check) {
        const bool low = check == 0 ? Wf[0] < prm.lim_eps_rho : Wf[4] < prm.lim_eps_p;
        if (low) {
            ++n;
            Wf[0] = q1[0];
            Wf[4] = q1[4];
            if (prm.limit_velocity) {
                Wf[1] = q1[1];
                Wf[2] = q1[2];
                Wf[3] = q1[3];
            }
        }
    }
    return n;
}

// Face states WL, WR, UL, UR of one face from its six-cell window (primitives for PRIMITIVE and
// CHAR-PRIMITIVE, conservatives otherwise); wl, wr are the primitives of the face's two cells,
// read for CHAR-CONSERVATIVE.  The positivity limiter is applied by the callers (limit_side).
template <int AXIS, int VAR, int STENCIL>
__device__ __forceinline__ void face_states(const double (&win)[Win<STENCIL>::n][5], const double wl[5],
                                            const double wr[5], const Params& prm, double WL[5], double WR[5],
                                            double UL[5], double UR[5]) {
    constexpr int v = Ax<AXIS>::v, m0 = Ax<AXIS>::m0, m1 = Ax<AXIS>::m1;
    constexpr int NW = Win<STENCIL>::n, H = Win<STENCIL>::h;
    if constexpr (VAR == JFA_PRIMITIVE || VAR == JFA_CONSERVATIVE) {
        double* QL = VAR == JFA_PRIMITIVE ? WL : UL;
        double* QR = VAR == JFA_PRIMITIVE ? WR : UR;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            // WENO5Base slices j=0: cells -3..1, j=1: cells 2..-2 (weno5_base.py:30)
            double cw[NW];
#pragma unroll
            for (int k = 0; k < NW; ++k) cw[k] = win[k][i];
            recon_pair<STENCIL>(cw, prm, prm.dx2[AXIS], QL[i], QR[i]);
        }
        if constexpr (VAR == JFA_PRIMITIVE) {
            cons_from_prim(WL, prm, UL);  // :279-280
            cons_from_prim(WR, prm, UR);
        } else {
            prim_from_cons(UL, prm, WL);  // :295-296
            prim_from_cons(UR, prm, WR);
        }
    } else if constexpr (VAR == JFA_CHAR_PRIMITIVE) {
        // frozen state from the cells left/right of the face (eigendecomposition.py:139-140)
        Frozen f;
        frozen_state<false>(win[H - 1], win[H], prm, f);
        // get_characteristics_from_primitives (eigendecomposition.py:425-431):
        // k1 = 0.5/c, k2 = 0.5/(c^2 rho), k3 = 1/c^2
        const double k1 = 0.5 * f.inv_c;
        const double k3 = f.inv_c * f.inv_c;
        const double k2 = 0.5 * rcp1(f.rcc);
        // one characteristic field at a time (x0 = k2 p - k1 u_n, x1 = rho - k3 p,
        // x2, x3 = tangential velocities, x4 = k2 p + k1 u_n), each reconstructed L and R
        double cLv[5], cRv[5], t[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) t[k] = k2 * win[k][4] - k1 * win[k][v];
        recon_pair<STENCIL>(t, prm, prm.dx2[AXIS], cLv[0], cRv[0]);
#pragma unroll
        for (int k = 0; k < NW; ++k) t[k] = win[k][0] - k3 * win[k][4];
        recon_pair<STENCIL>(t, prm, prm.dx2[AXIS], cLv[1], cRv[1]);
#pragma unroll
        for (int k = 0; k < NW; ++k) t[k] = win[k][m0];
        recon_pair<STENCIL>(t, prm, prm.dx2[AXIS], cLv[2], cRv[2]);
#pragma unroll
        for (int k = 0; k < NW; ++k) t[k] = win[k][m1];
        recon_pair<STENCIL>(t, prm, prm.dx2[AXIS], cLv[3], cRv[3]);
#pragma unroll
        for (int k = 0; k < NW; ++k) t[k] = k2 * win[k][4] + k1 * win[k][v];
        recon_pair<STENCIL>(t, prm, prm.dx2[AXIS], cLv[4], cRv[4]);
        // get_primitives_from_characteristics (eigendecomposition.py:517-521)
        const double sL = cLv[0] + cLv[4], sR = cRv[0] + cRv[4];
        WL[0] = f.P[0] * sL + cLv[1];
        WL[v] = f.c * (cLv[4] - cLv[0]);
        WL[m0] = cLv[2];
        WL[m1] = cLv[3];
        WL[4] = f.rcc * sL;
        WR[0] = f.P[0] * sR + cRv[1];
        WR[v] = f.c * (cRv[4] - cRv[0]);
        WR[m0] = cRv[2];
        WR[m1] = cRv[3];
        WR[4] = f.rcc * sR;
        cons_from_prim(WL, prm, UL);  // :401-402
        cons_from_prim(WR, prm, UR);
    } else {  // CHAR-CONSERVATIVE (:404-417)
        Frozen f;
        frozen_state<true>(wl, wr, prm, f);
        double L[5][5], R[5][5];
        conservative_eigenvectors<AXIS>(f, L, R);
        // transformtochar: einsum over j (eigendecomposition.py:729-730)
        double ch[NW][5];
#pragma unroll
        for (int k = 0; k < NW; ++k) {
#pragma unroll
            for (int a = 0; a < 5; ++a) {
                double acc = L[a][0] * win[k][0];
#pragma unroll
                for (int b = 1; b < 5; ++b) acc = acc + L[a][b] * win[k][b];
                ch[k][a] = acc;
            }
        }
        double cLv[5], cRv[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            double cw[NW];
#pragma unroll
            for (int k = 0; k < NW; ++k) cw[k] = ch[k][i];
            recon_pair<STENCIL>(cw, prm, prm.dx2[AXIS], cLv[i], cRv[i]);
        }
        // transformtophysical (eigendecomposition.py:742)
#pragma unroll
        for (int a = 0; a < 5; ++a) {
            double accL = R[a][0] * cLv[0], accR = R[a][0] * cRv[0];
#pragma unroll
            for (int b = 1; b < 5; ++b) {
                accL = accL + R[a][b] * cLv[b];
                accR = accR + R[a][b] * cRv[b];
            }
            UL[a] = accL;
            UR[a] = accR;
        }
        prim_from_cons(UL, prm, WL);  // :416-417
        prim_from_cons(UR, prm, WR);
    }
}

// The limiter on one face state (high_order_godunov.py:163-176), given the primitives q1 of the
// cell upwind of it on its side (the face's left cell for the left state), then U = cons(W)
// (limiter_interpolation.py:206; for primitive reconstruction U already is cons(W) where nothing
// was replaced).  Callers load q1 inside their `prm.limiter` branch, one side at a time, so no
// register holds cell values across the reconstruction.  Returns the number of replacements.
template <int VAR>
__device__ __forceinline__ int limit_side(const double q1[5], const Params& prm, double Wf[5], double Uf[5]) {
    constexpr bool cons_win = VAR == JFA_CONSERVATIVE || VAR == JFA_CHAR_CONSERVATIVE;
    const int n = limit_face_state(Wf, q1, prm);
    if (cons_win || n) cons_from_prim(Wf, prm, Uf);
    return n;
}

// ---------------------------------------------------------------- flux splitting
// FluxSplittingScheme.compute_flux_xi (solvers/convective_fluxes/flux_splitting_scheme.py:65-111),
// single phase, on one face.  win[k] holds the primitives of cell k-3 relative to the cell right of
// the face.  With the conservative eigenvectors L, R of the frozen state of the face's two cells
// and the eigenvalue magnitudes g_a of the splitting (eigendecomposition.py:672-696), each
// characteristic field a of the window splits into f+ = (L F + g L U)/2 and f- = (L F - g L U)/2
// (flux_splitting_scheme.py:91-92; F the cells' physical flux, U their conservatives); f+ is
// reconstructed from the left (j = 0), f- from the right (j = 1), and the face flux is R (f+ + f-).
// Instantiated as the kFluxSplit "Riemann solver" of the PRIMITIVE-window kernels; prm.riemann
// names the eigenvalues (JFA_FS_ROE / JFA_FS_CLLF / JFA_FS_LLF).
constexpr int kFluxSplit = 200;

template <int AXIS, int STENCIL>
__device__ __forceinline__ void flux_split_face(const double (&win)[6][5], const Params& prm, double F[5]) {
    constexpr int v = Ax<AXIS>::v;
    Frozen f;
    frozen_state<true>(win[2], win[3], prm, f);
    double L[5][5], R[5][5];
    conservative_eigenvectors<AXIS>(f, L, R);
    double g[5];
    if (prm.riemann == JFA_FS_ROE) {  // :674-677, at the frozen state
        g[0] = fabs(f.P[v] - f.c);
        g[1] = g[2] = g[3] = fabs(f.P[v]);
        g[4] = fabs(f.P[v] + f.c);
    } else {
        const double uL = win[2][v], uR = win[3][v];
        const double cL = sound_speed(win[2][4], win[2][0], prm.gamma);
        const double cR = sound_speed(win[3][4], win[3][0], prm.gamma);
        if (prm.riemann == JFA_FS_CLLF) {  // :680-688
            g[0] = fmax(fabs(uL - cL), fabs(uR - cR));
            g[1] = g[2] = g[3] = fmax(fabs(uL), fabs(uR));
            g[4] = fmax(fabs(uL + cL), fabs(uR + cR));
        } else {  // LLF :690-695
            g[0] = g[1] = g[2] = g[3] = g[4] = fmax(fabs(uL) + cL, fabs(uR) + cR);
        }
    }
    double Uk[6][5], Fk[6][5];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        cons_from_prim(win[k], prm, Uk[k]);
        physical_flux<AXIS>(win[k], Uk[k], Fk[k]);
    }
    double fc[5];
#pragma unroll
    for (int a = 0; a < 5; ++a) {
        // transformtochar of the window's conservatives and fluxes (eigendecomposition.py:717-730)
        double pos[6], neg[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double cu = L[a][0] * Uk[k][0], cf = L[a][0] * Fk[k][0];
#pragma unroll
            for (int b = 1; b < 5; ++b) {
                cu = cu + L[a][b] * Uk[k][b];
                cf = cf + L[a][b] * Fk[k][b];
            }
            const double gc = g[a] * cu;
            pos[k] = 0.5 * (cf + gc);
            neg[k] = 0.5 * (cf - gc);
        }
        fc[a] = recon_side<STENCIL>(pos[0], pos[1], pos[2], pos[3], pos[4], pos[5], prm, prm.dx2[AXIS]) +
                recon_side<STENCIL>(neg[5], neg[4], neg[3], neg[2], neg[1], neg[0], prm, prm.dx2[AXIS]);
    }
    // transformtophysical (eigendecomposition.py:732-742)
#pragma unroll
    for (int a = 0; a < 5; ++a) {
        double acc = R[a][0] * fc[0];
#pragma unroll
        for (int b = 1; b < 5; ++b) acc = acc + R[a][b] * fc[b];
        F[a] = acc;
    }
}

}  // namespace jfa
