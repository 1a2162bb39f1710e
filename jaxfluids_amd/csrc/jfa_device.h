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

struct Params {
    double gamma;
    double gm1;        // gamma - 1
    double inv_gm1;    // 1 / (gamma - 1)
    double eps;        // WENO epsilon (config/precision.py:39,53)
    int signal_speed;  // JFA_SS_*
    int frozen;        // JFA_FROZEN_*
};

// ---------------------------------------------------------------- refined hardware estimates
// Newton-Raphson for 1/x: r <- r + r(1 - x r), twice from the v_rcp_f64 estimate.
__device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
// a / b with a final residual correction q <- q + r(a - b q).
__device__ __forceinline__ double fdiv(double a, double b) {
    const double r = rcp(b);
    const double q = a * r;
    return fma(r, fma(-b, q, a), q);
}
// Newton-Raphson for 1/sqrt(x): y <- y + y(1/2 - (x/2) y^2), twice from v_rsq_f64.
__device__ __forceinline__ double rsq(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double t = fma(-h * y, y, 0.5);
    y = fma(y, t, y);
    t = fma(-h * y, y, 0.5);
    return fma(y, t, y);
}
// sqrt(x) = x rsq(x) with one residual correction g <- g + (y/2)(x - g^2).
__device__ __forceinline__ double fsqrt(double x) {
    const double y = rsq(x);
    const double g = x * y;
    return fma(0.5 * y, fma(-g, g, x), g);
}

// ---------------------------------------------------------------- WENO5
// weno5_base.py:34-50 smoothness / candidate polynomials; weno5_js.py:39-58; weno5_z.py:152-166.
// Arguments are (u_imm, u_im, u_i, u_ip, u_ipp) of one side.
template <int STENCIL>
__device__ __forceinline__ double weno5(double a, double b, double c, double d, double e, double eps) {
    const double t0 = a - 2.0 * b + c;
    const double t1 = a - 4.0 * b + 3.0 * c;
    const double t2 = b - 2.0 * c + d;
    const double t3 = b - d;
    const double t4 = c - 2.0 * d + e;
    const double t5 = 3.0 * c - 4.0 * d + e;
    const double beta0 = (13.0 / 12.0) * (t0 * t0) + 0.25 * (t1 * t1);
    const double beta1 = (13.0 / 12.0) * (t2 * t2) + 0.25 * (t3 * t3);
    const double beta2 = (13.0 / 12.0) * (t4 * t4) + 0.25 * (t5 * t5);
    // candidate values (weno5_base.py:47-49)
    const double p0 = (1.0 / 3.0) * a + (-7.0 / 6.0) * b + (11.0 / 6.0) * c;
    const double p1 = (-1.0 / 6.0) * b + (5.0 / 6.0) * c + (1.0 / 3.0) * d;
    const double p2 = (1.0 / 3.0) * c + (5.0 / 6.0) * d + (-1.0 / 6.0) * e;
    double a0, a1, a2;
    if constexpr (STENCIL == JFA_WENO5_JS) {
        // alpha_k = d_k / (beta_k^2 + eps), multiplied by prod_k (beta_k^2 + eps)
        const double q0 = beta0 * beta0 + eps;
        const double q1 = beta1 * beta1 + eps;
        const double q2 = beta2 * beta2 + eps;
        a0 = 0.1 * (q1 * q2);
        a1 = 0.6 * (q0 * q2);
        a2 = 0.3 * (q0 * q1);
    } else {
        // alpha_k = d_k (1 + tau5/(beta_k + eps)) = d_k (beta_k + eps + tau5)/(beta_k + eps)
        const double tau5 = fabs(beta0 - beta2);
        const double q0 = beta0 + eps;
        const double q1 = beta1 + eps;
        const double q2 = beta2 + eps;
        a0 = (0.1 * (q0 + tau5)) * (q1 * q2);
        a1 = (0.6 * (q1 + tau5)) * (q0 * q2);
        a2 = (0.3 * (q2 + tau5)) * (q0 * q1);
    }
    return (a0 * p0 + a1 * p1 + a2 * p2) * rcp(a0 + a1 + a2);
}

// ---------------------------------------------------------------- ideal gas
// materials/single_materials/ideal_gas.py:69-88, equation_manager.py:93-101,164-171
// c = sqrt(gamma p / rho) = gamma p / sqrt(gamma p rho)
__device__ __forceinline__ double sound_speed(double p, double rho, double gamma) {
    const double gp = gamma * p;
    return gp * rsq(gp * rho);
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
        const double r = rsq(gp * f.P[0]);
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
        const double rho_div = rcp(sL + sR);
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
        f.inv_c = rsq(f.cc);
        f.c = f.cc * f.inv_c;                                                      // :275
        f.rcc = f.cc * f.P[0];
        if (FULL) {
            f.H = (sL * total_enthalpy(WL, prm) + sR * total_enthalpy(WR, prm)) * rho_div;  // :250
            f.q2 = f.P[1] * f.P[1] + f.P[2] * f.P[2] + f.P[3] * f.P[3];                    // :271
        }
    }
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
            const double yL = rsq(rL), yR = rsq(rR);
            const double sL = rL * yL, sR = rR * yR;  // sqrt(rho)
            const double one_dens = rcp(sL + sR);
            const double eta2 = 0.5 * sL * sR * one_dens * one_dens;
            const double u_bar = (sL * uL + sR * uR) * one_dens;
            const double du = uR - uL;
            const double d_bar = fsqrt((sL * aL * aL + sR * aR * aR) * one_dens + eta2 * (du * du));
            SL = fmin(u_bar - d_bar, uL - aL);
            SR = fmax(u_bar + d_bar, uR + aR);
        } break;
    }
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
    const double inv = rcp(SK - Ss);
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

template <int AXIS, int RIEMANN>
__device__ __forceinline__ void riemann(const double WL[5], const double WR[5], const double UL[5],
                                        const double UR[5], const Params& prm, double F[5]) {
    if constexpr (RIEMANN == JFA_HLLC)
        hllc<AXIS>(WL, WR, UL, UR, prm, F);
    else
        rusanov<AXIS>(WL, WR, UL, UR, prm, F);
}

// ---------------------------------------------------------------- reconstruction
// HighOrderGodunov.reconstruct_xi (high_order_godunov.py:233-419) for one face.
// win[k][i]: variable i of cell k-3 relative to the cell right of the face (k = 0..5),
// i.e. the 6-cell stencil window of eigendecomposition.py:75-95.  For VAR PRIMITIVE and
// CHAR-PRIMITIVE `win` holds primitives; for CONSERVATIVE and CHAR-CONSERVATIVE it holds
// conservatives and `wl`, `wr` hold the primitives of the two cells adjacent to the face.
// `win` is read only: the marching sweeps keep it in registers across faces.
template <int AXIS, int VAR, int STENCIL>
__device__ __forceinline__ void face_states(const double (&win)[6][5], const double wl[5], const double wr[5],
                                            const Params& prm, double WL[5], double WR[5], double UL[5],
                                            double UR[5]) {
    constexpr int v = Ax<AXIS>::v, m0 = Ax<AXIS>::m0, m1 = Ax<AXIS>::m1;
    const double eps = prm.eps;
    if constexpr (VAR == JFA_PRIMITIVE || VAR == JFA_CONSERVATIVE) {
        double* QL = VAR == JFA_PRIMITIVE ? WL : UL;
        double* QR = VAR == JFA_PRIMITIVE ? WR : UR;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            // WENO5Base slices j=0: cells -3..1, j=1: cells 2..-2 (weno5_base.py:30)
            QL[i] = weno5<STENCIL>(win[0][i], win[1][i], win[2][i], win[3][i], win[4][i], eps);
            QR[i] = weno5<STENCIL>(win[5][i], win[4][i], win[3][i], win[2][i], win[1][i], eps);
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
        frozen_state<false>(win[2], win[3], prm, f);
        // get_characteristics_from_primitives (eigendecomposition.py:425-431):
        // k1 = 0.5/c, k2 = 0.5/(c^2 rho), k3 = 1/c^2
        const double k1 = 0.5 * f.inv_c;
        const double k3 = f.inv_c * f.inv_c;
        const double k2 = 0.5 * rcp(f.rcc);
        // one characteristic field at a time (x0 = k2 p - k1 u_n, x1 = rho - k3 p,
        // x2, x3 = tangential velocities, x4 = k2 p + k1 u_n), each reconstructed L and R
        double cLv[5], cRv[5], t[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) t[k] = k2 * win[k][4] - k1 * win[k][v];
        cLv[0] = weno5<STENCIL>(t[0], t[1], t[2], t[3], t[4], eps);
        cRv[0] = weno5<STENCIL>(t[5], t[4], t[3], t[2], t[1], eps);
#pragma unroll
        for (int k = 0; k < 6; ++k) t[k] = win[k][0] - k3 * win[k][4];
        cLv[1] = weno5<STENCIL>(t[0], t[1], t[2], t[3], t[4], eps);
        cRv[1] = weno5<STENCIL>(t[5], t[4], t[3], t[2], t[1], eps);
        cLv[2] = weno5<STENCIL>(win[0][m0], win[1][m0], win[2][m0], win[3][m0], win[4][m0], eps);
        cRv[2] = weno5<STENCIL>(win[5][m0], win[4][m0], win[3][m0], win[2][m0], win[1][m0], eps);
        cLv[3] = weno5<STENCIL>(win[0][m1], win[1][m1], win[2][m1], win[3][m1], win[4][m1], eps);
        cRv[3] = weno5<STENCIL>(win[5][m1], win[4][m1], win[3][m1], win[2][m1], win[1][m1], eps);
#pragma unroll
        for (int k = 0; k < 6; ++k) t[k] = k2 * win[k][4] + k1 * win[k][v];
        cLv[4] = weno5<STENCIL>(t[0], t[1], t[2], t[3], t[4], eps);
        cRv[4] = weno5<STENCIL>(t[5], t[4], t[3], t[2], t[1], eps);
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
        // eigendecomposition_conservatives (eigendecomposition.py:595-658)
        const double one_cc = f.inv_c * f.inv_c;
        const double one_rho = rcp(f.P[0]);
        const double* P = f.P;
        const double g = f.g, H = f.H, c = f.c, q2 = f.q2;
        double L[5][5], R[5][5];
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
        // transformtochar: einsum over j (eigendecomposition.py:729-730)
        double ch[6][5];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
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
            cLv[i] = weno5<STENCIL>(ch[0][i], ch[1][i], ch[2][i], ch[3][i], ch[4][i], eps);
            cRv[i] = weno5<STENCIL>(ch[5][i], ch[4][i], ch[3][i], ch[2][i], ch[1][i], eps);
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

}  // namespace jfa
