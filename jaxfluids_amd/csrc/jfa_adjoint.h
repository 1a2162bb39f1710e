// jfa_adjoint.h -- reverse mode (vector-Jacobian products) of the RK3 stage at fixed dt for
// primitive-variable WENO5 reconstruction with the HLLC or Rusanov Riemann solver.
//
// The forward stage is   U_out = c0 U_in + c1 U_n + ts R(U_in),   ts = m dt,
// with R the flux divergence of prim(U_in) after the face-halo fill (RK3.py:27-62,
// space_solver.py:266-314).  Given lam = dL/dU_out on the interior, one backward stage adds
//   lam_in += c0 lam + (dW/dU)^T  halo^T  sum_faces (dF/dW_window)^T mu_f,
//   lam_n  += c1 lam,           mu_f = ts (lam_{f} - lam_{f-1}) / dx   (face f between cells f-1, f)
// where halo^T sends the cotangent of every face-halo cell back to the interior cell the halo
// fill copied it from (with the symmetric-wall sign flip).
//
// Kernels mirror the forward sweeps, transposed:
// * adj_march_kernel (every active axis but the innermost): lanes along the innermost axis,
//   each thread walks the faces of a column and accumulates window cotangents in a 6-cell
//   register window; each chunk owns a range of cells and evaluates every face that touches
//   them (5 extra faces per chunk), so all writes are owned -- no atomics.
// * adj_lane_kernel (the innermost axis, and any axis whose marching sweep is too small to fill
//   the device): one face per lane; a cell gathers the six window cotangents of the faces
//   f = c-2..c+3 from lanes l..l-5 by wave shuffles; a wave owns 59 cells (lanes 5..63).
// * adj_dissip_cell_kernel (with viscous / heat fluxes): per cell, the VJPs of its two faces on
//   every axis into gW, and their transverse-derivative terms into X, a 9-field per-cell sum
//   gathered from the periodic neighbours by adj_update_kernel.
// * adj_update_kernel: halo^T, (dW/dU)^T and the stage blend per interior cell.
// Face states are recomputed with the forward's own face_states, so branch choices (the sign
// of S*, the min/max of the wave-speed estimates) are those of the forward evaluation.
#pragma once

#include "jfa_common.h"
#include "jfa_dissipative.h"

namespace jfa {

// ------------------------------------------------------------------ WENO5 VJP
// Adds vbar * d v / d(a..e) to g for v = WENO5(a, b, c, d, e), the reference's formulas
// (weno5_base.py:34-50, weno5_js.py:39-58, weno5_z.py:38-59).
template <int STENCIL>
__device__ __forceinline__ void weno5_vjp(double a, double b, double c, double d, double e, double eps, double vbar,
                                          double& ga, double& gb, double& gc, double& gd, double& ge) {
    const double D0 = a - 2.0 * b + c, E0 = a - 4.0 * b + 3.0 * c;
    const double D1 = b - 2.0 * c + d, E1 = b - d;
    const double D2 = c - 2.0 * d + e, E2 = 3.0 * c - 4.0 * d + e;
    const double b0 = 13.0 / 12.0 * D0 * D0 + 0.25 * E0 * E0;
    const double b1 = 13.0 / 12.0 * D1 * D1 + 0.25 * E1 * E1;
    const double b2 = 13.0 / 12.0 * D2 * D2 + 0.25 * E2 * E2;
    const double p0 = (1.0 / 3.0) * a - (7.0 / 6.0) * b + (11.0 / 6.0) * c;
    const double p1 = (-1.0 / 6.0) * b + (5.0 / 6.0) * c + (1.0 / 3.0) * d;
    const double p2 = (1.0 / 3.0) * c + (5.0 / 6.0) * d - (1.0 / 6.0) * e;
    constexpr double d0 = 0.1, d1 = 0.6, d2 = 0.3;
    double al0, al1, al2, da0, da1, da2;  // alpha_k and d alpha_k / d beta_k
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, tau = 0.0;
    if constexpr (STENCIL == JFA_WENO5_JS) {
        const double q0 = 1.0 / (b0 * b0 + eps), q1 = 1.0 / (b1 * b1 + eps), q2 = 1.0 / (b2 * b2 + eps);
        al0 = d0 * q0;
        al1 = d1 * q1;
        al2 = d2 * q2;
        da0 = -2.0 * b0 * al0 * q0;
        da1 = -2.0 * b1 * al1 * q1;
        da2 = -2.0 * b2 * al2 * q2;
    } else {
        tau = fabs(b0 - b2);
        r0 = 1.0 / (b0 + eps);
        r1 = 1.0 / (b1 + eps);
        r2 = 1.0 / (b2 + eps);
        al0 = d0 * (1.0 + tau * r0);
        al1 = d1 * (1.0 + tau * r1);
        al2 = d2 * (1.0 + tau * r2);
        da0 = -d0 * tau * r0 * r0;
        da1 = -d1 * tau * r1 * r1;
        da2 = -d2 * tau * r2 * r2;
    }
    const double iS = 1.0 / (al0 + al1 + al2);
    const double w0 = al0 * iS, w1 = al1 * iS, w2 = al2 * iS;
    const double v = w0 * p0 + w1 * p1 + w2 * p2;
    // omega_k = alpha_k / S:  abar_k = vbar (p_k - v) / S
    const double ab0 = vbar * (p0 - v) * iS, ab1 = vbar * (p1 - v) * iS, ab2 = vbar * (p2 - v) * iS;
    double bb0 = ab0 * da0, bb1 = ab1 * da1, bb2 = ab2 * da2;
    if constexpr (STENCIL == JFA_WENO5_Z) {
        const double tb = ab0 * d0 * r0 + ab1 * d1 * r1 + ab2 * d2 * r2;
        const double s = b0 > b2 ? 1.0 : (b0 < b2 ? -1.0 : 0.0);
        bb0 += tb * s;
        bb2 -= tb * s;
    }
    const double pb0 = vbar * w0, pb1 = vbar * w1, pb2 = vbar * w2;
    ga += (1.0 / 3.0) * pb0;
    gb += (-7.0 / 6.0) * pb0 + (-1.0 / 6.0) * pb1;
    gc += (11.0 / 6.0) * pb0 + (5.0 / 6.0) * pb1 + (1.0 / 3.0) * pb2;
    gd += (1.0 / 3.0) * pb1 + (5.0 / 6.0) * pb2;
    ge += (-1.0 / 6.0) * pb2;
    const double t0 = 13.0 / 6.0 * D0 * bb0, u0 = 0.5 * E0 * bb0;
    ga += t0 + u0;
    gb += -2.0 * t0 - 4.0 * u0;
    gc += t0 + 3.0 * u0;
    const double t1 = 13.0 / 6.0 * D1 * bb1, u1 = 0.5 * E1 * bb1;
    gb += t1 + u1;
    gc += -2.0 * t1;
    gd += t1 - u1;
    const double t2 = 13.0 / 6.0 * D2 * bb2, u2 = 0.5 * E2 * bb2;
    gc += t2 + 3.0 * u2;
    gd += -2.0 * t2 - 4.0 * u2;
    ge += t2 + u2;
}

// ------------------------------------------------------------------ pointwise VJPs
// U = cons(W) (equation_manager.py:93-101): adds (dU/dW)^T gU to gW.
__device__ __forceinline__ void cons_vjp(const double W[5], const double gU[5], const Params& prm, double gW[5]) {
    const double q2 = W[1] * W[1] + W[2] * W[2] + W[3] * W[3];
    gW[0] += gU[0] + gU[1] * W[1] + gU[2] * W[2] + gU[3] * W[3] + gU[4] * 0.5 * q2;
#pragma unroll
    for (int j = 1; j < 4; ++j) gW[j] += (gU[j] + gU[4] * W[j]) * W[0];
    gW[4] += gU[4] * prm.inv_gm1;
}

// W = prim(U) (equation_manager.py:164-171): returns (dW/dU)^T gW, written with W.
__device__ __forceinline__ void prim_vjp(const double W[5], const double gW[5], const Params& prm, double gU[5]) {
    const double ir = 1.0 / W[0];
    const double q2 = W[1] * W[1] + W[2] * W[2] + W[3] * W[3];
    gU[0] = gW[0] - (W[1] * gW[1] + W[2] * gW[2] + W[3] * gW[3]) * ir + prm.gm1 * 0.5 * q2 * gW[4];
#pragma unroll
    for (int j = 1; j < 4; ++j) gU[j] = gW[j] * ir - prm.gm1 * W[j] * gW[4];
    gU[4] = prm.gm1 * gW[4];
}

// physical flux along AXIS (equation_manager.py:237-252): adds its VJP to gW, gU.
template <int AXIS>
__device__ __forceinline__ void physical_flux_vjp(const double W[5], const double U[5], const double gF[5],
                                                  double gW[5], double gU[5]) {
    constexpr int v = Ax<AXIS>::v;
    gU[v] += gF[0] + gF[1] * W[1] + gF[2] * W[2] + gF[3] * W[3];
#pragma unroll
    for (int j = 1; j < 4; ++j) gW[j] += gF[j] * U[v];
    gW[4] += gF[v];
    gW[v] += gF[4] * (U[4] + W[4]);
    gU[4] += gF[4] * W[v];
    gW[4] += gF[4] * W[v];
}

// c = sqrt(gamma p / rho) (ideal_gas.py:69-71)
__device__ __forceinline__ void sound_speed_vjp(double c, double rho, double gamma, double gc, double& grho,
                                                double& gp) {
    gp += gc * gamma / (2.0 * c * rho);
    grho -= gc * c / (2.0 * rho);
}

// Shares of a min / max cotangent g for its operands x, y: all of it to the one the forward
// takes, half each on a tie (jax's balanced rule for lax.min / lax.max; a NaN operand gets the
// halves too).
__device__ __forceinline__ void minmax_vjp(bool is_min, double x, double y, double g, double& gx, double& gy) {
    const bool x_wins = is_min ? x < y : x > y, y_wins = is_min ? y < x : y > x;
    // share of x as a value, not a branch per output: a branch that adds g to one of two
    // references made the compiler keep the pair in private (scratch) memory
    const double wx = x_wins ? 1.0 : (y_wins ? 0.0 : 0.5);
    gx += wx * g;
    gy += (1.0 - wx) * g;
}

// Wave-speed estimates (signal_speeds.py): adds the VJP of (SL, SR) to u, a (sound speed), rho
// and (TORO) p.
__device__ __forceinline__ void signal_speeds_vjp(int kind, double uL, double uR, double aL, double aR, double rL,
                                                  double rR, double pL, double pR, double gamma, double gSL,
                                                  double gSR, double& guL, double& guR, double& gaL, double& gaR,
                                                  double& grL, double& grR, double& gpL, double& gpR) {
    switch (kind) {
        case JFA_SS_TORO: {  // signal_speeds.py:135-157, estimate_pressure :201-231
            const double rho_bar = 0.5 * (rL + rR), a_bar = 0.5 * (aL + aR);
            const double p_pvrs = 0.5 * (pL + pR) - 0.5 * (uR - uL) * rho_bar * a_bar;
            const double ps = fmax(0.0, p_pvrs);
            const double g_ = (gamma + 1.0) * 0.5 / gamma;
            // q_K = 1 (p* <= p_K) or sqrt(1 + g_ (p*/p_K - 1)); S_L = uL - aL qL, S_R = uR + aR qR
            double gps = 0.0;
            guL += gSL;
            guR += gSR;
            if (ps > pL) {
                const double qL = sqrt(1.0 + g_ * (ps / pL - 1.0));
                gaL -= gSL * qL;
                const double gq = -gSL * aL * g_ / (2.0 * qL * pL);
                gps += gq;
                gpL -= gq * ps / pL;
            } else {
                gaL -= gSL;
            }
            if (ps > pR) {
                const double qR = sqrt(1.0 + g_ * (ps / pR - 1.0));
                gaR += gSR * qR;
                const double gq = gSR * aR * g_ / (2.0 * qR * pR);
                gps += gq;
                gpR -= gq * ps / pR;
            } else {
                gaR += gSR;
            }
            double gzero = 0.0, gpv = 0.0;
            minmax_vjp(false, 0.0, p_pvrs, gps, gzero, gpv);  // p* = max(0, p_pvrs)
            gpL += 0.5 * gpv;
            gpR += 0.5 * gpv;
            const double du = uR - uL;
            guR -= 0.5 * rho_bar * a_bar * gpv;
            guL += 0.5 * rho_bar * a_bar * gpv;
            const double grb = -0.5 * du * a_bar * gpv, gab = -0.5 * du * rho_bar * gpv;
            grL += 0.5 * grb;
            grR += 0.5 * grb;
            gaL += 0.5 * gab;
            gaR += 0.5 * gab;
        } break;
        case JFA_SS_ARITHMETIC: {
            const double um = 0.5 * (uL + uR), am = 0.5 * (aL + aR);
            double gum = 0.0, gam = 0.0, g1 = 0.0, g2 = 0.0;
            minmax_vjp(true, um - am, uL - aL, gSL, g1, g2);  // S_L = min(um - am, uL - aL)
            gum += g1;
            gam -= g1;
            guL += g2;
            gaL -= g2;
            g1 = g2 = 0.0;
            minmax_vjp(false, um + am, uR + aR, gSR, g1, g2);  // S_R = max(um + am, uR + aR)
            gum += g1;
            gam += g1;
            guR += g2;
            gaR += g2;
            guL += 0.5 * gum;
            guR += 0.5 * gum;
            gaL += 0.5 * gam;
            gaR += 0.5 * gam;
        } break;
        case JFA_SS_RUSANOV: {
            double g1 = 0.0, g2 = 0.0;  // S_+ = max(|uL| + aL, |uR| + aR), (S_L, S_R) = (-S_+, S_+)
            minmax_vjp(false, fabs(uL) + aL, fabs(uR) + aR, gSR - gSL, g1, g2);
            guL += g1 * (uL > 0.0 ? 1.0 : (uL < 0.0 ? -1.0 : 0.0));
            gaL += g1;
            guR += g2 * (uR > 0.0 ? 1.0 : (uR < 0.0 ? -1.0 : 0.0));
            gaR += g2;
        } break;
        case JFA_SS_DAVIS: {
            double g1 = 0.0, g2 = 0.0;
            minmax_vjp(true, uL - aL, uR - aR, gSL, g1, g2);
            guL += g1;
            gaL -= g1;
            guR += g2;
            gaR -= g2;
            g1 = g2 = 0.0;
            minmax_vjp(false, uL + aL, uR + aR, gSR, g1, g2);
            guL += g1;
            gaL += g1;
            guR += g2;
            gaR += g2;
        } break;
        default: {  // EINFELDT
            const double sL = sqrt(rL), sR = sqrt(rR);
            const double od = 1.0 / (sL + sR);
            const double eta2 = 0.5 * sL * sR * od * od;
            const double t2 = sL * uL + sR * uR;
            const double ub = t2 * od;
            const double du = uR - uL;
            const double t1 = sL * aL * aL + sR * aR * aR;
            const double db = sqrt(t1 * od + eta2 * du * du);
            double gub = 0.0, gdb = 0.0, g1 = 0.0, g2 = 0.0;
            minmax_vjp(true, ub - db, uL - aL, gSL, g1, g2);
            gub += g1;
            gdb -= g1;
            guL += g2;
            gaL -= g2;
            g1 = g2 = 0.0;
            minmax_vjp(false, ub + db, uR + aR, gSR, g1, g2);
            gub += g1;
            gdb += g1;
            guR += g2;
            gaR += g2;
            const double gQ = db > 0.0 ? gdb / (2.0 * db) : 0.0;
            const double gt1 = gQ * od;
            double god = gQ * t1;
            const double geta2 = gQ * du * du;
            const double gdu = gQ * 2.0 * eta2 * du;
            guR += gdu;
            guL -= gdu;
            double gsL = gt1 * aL * aL, gsR = gt1 * aR * aR;
            gaL += gt1 * 2.0 * sL * aL;
            gaR += gt1 * 2.0 * sR * aR;
            const double gt2 = gub * od;
            god += gub * t2;
            gsL += gt2 * uL;
            guL += gt2 * sL;
            gsR += gt2 * uR;
            guR += gt2 * sR;
            gsL += geta2 * 0.5 * sR * od * od;
            gsR += geta2 * 0.5 * sL * od * od;
            god += geta2 * sL * sR * od;
            const double gsum = -god * od * od;
            gsL += gsum;
            gsR += gsum;
            grL += gsL / (2.0 * sL);
            grR += gsR / (2.0 * sR);
        } break;
    }
}

// HLLC star state of side K (HLLC.py:52-63, hllc_star_state): d = SK - u, inv = 1/(SK - Ss),
// pre = rho d inv, u* = (pre, pre Ss | pre w_minor, inv (d E + (Ss - u)(rho d Ss + p))).
// Adds its VJP for cotangent gus to gW, gU, gSK, gSs.
template <int AXIS>
__device__ __forceinline__ void hllc_state_vjp(const double W[5], const double U[5], double SK, double Ss,
                                               const double gus[5], double gW[5], double gU[5], double& gSK,
                                               double& gSs) {
    constexpr int v = Ax<AXIS>::v, m0 = Ax<AXIS>::m0, m1 = Ax<AXIS>::m1;
    const double u = W[v];
    const double d = SK - u;
    const double inv = 1.0 / (SK - Ss);
    const double rd = W[0] * d;
    const double pre = rd * inv;
    const double A = Ss - u;
    const double B = rd * Ss + W[4];
    const double Q = d * U[4] + A * B;
    const double gpre = gus[0] + gus[v] * Ss + gus[m0] * W[m0] + gus[m1] * W[m1];
    gSs += gus[v] * pre;
    gW[m0] += gus[m0] * pre;
    gW[m1] += gus[m1] * pre;
    double ginv = gus[4] * Q;
    const double gQ = gus[4] * inv;
    double gd = gQ * U[4];
    gU[4] += gQ * d;
    const double gA = gQ * B, gB = gQ * A;
    double grd = gB * Ss;
    gSs += gB * rd;
    gW[4] += gB;
    gSs += gA;
    double gu = -gA;
    grd += gpre * inv;
    ginv += gpre * rd;
    gW[0] += grd * d;
    gd += grd * W[0];
    const double gM = -ginv * inv * inv;
    gSK += gM;
    gSs -= gM;
    gSK += gd;
    gu -= gd;
    gW[v] += gu;
}

// HLLC star flux of side K (HLLC.py:41-78, the form of hllc_star_flux):
//   F = phys(W, U) + Sc (u* - U) with the star state u* above.
// Adds its VJP for cotangent mu to gW, gU, gSK, gSs.
template <int AXIS, bool LEFT>
__device__ __forceinline__ void hllc_star_vjp(const double W[5], const double U[5], double SK, double Ss,
                                              const double mu[5], double gW[5], double gU[5], double& gSK,
                                              double& gSs) {
    double us[5];
    hllc_star_state<AXIS>(W, U, SK, Ss, us);
    const double Sc = LEFT ? fmin(SK, 0.0) : fmax(SK, 0.0);
    double gSc = 0.0, gus[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        gSc += mu[i] * (us[i] - U[i]);
        gus[i] = Sc * mu[i];
        gU[i] -= Sc * mu[i];
    }
    double gzero = 0.0;  // Sc = min(SK, 0) (left) / max(SK, 0) (right)
    minmax_vjp(LEFT, SK, 0.0, gSc, gSK, gzero);
    hllc_state_vjp<AXIS>(W, U, SK, Ss, gus, gW, gU, gSK, gSs);
    physical_flux_vjp<AXIS>(W, U, mu, gW, gU);
}

// HLLC (HLLC.py:80-126): adds the VJP of the face flux for cotangent mu to gWL, gWR.
template <int AXIS>
__device__ __forceinline__ void hllc_vjp(const double WL[5], const double WR[5], const double UL[5],
                                         const double UR[5], const double mu[5], const Params& prm, double gWL[5],
                                         double gWR[5]) {
    constexpr int v = Ax<AXIS>::v;
    const double g = prm.gamma;
    const double cL = sqrt(g * WL[4] / WL[0]), cR = sqrt(g * WR[4] / WR[0]);
    double SL, SR;
    signal_speeds(prm.signal_speed, WL[v], WR[v], cL, cR, WL[0], WR[0], WL[4], WR[4], g, SL, SR);
    const double aL = WL[0] * (SL - WL[v]), aR = WR[0] * (SR - WR[v]);
    const double num = (WR[4] - WL[4]) + (WL[v] * aL - WR[v] * aR);
    const double den = aL - aR;
    const double Ss = num / den;
    double gUL[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gUR[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double gSL = 0.0, gSR = 0.0, gSs = 0.0;
    if (Ss >= 0.0) {
        double m[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) m[i] = Ss == 0.0 ? 0.5 * mu[i] : mu[i];
        hllc_star_vjp<AXIS, true>(WL, UL, SL, Ss, m, gWL, gUL, gSL, gSs);
    }
    if (Ss <= 0.0) {
        double m[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) m[i] = Ss == 0.0 ? 0.5 * mu[i] : mu[i];
        hllc_star_vjp<AXIS, false>(WR, UR, SR, Ss, m, gWR, gUR, gSR, gSs);
    }
    // S* = num / den (signal_speeds.py:194-197)
    const double gnum = gSs / den, gden = -gSs * Ss / den;
    gWR[4] += gnum;
    gWL[4] -= gnum;
    gWL[v] += gnum * aL;
    gWR[v] -= gnum * aR;
    const double gaL = gnum * WL[v] + gden, gaR = -gnum * WR[v] - gden;
    gWL[0] += gaL * (SL - WL[v]);
    gSL += gaL * WL[0];
    gWL[v] -= gaL * WL[0];
    gWR[0] += gaR * (SR - WR[v]);
    gSR += gaR * WR[0];
    gWR[v] -= gaR * WR[0];
    double gcL = 0.0, gcR = 0.0;
    signal_speeds_vjp(prm.signal_speed, WL[v], WR[v], cL, cR, WL[0], WR[0], WL[4], WR[4], g, gSL, gSR, gWL[v],
                      gWR[v], gcL, gcR, gWL[0], gWR[0], gWL[4], gWR[4]);
    sound_speed_vjp(cL, WL[0], g, gcL, gWL[0], gWL[4]);
    sound_speed_vjp(cR, WR[0], g, gcR, gWR[0], gWR[4]);
    cons_vjp(WL, gUL, prm, gWL);
    cons_vjp(WR, gUR, prm, gWR);
}

// HLL (HLL.py:29-61): F = N / D with N = s_R F_L - s_L F_R + s_L s_R (U_R - U_L),
// D = s_R - s_L + eps, s_L = min(S_L, 0), s_R = max(S_R, 0); adds the VJP for cotangent mu.
template <int AXIS>
__device__ __forceinline__ void hll_vjp(const double WL[5], const double WR[5], const double UL[5],
                                        const double UR[5], const double mu[5], const Params& prm, double gWL[5],
                                        double gWR[5]) {
    constexpr int v = Ax<AXIS>::v;
    const double g = prm.gamma;
    const double cL = sqrt(g * WL[4] / WL[0]), cR = sqrt(g * WR[4] / WR[0]);
    double SL, SR;
    signal_speeds(prm.signal_speed, WL[v], WR[v], cL, cR, WL[0], WR[0], WL[4], WR[4], g, SL, SR);
    const double sl = fmin(SL, 0.0), sr = fmax(SR, 0.0);
    const double inv = 1.0 / (sr - sl + prm.eps_precision);
    double FL[5], FR[5];
    physical_flux<AXIS>(WL, UL, FL);
    physical_flux<AXIS>(WR, UR, FR);
    double gUL[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gUR[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gFL[5], gFR[5];
    double gsl = 0.0, gsr = 0.0, gD = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const double gN = mu[i] * inv;
        const double dU = UR[i] - UL[i];
        gD -= gN * ((sr * FL[i] - sl * FR[i] + sl * sr * dU) * inv);
        gFL[i] = sr * gN;
        gFR[i] = -sl * gN;
        gUR[i] += sl * sr * gN;
        gUL[i] -= sl * sr * gN;
        gsr += gN * (FL[i] + sl * dU);
        gsl += gN * (sr * dU - FR[i]);
    }
    gsr += gD;
    gsl -= gD;
    physical_flux_vjp<AXIS>(WL, UL, gFL, gWL, gUL);
    physical_flux_vjp<AXIS>(WR, UR, gFR, gWR, gUR);
    double gSL = 0.0, gSR = 0.0, gz = 0.0;
    minmax_vjp(true, SL, 0.0, gsl, gSL, gz);   // s_L = min(S_L, 0)
    minmax_vjp(false, SR, 0.0, gsr, gSR, gz);  // s_R = max(S_R, 0)
    double gcL = 0.0, gcR = 0.0;
    signal_speeds_vjp(prm.signal_speed, WL[v], WR[v], cL, cR, WL[0], WR[0], WL[4], WR[4], g, gSL, gSR, gWL[v],
                      gWR[v], gcL, gcR, gWL[0], gWR[0], gWL[4], gWR[4]);
    sound_speed_vjp(cL, WL[0], g, gcL, gWL[0], gWL[4]);
    sound_speed_vjp(cR, WR[0], g, gcR, gWR[0], gWR[4]);
    cons_vjp(WL, gUL, prm, gWL);
    cons_vjp(WR, gUR, prm, gWR);
}

// HLLC-LM (HLLCLM.py:59-136; hllc_lm of the forward): F = bL F_L + bR F_R + bS F* with the
// sign blends bL, bR, bS (piecewise constant: jax's derivative of sign is zero) and
//   F* = (F_L + F_R)/2 + (phi S_L (u*_L - U_L) + |S*| (u*_L - u*_R) + phi S_R (u*_R - U_R))/2,
// phi = sin(min(1, Ma / 0.1) pi / 2), Ma = max(|u_L / c_L|, |u_R / c_R|).  Adds the VJP for
// cotangent mu.
template <int AXIS>
__device__ __forceinline__ void hllc_lm_vjp(const double WL[5], const double WR[5], const double UL[5],
                                            const double UR[5], const double mu[5], const Params& prm,
                                            double gWL[5], double gWR[5]) {
    constexpr int v = Ax<AXIS>::v;
    const double g = prm.gamma;
    const double cL = sqrt(g * WL[4] / WL[0]), cR = sqrt(g * WR[4] / WR[0]);
    double SL, SR;
    signal_speeds(prm.signal_speed, WL[v], WR[v], cL, cR, WL[0], WR[0], WL[4], WR[4], g, SL, SR);
    const double aL = WL[0] * (SL - WL[v]), aR = WR[0] * (SR - WR[v]);
    const double num = (WR[4] - WL[4]) + (WL[v] * aL - WR[v] * aR);
    const double den = aL - aR;
    const double Ss = num / den;
    double usL[5], usR[5];
    hllc_star_state<AXIS>(WL, UL, SL, Ss, usL);
    hllc_star_state<AXIS>(WR, UR, SR, Ss, usR);
    const double xL = WL[v] / cL, xR = WR[v] / cR;
    const double Ma = fmax(fabs(xL), fabs(xR));
    const double mq = fmin(1.0, Ma / 0.1);
    const double hp = 0.5 * 3.141592653589793;
    const double phi = sin(mq * hp);
    const double wL = phi * SL, wR = phi * SR, aS = fabs(Ss);
    const double bL = 0.5 * (1.0 + sgn(SL)), bR = 0.5 * (1.0 - sgn(SR));
    const double bS = 0.25 * (1.0 - sgn(SL)) * (1.0 + sgn(SR));
    double gUL[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gUR[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double gFL[5], gFR[5], gusL[5], gusR[5];
    double gwL = 0.0, gwR = 0.0, gaS = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const double gFs = bS * mu[i];
        gFL[i] = bL * mu[i] + 0.5 * gFs;
        gFR[i] = bR * mu[i] + 0.5 * gFs;
        gusL[i] = 0.5 * gFs * (wL + aS);
        gusR[i] = 0.5 * gFs * (wR - aS);
        gUL[i] -= 0.5 * gFs * wL;
        gUR[i] -= 0.5 * gFs * wR;
        gwL += 0.5 * gFs * (usL[i] - UL[i]);
        gwR += 0.5 * gFs * (usR[i] - UR[i]);
        gaS += 0.5 * gFs * (usL[i] - usR[i]);
    }
    physical_flux_vjp<AXIS>(WL, UL, gFL, gWL, gUL);
    physical_flux_vjp<AXIS>(WR, UR, gFR, gWR, gUR);
    double gSL = gwL * phi, gSR = gwR * phi, gSs = gaS * sgn(Ss);
    const double gphi = gwL * SL + gwR * SR;
    double gone = 0.0, gq = 0.0, g1 = 0.0, g2 = 0.0;
    minmax_vjp(true, 1.0, Ma / 0.1, gphi * cos(mq * hp) * hp, gone, gq);  // min(1, Ma / 0.1)
    minmax_vjp(false, fabs(xL), fabs(xR), gq / 0.1, g1, g2);             // Ma = max(|xL|, |xR|)
    const double gxL = g1 * sgn(xL), gxR = g2 * sgn(xR);
    double gcL = -gxL * xL / cL, gcR = -gxR * xR / cR;
    gWL[v] += gxL / cL;
    gWR[v] += gxR / cR;
    hllc_state_vjp<AXIS>(WL, UL, SL, Ss, gusL, gWL, gUL, gSL, gSs);
    hllc_state_vjp<AXIS>(WR, UR, SR, Ss, gusR, gWR, gUR, gSR, gSs);
    // S* = num / den (signal_speeds.py:194-197)
    const double gnum = gSs / den, gden = -gSs * Ss / den;
    gWR[4] += gnum;
    gWL[4] -= gnum;
    gWL[v] += gnum * aL;
    gWR[v] -= gnum * aR;
    const double gaL = gnum * WL[v] + gden, gaR = -gnum * WR[v] - gden;
    gWL[0] += gaL * (SL - WL[v]);
    gSL += gaL * WL[0];
    gWL[v] -= gaL * WL[0];
    gWR[0] += gaR * (SR - WR[v]);
    gSR += gaR * WR[0];
    gWR[v] -= gaR * WR[0];
    signal_speeds_vjp(prm.signal_speed, WL[v], WR[v], cL, cR, WL[0], WR[0], WL[4], WR[4], g, gSL, gSR, gWL[v],
                      gWR[v], gcL, gcR, gWL[0], gWR[0], gWL[4], gWR[4]);
    sound_speed_vjp(cL, WL[0], g, gcL, gWL[0], gWL[4]);
    sound_speed_vjp(cR, WR[0], g, gcR, gWR[0], gWR[4]);
    cons_vjp(WL, gUL, prm, gWL);
    cons_vjp(WR, gUR, prm, gWR);
}

// AUSM+ (AUSMP.py:30-103; ausmp of the forward): F = a (M+ phi_L + M- phi_R) + (P+ p_L + P- p_R) e_n
// with a = (c_L + c_R)/2, M = M+(M_L) + M-(M_R), M+- = (M +- |M|)/2, phi = (rho, rho u, E + p) and
// the split functions of M_L = u_L / a, M_R = u_R / a, each piecewise (|M| >= 1: supersonic
// branch, jnp.where gives its derivative only).  Adds the VJP for cotangent mu.
template <int AXIS>
__device__ __forceinline__ void ausmp_vjp(const double WL[5], const double WR[5], const double UL[5],
                                          const double UR[5], const double mu[5], const Params& prm,
                                          double gWL[5], double gWR[5]) {
    constexpr int v = Ax<AXIS>::v;
    constexpr double alpha = 3.0 / 16.0, beta = 1.0 / 8.0;
    const double g = prm.gamma;
    const double cL = sqrt(g * WL[4] / WL[0]), cR = sqrt(g * WR[4] / WR[0]);
    const double a = 0.5 * (cL + cR);
    const double Ml = WL[v] / a, Mr = WR[v] / a;
    const double ql = Ml * Ml - 1.0, qr = Mr * Mr - 1.0;
    const double lp = Ml + 1.0, rm = Mr - 1.0;
    const bool sl = fabs(Ml) >= 1.0, sr = fabs(Mr) >= 1.0;
    const double Mp = sl ? 0.5 * (Ml + fabs(Ml)) : 0.25 * (lp * lp) + beta * (ql * ql);
    const double Mm = sr ? 0.5 * (Mr - fabs(Mr)) : -0.25 * (rm * rm) - beta * (qr * qr);
    const double M = Mp + Mm;
    const double Map = 0.5 * (M + fabs(M)), Mam = 0.5 * (M - fabs(M));
    const double Pp = sl ? 0.5 * (1.0 + sgn(Ml)) : 0.25 * (lp * lp) * (2.0 - Ml) + alpha * Ml * (ql * ql);
    const double Pm = sr ? 0.5 * (1.0 - sgn(Mr)) : 0.25 * (rm * rm) * (2.0 + Mr) - alpha * Mr * (qr * qr);
    double gUL[5], gUR[5];
    double ga = 0.0, gMap = 0.0, gMam = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const double phl = i == 4 ? UL[4] + WL[4] : UL[i];
        const double phr = i == 4 ? UR[4] + WR[4] : UR[i];
        ga += mu[i] * (Map * phl + Mam * phr);
        gMap += a * mu[i] * phl;
        gMam += a * mu[i] * phr;
        gUL[i] = a * Map * mu[i];
        gUR[i] = a * Mam * mu[i];
    }
    gWL[4] += gUL[4];  // phi_4 = E + p
    gWR[4] += gUR[4];
    const double gpa = mu[v];
    const double gPp = gpa * WL[4], gPm = gpa * WR[4];
    gWL[4] += gpa * Pp;
    gWR[4] += gpa * Pm;
    const double sM = sgn(M);
    const double gM = gMap * 0.5 * (1.0 + sM) + gMam * 0.5 * (1.0 - sM);
    const double dMp = sl ? 0.5 * (1.0 + sgn(Ml)) : 0.5 * lp + 4.0 * beta * Ml * ql;
    const double dMm = sr ? 0.5 * (1.0 - sgn(Mr)) : -0.5 * rm - 4.0 * beta * Mr * qr;
    const double dPp = sl ? 0.0 : 0.5 * lp * (2.0 - Ml) - 0.25 * (lp * lp) + alpha * (ql * ql) + 4.0 * alpha * Ml * Ml * ql;
    const double dPm = sr ? 0.0 : 0.5 * rm * (2.0 + Mr) + 0.25 * (rm * rm) - alpha * (qr * qr) - 4.0 * alpha * Mr * Mr * qr;
    const double gMl = gM * dMp + gPp * dPp, gMr = gM * dMm + gPm * dPm;
    gWL[v] += gMl / a;
    gWR[v] += gMr / a;
    ga -= gMl * Ml / a + gMr * Mr / a;
    double gcL = 0.5 * ga, gcR = 0.5 * ga;
    sound_speed_vjp(cL, WL[0], g, gcL, gWL[0], gWL[4]);
    sound_speed_vjp(cR, WR[0], g, gcR, gWR[0], gWR[4]);
    cons_vjp(WL, gUL, prm, gWL);
    cons_vjp(WR, gUR, prm, gWR);
}

// Rusanov (Rusanov.py:35-45)
template <int AXIS>
__device__ __forceinline__ void rusanov_vjp(const double WL[5], const double WR[5], const double UL[5],
                                            const double UR[5], const double mu[5], const Params& prm,
                                            double gWL[5], double gWR[5]) {
    constexpr int v = Ax<AXIS>::v;
    const double g = prm.gamma;
    const double cL = sqrt(g * WL[4] / WL[0]), cR = sqrt(g * WR[4] / WR[0]);
    const double alpha = fmax(fabs(WL[v]) + cL, fabs(WR[v]) + cR);
    double gUL[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gUR[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, half[5];
    double galpha = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        half[i] = 0.5 * mu[i];
        galpha -= 0.5 * mu[i] * (UR[i] - UL[i]);
        gUR[i] -= 0.5 * alpha * mu[i];
        gUL[i] += 0.5 * alpha * mu[i];
    }
    physical_flux_vjp<AXIS>(WL, UL, half, gWL, gUL);
    physical_flux_vjp<AXIS>(WR, UR, half, gWR, gUR);
    double gcL = 0.0, gcR = 0.0, g1 = 0.0, g2 = 0.0;
    minmax_vjp(false, fabs(WL[v]) + cL, fabs(WR[v]) + cR, galpha, g1, g2);  // alpha = max(., .)
    gWL[v] += g1 * (WL[v] > 0.0 ? 1.0 : (WL[v] < 0.0 ? -1.0 : 0.0));
    gcL += g1;
    gWR[v] += g2 * (WR[v] > 0.0 ? 1.0 : (WR[v] < 0.0 ? -1.0 : 0.0));
    gcR += g2;
    sound_speed_vjp(cL, WL[0], g, gcL, gWL[0], gWL[4]);
    sound_speed_vjp(cR, WR[0], g, gcR, gWR[0], gWR[4]);
    cons_vjp(WL, gUL, prm, gWL);
    cons_vjp(WR, gUR, prm, gWR);
}

// ------------------------------------------------------------------ CHAR-PRIMITIVE projection
// With an ARITHMETIC frozen state a = rho_ave, q = p_ave of the face's two cells
// (eigendecomposition.py:139-148,223-231) the projection of a window cell w
// (get_characteristics_from_primitives, :421-433) and the face states from the reconstructed
// characteristics x (get_primitives_from_characteristics, :512-523) are
//   x0 = k2 p - k1 u_n,  x1 = rho - k3 p,  x2 = u_m0,  x3 = u_m1,  x4 = k2 p + k1 u_n,
//   W_rho = a (x0 + x4) + x1,  W_n = c (x4 - x0),  W_m0 = x2,  W_m1 = x3,  W_p = rcc (x0 + x4),
// with c = sqrt(gamma q / a), k1 = 1/(2c), k2 = 1/(2 rcc), k3 = 1/c^2 = a / rcc, rcc = gamma q
// (the forward's face_states evaluates the same values).  As functions of (a, q):
//   dk1 = k1 (da/a - dq/q)/2,  dk2 = -k2 dq/q,  dk3 = k3 (da/a - dq/q),  dc = c (dq/q - da/a)/2.
struct CharK {
    double a, q, c, rcc, k1, k2, k3;
};

__device__ __forceinline__ CharK char_frozen(const double wl[5], const double wr[5], const Params& prm) {
    Frozen f;
    frozen_state<false>(wl, wr, prm, f);
    CharK K;
    K.a = f.P[0];
    // ARITHMETIC: q is the mean pressure.  ROE: the projection depends on the Roe density a and
    // c^2 only, written as q = a c^2 / gamma so that c^2 = gamma q / a and rcc = gamma q hold
    // for both states
    K.q = prm.frozen == JFA_FROZEN_ARITHMETIC ? f.P[4] : f.rcc / prm.gamma;
    K.c = f.c;
    K.rcc = f.rcc;
    K.k1 = 0.5 * f.inv_c;
    K.k3 = f.inv_c * f.inv_c;
    K.k2 = 0.5 * rcp1(f.rcc);
    return K;
}

// characteristic variable k of window cell w (the expressions of face_states)
template <int AXIS>
__device__ __forceinline__ double char_value(int k, const double w[5], const CharK& K) {
    constexpr int v = Ax<AXIS>::v, m0 = Ax<AXIS>::m0, m1 = Ax<AXIS>::m1;
    if (k == 0) return K.k2 * w[4] - K.k1 * w[v];
    if (k == 1) return w[0] - K.k3 * w[4];
    if (k == 2) return w[m0];
    if (k == 3) return w[m1];
    return K.k2 * w[4] + K.k1 * w[v];
}

// The primitive variables characteristic k reads and their coefficients: x_k = c0 w[i0] + c1 w[i1]
// (n = 1 or 2 terms).
template <int AXIS>
__device__ __forceinline__ int char_terms(int k, const CharK& K, int& i0, double& c0, int& i1, double& c1) {
    constexpr int v = Ax<AXIS>::v, m0 = Ax<AXIS>::m0, m1 = Ax<AXIS>::m1;
    if (k == 0) { i0 = 4; c0 = K.k2; i1 = v; c1 = -K.k1; return 2; }
    if (k == 1) { i0 = 0; c0 = 1.0; i1 = 4; c1 = -K.k3; return 2; }
    if (k == 2) { i0 = m0; c0 = 1.0; i1 = m0; c1 = 0.0; return 1; }
    if (k == 3) { i0 = m1; c0 = 1.0; i1 = m1; c1 = 0.0; return 1; }
    i0 = 4; c0 = K.k2; i1 = v; c1 = K.k1;
    return 2;
}

// Adds to gk = (g k1, g k2, g k3) the cotangents of the projection coefficients from the
// window cotangent g6 of characteristic k (cells w6[j][.]).
template <int AXIS>
__device__ __forceinline__ void char_coeff_vjp(int k, const double p6[6], const double u6[6], const double g6[6],
                                               double gk[3]) {
    if (k == 0 || k == 4) {
        const double sg = k == 0 ? -1.0 : 1.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            gk[1] += p6[j] * g6[j];
            gk[0] += sg * u6[j] * g6[j];
        }
    } else if (k == 1) {
#pragma unroll
        for (int j = 0; j < 6; ++j) gk[2] -= p6[j] * g6[j];
    }
}

// Cotangents (g a, g q) of the frozen state from those of the projection coefficients gk and of
// the back projection fz = (g a, g q, g c) (face_state_vjp).
__device__ __forceinline__ void char_frozen_vjp(const CharK& K, const double gk[3], const double fz[3], double& ga,
                                                double& gq) {
    ga = fz[0] + gk[0] * K.k1 / (2.0 * K.a) + gk[2] * K.k3 / K.a - fz[2] * K.c / (2.0 * K.a);
    gq = fz[1] - gk[0] * K.k1 / (2.0 * K.q) - gk[1] * K.k2 / K.q - gk[2] * K.k3 / K.q + fz[2] * K.c / (2.0 * K.q);
}

// ROE averages of the face's two cells (eigendecomposition.py:233-275): a = sqrt(rho_L rho_R),
// u~ = (s_L u_L + s_R u_R) r, H~ = (s_L H_L + s_R H_R) r and
// c^2 = psi + g (psi + a r^2 |u_R - u_L|^2 / 2), psi = (p_L / s_L + p_R / s_R) r, with
// s_K = sqrt(rho_K), r = 1 / (s_L + s_R); the Grueneisen average g = (s_L g + s_R g) r equals
// gamma - 1 for a single ideal gas and is taken as that constant.  Adds to gwl, gwr the VJP for
// the cotangents ga (of a, also where c^2 reads it), gs (of c^2), gu (of u~, may be null) and gH
// (of H~), given u~ (ua) and H~ (Ha).
__device__ __forceinline__ void roe_frozen_vjp(const double wl[5], const double wr[5], double a, double ga, double gs,
                                               const double* gu, const double* ua, double gH, double Ha,
                                               const Params& prm, double gwl[5], double gwr[5]) {
    const double g = prm.gm1;
    const double sL = sqrt(wl[0]), sR = sqrt(wr[0]);
    const double r = 1.0 / (sL + sR);
    double dv[3], dv2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        dv[k] = wr[1 + k] - wl[1 + k];
        dv2 += dv[k] * dv[k];
    }
    const double gpsi = gs * (1.0 + g);
    const double gd = gs * 0.5 * g * a * r * r;  // of |du|^2
    const double gat = ga + gs * 0.5 * g * r * r * dv2;
    double gr = gs * g * a * r * dv2;
    const double tL = wl[4] / sL, tR = wr[4] / sR;  // psi = (tL + tR) r
    gr += gpsi * (tL + tR);
    gwl[4] += gpsi * r / sL;
    gwr[4] += gpsi * r / sR;
    double gsL = -gpsi * r * tL / sL, gsR = -gpsi * r * tR / sR;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        gwl[1 + k] -= 2.0 * gd * dv[k];
        gwr[1 + k] += 2.0 * gd * dv[k];
    }
    if (gu != nullptr) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            gwl[1 + k] += gu[k] * sL * r;
            gwr[1 + k] += gu[k] * sR * r;
            gsL += gu[k] * (wl[1 + k] - ua[k]) * r;
            gsR += gu[k] * (wr[1 + k] - ua[k]) * r;
        }
        // H_K = gamma p_K / ((gamma - 1) rho_K) + |u_K|^2 / 2 (ideal_gas.py:90-110)
        const double hf = prm.gamma * prm.inv_gm1;
        const double* w2[2] = {wl, wr};
        double* g2[2] = {gwl, gwr};
#pragma unroll
        for (int K = 0; K < 2; ++K) {
            const double* w = w2[K];
            const double sK = K == 0 ? sL : sR;
            const double HK = hf * w[4] / w[0] + 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
            const double gHK = gH * sK * r;
            (K == 0 ? gsL : gsR) += gH * (HK - Ha) * r;
            g2[K][4] += gHK * hf / w[0];
            g2[K][0] -= gHK * hf * w[4] / (w[0] * w[0]);
#pragma unroll
            for (int k = 1; k < 4; ++k) g2[K][k] += gHK * w[k];
        }
    }
    gsL -= gr * r * r;
    gsR -= gr * r * r;
    gwl[0] += gsL / (2.0 * sL) + gat * a / (2.0 * wl[0]);
    gwr[0] += gsR / (2.0 * sR) + gat * a / (2.0 * wr[0]);
}

// Cotangents of the face's left / right cell primitives (gwl, gwr, overwritten) from those of
// the CHAR-PRIMITIVE frozen state (g a, g q).  ARITHMETIC: a, q are the means of rho and p.
// ROE: a is the Roe density and q = a c^2 / gamma (roe_frozen_vjp).
__device__ __forceinline__ void frozen_cells_vjp(const double wl[5], const double wr[5], const CharK& K, double ga,
                                                 double gq, const Params& prm, double gwl[5], double gwr[5]) {
#pragma unroll
    for (int i = 0; i < 5; ++i) gwl[i] = gwr[i] = 0.0;
    if (prm.frozen == JFA_FROZEN_ARITHMETIC) {
        gwl[0] = gwr[0] = 0.5 * ga;
        gwl[4] = gwr[4] = 0.5 * gq;
        return;
    }
    const double ig = 1.0 / prm.gamma;
    roe_frozen_vjp(wl, wr, K.a, ga + gq * K.c * K.c * ig, gq * K.a * ig, nullptr, nullptr, 0.0, 0.0, prm, gwl, gwr);
}

// ------------------------------------------------------------------ CHAR-CONSERVATIVE
// Cotangents of the frozen state read by the conservative eigenvectors
// (eigendecomposition.py:595-666), fz[0..9] = g of (rho~, u~, v~, w~, H~, c, c^2, 1/c^2,
// 1/rho~, |u~|^2); the Grueneisen average is a constant of the single ideal gas.
enum { FZ_P0 = 0, FZ_H = 4, FZ_C = 5, FZ_CC = 6, FZ_OC = 7, FZ_OR = 8, FZ_Q2 = 9, FZ_N = 10 };

// right eigenvectors: adds the VJP of R[i][a] for cotangents gR[i][a] = gA[i] xA[a] + gB[i] xB[a]
template <int AXIS>
__device__ __forceinline__ void eigen_r_vjp(const Frozen& f, const double gA[5], const double xA[5],
                                            const double gB[5], const double xB[5], double fz[FZ_N]) {
    constexpr int v = Ax<AXIS>::v, m0 = Ax<AXIS>::m0, m1 = Ax<AXIS>::m1;
    auto G = [&](int i, int a) { return gA[i] * xA[a] + gB[i] * xB[a]; };
    const double* P = f.P;
    const double g = f.g, c = f.c;
    double t;
    t = G(v, 0); fz[v] += t; fz[FZ_C] -= t;                                      // P_v - c
    fz[m0] += G(m0, 0);                                                          // P_m0
    fz[m1] += G(m1, 0);                                                          // P_m1
    t = G(4, 0); fz[FZ_H] += t; fz[v] -= t * c; fz[FZ_C] -= t * P[v];            // H - P_v c
    fz[1] += G(1, v) * g;                                                        // g P_1
    fz[2] += G(2, v) * g;
    fz[3] += G(3, v) * g;
    t = G(4, v); fz[FZ_H] += t * g; fz[FZ_CC] -= t;                              // g H - c^2
    fz[FZ_P0] -= G(m0, m0);                                                      // -rho
    t = G(4, m0); fz[FZ_P0] -= t * P[m0]; fz[m0] -= t * P[0];                    // -rho P_m0
    fz[FZ_P0] += G(m1, m1);                                                      // rho
    t = G(4, m1); fz[FZ_P0] += t * P[m1]; fz[m1] += t * P[0];                    // rho P_m1
    t = G(v, 4); fz[v] += t; fz[FZ_C] += t;                                      // P_v + c
    fz[m0] += G(m0, 4);
    fz[m1] += G(m1, 4);
    t = G(4, 4); fz[FZ_H] += t; fz[v] += t * c; fz[FZ_C] += t * P[v];            // H + P_v c
}

// left eigenvectors: adds the VJP of L[a][b] for cotangents gL[a][b]
template <int AXIS>
__device__ __forceinline__ void eigen_l_vjp(const Frozen& f, const double (&gL)[5][5], double fz[FZ_N]) {
    constexpr int v = Ax<AXIS>::v, m0 = Ax<AXIS>::m0, m1 = Ax<AXIS>::m1;
    const double* P = f.P;
    const double g = f.g, c = f.c, H = f.H, q2 = f.q2;
    const double oc = f.inv_c * f.inv_c, orr = 1.0 / P[0];
    double t;
    // rows 0 and 4: 0.5 oc (g q2 - g H +- (P_v +- c) c), 0.5 oc (-P_v g -+ c), 0.5 oc (-P_m g), 0.5 oc g
#pragma unroll
    for (int a = 0; a < 5; a += 4) {
        const double sg = a == 0 ? 1.0 : -1.0;
        t = gL[a][0];
        fz[FZ_OC] += 0.5 * t * (g * q2 - g * H + sg * (P[v] + sg * c) * c);
        t *= 0.5 * oc;
        fz[FZ_Q2] += t * g;
        fz[FZ_H] -= t * g;
        fz[v] += sg * t * c;
        fz[FZ_C] += t * (sg * P[v] + 2.0 * c);
        t = gL[a][v];
        fz[FZ_OC] += 0.5 * t * (-P[v] * g - sg * c);
        fz[v] -= 0.5 * oc * g * t;
        fz[FZ_C] -= sg * 0.5 * oc * t;
        t = gL[a][m0];
        fz[FZ_OC] -= 0.5 * t * P[m0] * g;
        fz[m0] -= 0.5 * oc * g * t;
        t = gL[a][m1];
        fz[FZ_OC] -= 0.5 * t * P[m1] * g;
        fz[m1] -= 0.5 * oc * g * t;
        fz[FZ_OC] += 0.5 * g * gL[a][4];
    }
    // row v: oc (H - q2), P_k oc, -oc
    t = gL[v][0];
    fz[FZ_OC] += t * (H - q2);
    fz[FZ_H] += t * oc;
    fz[FZ_Q2] -= t * oc;
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        fz[k] += gL[v][k] * oc;
        fz[FZ_OC] += gL[v][k] * P[k];
    }
    fz[FZ_OC] -= gL[v][4];
    // rows m0, m1: +-P_m / rho, -+1 / rho
    fz[m0] += gL[m0][0] * orr;
    fz[FZ_OR] += gL[m0][0] * P[m0] - gL[m0][m0];
    fz[m1] -= gL[m1][0] * orr;
    fz[FZ_OR] += -gL[m1][0] * P[m1] + gL[m1][m1];
}

// Cotangents of the face's two cells (gwl, gwr, overwritten) from the frozen-state cotangents
// fz: 1/c^2, 1/rho~, c and |u~|^2 first folded into c^2, rho~ and u~, then ARITHMETIC means
// (c^2 = gamma p~ / rho~, H~ = gamma p~ / ((gamma - 1) rho~) + |u~|^2 / 2) or ROE averages.
__device__ __forceinline__ void char_cons_frozen_vjp(const double wl[5], const double wr[5], const Frozen& f,
                                                     double fz[FZ_N], const Params& prm, double gwl[5],
                                                     double gwr[5]) {
    const double oc = f.inv_c * f.inv_c, orr = 1.0 / f.P[0];
    double gcc = fz[FZ_CC] - fz[FZ_OC] * oc * oc + fz[FZ_C] / (2.0 * f.c);
    double gP0 = fz[FZ_P0] - fz[FZ_OR] * orr * orr;
    double gu[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) gu[k] = fz[1 + k] + 2.0 * f.P[1 + k] * fz[FZ_Q2];
#pragma unroll
    for (int i = 0; i < 5; ++i) gwl[i] = gwr[i] = 0.0;
    if (prm.frozen == JFA_FROZEN_ARITHMETIC) {
        const double gH = fz[FZ_H];
        const double hf = prm.gamma * prm.inv_gm1;
        double gP4 = gcc * prm.gamma * orr + gH * hf * orr;
        gP0 -= gcc * f.cc * orr + gH * (f.H - 0.5 * f.q2) * orr;
#pragma unroll
        for (int k = 0; k < 3; ++k) gu[k] += gH * f.P[1 + k];
        gwl[0] = gwr[0] = 0.5 * gP0;
        gwl[4] = gwr[4] = 0.5 * gP4;
#pragma unroll
        for (int k = 0; k < 3; ++k) gwl[1 + k] = gwr[1 + k] = 0.5 * gu[k];
        return;
    }
    const double ua[3] = {f.P[1], f.P[2], f.P[3]};
    roe_frozen_vjp(wl, wr, f.P[0], gP0, gcc, gu, ua, fz[FZ_H], f.H, prm, gwl, gwr);
}

// VJP of one face flux, first part: cotangents of the reconstructed face states.  PRIMITIVE:
// gl, gr are those of WL, WR.  CHAR-PRIMITIVE (ARITHMETIC frozen state): gl, gr are those of
// the reconstructed characteristics and fz = (g a, g q, g c) the frozen state's share through
// the back projection (W_rho depends on a, W_n on c, W_p on rcc = gamma q).
template <int AXIS, int STENCIL, int RIEMANN, int VAR>
__device__ __forceinline__ void face_state_vjp(const double (&win)[6][5], const double wl[5], const double wr[5],
                                               const double mu[5], const Params& prm, double gl[5], double gr[5],
                                               double fz[FZ_N]) {
    double WL[5], WR[5], UL[5], UR[5];
    face_states<AXIS, VAR, STENCIL>(win, wl, wr, prm, WL, WR, UL, UR);
    double gWL[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gWR[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if constexpr (RIEMANN == JFA_HLLC) hllc_vjp<AXIS>(WL, WR, UL, UR, mu, prm, gWL, gWR);
    else if constexpr (RIEMANN == JFA_HLL) hll_vjp<AXIS>(WL, WR, UL, UR, mu, prm, gWL, gWR);
    else if constexpr (RIEMANN == JFA_HLLC_LM) hllc_lm_vjp<AXIS>(WL, WR, UL, UR, mu, prm, gWL, gWR);
    else if constexpr (RIEMANN == JFA_AUSMP) ausmp_vjp<AXIS>(WL, WR, UL, UR, mu, prm, gWL, gWR);
    else rusanov_vjp<AXIS>(WL, WR, UL, UR, mu, prm, gWL, gWR);
    if constexpr (VAR == JFA_PRIMITIVE) {
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            gl[i] = gWL[i];
            gr[i] = gWR[i];
        }
        fz[0] = fz[1] = fz[2] = 0.0;
    } else if constexpr (VAR == JFA_CONSERVATIVE) {
        // W_L = prim(U_L) of the reconstructed conservatives (high_order_godunov.py:282-296);
        // gWL already holds the flux's whole dependence on the face state (its U_L = cons(W_L))
        prim_vjp(WL, gWL, prm, gl);
        prim_vjp(WR, gWR, prm, gr);
        fz[0] = fz[1] = fz[2] = 0.0;
    } else if constexpr (VAR == JFA_CHAR_CONSERVATIVE) {
        // U_L = R x_L of the reconstructed characteristics x_L (high_order_godunov.py:404-417),
        // W_L = prim(U_L): gl = R^T gU_L, and the right eigenvectors' share of the frozen state.
        // x_L = L U_L (L = R^-1 to rounding) stands in for the forward's reconstructed values.
        double gUL[5], gUR[5];
        prim_vjp(WL, gWL, prm, gUL);
        prim_vjp(WR, gWR, prm, gUR);
        Frozen f;
        frozen_state<true>(wl, wr, prm, f);
        double Lm[5][5], Rm[5][5];
        conservative_eigenvectors<AXIS>(f, Lm, Rm);
        double xL[5], xR[5];
#pragma unroll
        for (int a = 0; a < 5; ++a) {
            xL[a] = xR[a] = gl[a] = gr[a] = 0.0;
#pragma unroll
            for (int b = 0; b < 5; ++b) {
                xL[a] += Lm[a][b] * UL[b];
                xR[a] += Lm[a][b] * UR[b];
                gl[a] += Rm[b][a] * gUL[b];
                gr[a] += Rm[b][a] * gUR[b];
            }
        }
#pragma unroll
        for (int i = 0; i < FZ_N; ++i) fz[i] = 0.0;
        eigen_r_vjp<AXIS>(f, gUL, xL, gUR, xR, fz);
    } else {
        constexpr int v = Ax<AXIS>::v, m0 = Ax<AXIS>::m0, m1 = Ax<AXIS>::m1;
        const CharK K = char_frozen(win[2], win[3], prm);
        // x0 + x4 = W_p / rcc, x4 - x0 = W_n / c
        gl[0] = K.a * gWL[0] - K.c * gWL[v] + K.rcc * gWL[4];
        gl[4] = K.a * gWL[0] + K.c * gWL[v] + K.rcc * gWL[4];
        gl[1] = gWL[0];
        gl[2] = gWL[m0];
        gl[3] = gWL[m1];
        gr[0] = K.a * gWR[0] - K.c * gWR[v] + K.rcc * gWR[4];
        gr[4] = K.a * gWR[0] + K.c * gWR[v] + K.rcc * gWR[4];
        gr[1] = gWR[0];
        gr[2] = gWR[m0];
        gr[3] = gWR[m1];
        const double sL = WL[4] / K.rcc, sR = WR[4] / K.rcc;
        fz[0] = sL * gWL[0] + sR * gWR[0];
        fz[1] = prm.gamma * (sL * gWL[4] + sR * gWR[4]);
        fz[2] = (WL[v] / K.c) * gWL[v] + (WR[v] / K.c) * gWR[v];
    }
}

// Second part, one variable at a time: cotangents of the six window cells x[0..5] of that
// variable (the left state reads cells -3..1 = window 0..4, the right state cells 2..-2 =
// window 5..1).
// WENO3 one side (weno3 of the forward, weno3_base.py:20-45, weno3_js.py / weno3_z.py):
// v = u_i + (e0 dm + e1 dp) / (2 (e0 + e1)) with dm = u_i - u_im, dp = u_ip - u_i and the
// weights multiplied through, e0 = (1/3)(b1^2 + eps), e1 = (2/3)(b0^2 + eps) (JS) or
// e0 = (1/3)(q0 + tau) q1, e1 = (2/3)(q1 + tau) q0, q_k = b_k + eps, tau = |b0 - b1| (Z);
// b0 = dm^2, b1 = dp^2.  Adds the VJP for cotangent g.
template <bool Z>
__device__ __forceinline__ void weno3_vjp(double um, double ui, double up, double eps, double g, double& gum,
                                          double& gui, double& gup) {
    const double dm = ui - um, dp = up - ui;
    const double b0 = dm * dm, b1 = dp * dp;
    double e0, e1, q0 = 0.0, q1 = 0.0, tau = 0.0;
    if constexpr (Z) {
        tau = fabs(b0 - b1);
        q0 = b0 + eps;
        q1 = b1 + eps;
        e0 = (1.0 / 3.0) * ((q0 + tau) * q1);
        e1 = (2.0 / 3.0) * ((q1 + tau) * q0);
    } else {
        e0 = (1.0 / 3.0) * (b1 * b1 + eps);
        e1 = (2.0 / 3.0) * (b0 * b0 + eps);
    }
    const double S = e0 + e1, N = e0 * dm + e1 * dp;
    const double gN = 0.5 * g / S, gS = -0.5 * g * N / (S * S);
    const double ge0 = gN * dm + gS, ge1 = gN * dp + gS;
    double gdm = gN * e0, gdp = gN * e1, gb0, gb1;
    if constexpr (Z) {
        const double gq0 = (1.0 / 3.0) * q1 * ge0 + (2.0 / 3.0) * (q1 + tau) * ge1;
        const double gq1 = (1.0 / 3.0) * (q0 + tau) * ge0 + (2.0 / 3.0) * q0 * ge1;
        const double gt = (1.0 / 3.0) * q1 * ge0 + (2.0 / 3.0) * q0 * ge1;
        const double sg = b0 > b1 ? 1.0 : (b0 < b1 ? -1.0 : 0.0);  // d|x|/dx = 0 at 0
        gb0 = gq0 + gt * sg;
        gb1 = gq1 - gt * sg;
    } else {
        gb0 = ge1 * (2.0 / 3.0) * 2.0 * b0;
        gb1 = ge0 * (1.0 / 3.0) * 2.0 * b1;
    }
    gdm += 2.0 * dm * gb0;
    gdp += 2.0 * dp * gb1;
    gui += g + gdm - gdp;
    gum -= gdm;
    gup += gdp;
}

// TENO5 / TENO5-A one side (teno5 of the forward, teno5.py:32-73, teno5_a.py:42-99): the
// weights w_k are d_k or 0 by the sharp cutoff, constant where the cutoff does not switch
// (jax's derivative of the comparison is zero), so v = c + sum_k w_k del_k / (sum w + eps) is
// linear in the stencil: del = ((5 d1 - 2 d0)/6, (d1 + 2 d2)/6, (4 d2 - d3)/6).
template <bool A>
__device__ __forceinline__ void teno5_vjp(double a, double b, double c, double d, double e, const Params& prm,
                                          double g, double& ga, double& gb, double& gc, double& gd, double& ge) {
    const double d0 = b - a, d1 = c - b, d2 = d - c, d3 = e - d;
    double w0, w1, w2;
    teno5_weights<A>(d0, d1, d2, d3, prm, w0, w1, w2);
    const double r = g / (w0 + w1 + w2 + prm.eps);
    const double g0 = w0 * r, g1 = w1 * r, g2 = w2 * r;  // of del_k
    const double gd0 = -(2.0 / 6.0) * g0, gd1 = (5.0 / 6.0) * g0 + (1.0 / 6.0) * g1;
    const double gd2 = (2.0 / 6.0) * g1 + (4.0 / 6.0) * g2, gd3 = -(1.0 / 6.0) * g2;
    ga -= gd0;
    gb += gd0 - gd1;
    gc += g + gd1 - gd2;
    gd += gd2 - gd3;
    ge += gd3;
}

// Second part of a face VJP, one variable at a time (see above); STENCIL = kStencilExt selects
// WENO1, WENO3-JS/-Z or TENO5 / TENO5-A at run time (prm.stencil), as the forward does.
template <int STENCIL>
__device__ __forceinline__ void window_vjp_var(const double x[6], double gl, double gr, const Params& prm,
                                               double out[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k] = 0.0;
    if constexpr (STENCIL == kStencilExt) {
        switch (prm.stencil) {
            case JFA_WENO1:  // the adjacent cells
                out[2] += gl;
                out[3] += gr;
                break;
            case JFA_WENO3_JS:  // cells -2..0 and 1..-1
                weno3_vjp<false>(x[1], x[2], x[3], prm.eps, gl, out[1], out[2], out[3]);
                weno3_vjp<false>(x[4], x[3], x[2], prm.eps, gr, out[4], out[3], out[2]);
                break;
            case JFA_WENO3_Z:
                weno3_vjp<true>(x[1], x[2], x[3], prm.eps, gl, out[1], out[2], out[3]);
                weno3_vjp<true>(x[4], x[3], x[2], prm.eps, gr, out[4], out[3], out[2]);
                break;
            case JFA_TENO5:
                teno5_vjp<false>(x[0], x[1], x[2], x[3], x[4], prm, gl, out[0], out[1], out[2], out[3], out[4]);
                teno5_vjp<false>(x[5], x[4], x[3], x[2], x[1], prm, gr, out[5], out[4], out[3], out[2], out[1]);
                break;
            default:  // TENO5-A (the entry point admits no other)
                teno5_vjp<true>(x[0], x[1], x[2], x[3], x[4], prm, gl, out[0], out[1], out[2], out[3], out[4]);
                teno5_vjp<true>(x[5], x[4], x[3], x[2], x[1], prm, gr, out[5], out[4], out[3], out[2], out[1]);
                break;
        }
    } else {
        weno5_vjp<STENCIL>(x[0], x[1], x[2], x[3], x[4], prm.eps, gl, out[0], out[1], out[2], out[3], out[4]);
        weno5_vjp<STENCIL>(x[5], x[4], x[3], x[2], x[1], prm.eps, gr, out[5], out[4], out[3], out[2], out[1]);
    }
}

// Face f (right cell at buffer index br, stride sa along the face normal): cotangents of the
// face states from a window loaded here, so the window is not held across the Riemann VJP.
template <int AXIS, int STENCIL, int RIEMANN, int VAR>
__device__ __forceinline__ void face_state_vjp_at(const double* __restrict__ W, int64_t br, int64_t sa, int64_t fs,
                                                  const double mu[5], const Params& prm, double gl[5], double gr[5],
                                                  double fz[FZ_N]) {
    double win[6][5], wl[5], wr[5];
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int i = 0; i < 5; ++i) win[k][i] = ld(W, i * fs + br + (int64_t)(k - 3) * sa);
#pragma unroll
    for (int i = 0; i < 5; ++i) {  // the face's two cells (read by the CHAR-CONSERVATIVE frozen state)
        wl[i] = win[2][i];
        wr[i] = win[3][i];
    }
    if constexpr (VAR == JFA_CONSERVATIVE || VAR == JFA_CHAR_CONSERVATIVE) {  // conservative windows
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double u[5];
            cons_from_prim(win[k], prm, u);
#pragma unroll
            for (int i = 0; i < 5; ++i) win[k][i] = u[i];
        }
    }
    face_state_vjp<AXIS, STENCIL, RIEMANN, VAR>(win, wl, wr, mu, prm, gl, gr, fz);
}

// CHAR-PRIMITIVE, second part for characteristic k of the face: its window (projected from the
// cells, reloaded; they are in cache), the WENO VJP of its two reconstructions into g6, and its
// share of the coefficient cotangents gk.
template <int AXIS, int STENCIL>
__device__ __forceinline__ void char_window_vjp(const double* __restrict__ W, int64_t br, int64_t sa, int64_t fs,
                                                int k, const CharK& K, double gl, double gr, const Params& prm,
                                                double g6[6], double gk[3]) {
    constexpr int v = Ax<AXIS>::v;
    double x6[6], p6[6], u6[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double w[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) w[i] = ld(W, i * fs + br + (int64_t)(j - 3) * sa);
        x6[j] = char_value<AXIS>(k, w, K);
        p6[j] = w[4];
        u6[j] = w[v];
    }
    window_vjp_var<STENCIL>(x6, gl, gr, prm, g6);
    char_coeff_vjp<AXIS>(k, p6, u6, g6, gk);
}

// The six window cells of variable i of that face (reloaded; they are in cache).
__device__ __forceinline__ void load_var6(const double* __restrict__ W, int64_t br, int64_t sa, int64_t fs, int i,
                                          double x[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) x[k] = ld(W, i * fs + br + (int64_t)(k - 3) * sa);
}

// CONSERVATIVE, second part for conserved variable i of the face: its window U_i of the six
// cells (cons of the primitive window, as the forward's U buffer holds it up to rounding), the
// WENO VJP of its two reconstructions into g6, and d U_i / d W of each cell in dU[j][.], so the
// caller adds g6[j] dU[j][t] to the cotangent of W_j[t] (U = cons(W), equation_manager.py:93-101).
// d U_i / d W_t is identically zero unless this holds
__host__ __device__ constexpr bool cons_dep(int i, int t) { return i == 4 || t == 0 || t == i; }

template <int STENCIL>
__device__ __forceinline__ void cons_window_vjp(const double* __restrict__ W, int64_t br, int64_t sa, int64_t fs,
                                                int i, const Params& prm, double gl, double gr, double g6[6],
                                                double dU[6][5]) {
    double x6[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double w[5], u[5];
#pragma unroll
        for (int t = 0; t < 5; ++t) w[t] = ld(W, t * fs + br + (int64_t)(j - 3) * sa);
        cons_from_prim(w, prm, u);
        x6[j] = u[i];
        // d U_i / d W: rho -> (1, 0, 0, 0, 0); rho u_k -> (u_k, rho e_k); E -> (|u|^2/2, rho u, 1/(gamma-1))
#pragma unroll
        for (int t = 0; t < 5; ++t) dU[j][t] = 0.0;
        if (i == 0) {
            dU[j][0] = 1.0;
        } else if (i < 4) {
            dU[j][0] = w[i];
            dU[j][i] = w[0];
        } else {
            dU[j][0] = 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
            dU[j][1] = w[0] * w[1];
            dU[j][2] = w[0] * w[2];
            dU[j][3] = w[0] * w[3];
            dU[j][4] = prm.inv_gm1;
        }
    }
    window_vjp_var<STENCIL>(x6, gl, gr, prm, g6);
}

// CHAR-CONSERVATIVE, second part: the window cotangents of all five characteristics
// x_a = sum_b L[a][b] U_j[b] of the six cells, taken back to their conservatives in gU6 (then
// to the primitives by the caller, U = cons(W)), and the left eigenvectors' share of the frozen
// state, added to fz.  f, Lm: the face's frozen state and left eigenvectors.
template <int AXIS, int STENCIL>
__device__ __forceinline__ void char_cons_window_vjp(const double* __restrict__ W, int64_t br, int64_t sa,
                                                     int64_t fs, const Frozen& f, const double (&Lm)[5][5],
                                                     const double gl[5], const double gr[5], const Params& prm,
                                                     double (&gU6)[6][5], double fz[FZ_N]) {
    double u6[6][5];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double w[5];
#pragma unroll
        for (int t = 0; t < 5; ++t) w[t] = ld(W, t * fs + br + (int64_t)(j - 3) * sa);
        cons_from_prim(w, prm, u6[j]);
#pragma unroll
        for (int t = 0; t < 5; ++t) gU6[j][t] = 0.0;
    }
    double gL[5][5];
#pragma unroll
    for (int a = 0; a < 5; ++a) {
        double x6[6], g6[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double acc = Lm[a][0] * u6[j][0];
#pragma unroll
            for (int b = 1; b < 5; ++b) acc += Lm[a][b] * u6[j][b];
            x6[j] = acc;
        }
        window_vjp_var<STENCIL>(x6, gl[a], gr[a], prm, g6);
#pragma unroll
        for (int b = 0; b < 5; ++b) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                acc += g6[j] * u6[j][b];
                gU6[j][b] += Lm[a][b] * g6[j];
            }
            gL[a][b] = acc;
        }
    }
    eigen_l_vjp<AXIS>(f, gL, fz);
}

// ------------------------------------------------------------------ dissipative fluxes
// d law / dT at T, where value = law_value(T, l): POWER n value / T; SUTHERLAND
// value (3/2 / T - 1 / (T + C)).
__device__ __forceinline__ double law_slope(double T, double value, const Law& l) {
    if (l.kind == JFA_LAW_SUTHERLAND) return value * (1.5 / T - 1.0 / (T + l.param));
    return l.param == 0.0 ? 0.0 : l.param * value / T;
}

// VJP of the dissipative flux (jfa_dissipative.h) of the face whose right cell is cr, for the
// cotangent m of the face flux.  Adds the cotangents of the left and right cells' primitives to
// gl, gr and returns Gt[j][k], j-th axis i != AXIS (ascending), velocity component k: the face
// sends +Gt to u_k at L + e_i and R + e_i and -Gt to u_k at L - e_i and R - e_i (the transverse
// central derivatives).  The clamp of T_f <= 0 has zero derivative.
template <int AXIS>
__device__ __forceinline__ void dissipative_face_vjp(const double* __restrict__ W, const Geo& G, const Params& prm,
                                                     const int cr[3], const double m[5], double gl[5], double gr[5],
                                                     double Gt[2][3]) {
    DissState s;
    dissipative_state<double, AXIS>(W, G, prm, cr, s);
    double mub = 0.0, Tfb = 0.0, TLb = 0.0, TRb = 0.0;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) Gt[j][k] = 0.0;
    if (prm.dissip & 1) {
        // forward: tau_k = mu (g[A][k] + g[k][A]) (+ mu2 div for k = A), D[1+k] = -tau_k,
        // D[4] = -sum_k tau_k u_f,k
        const double mu2 = prm.bulk - 2.0 / 3.0 * s.mu;
        double tau[3], uf[3], div = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            tau[k] = G.active[k] ? s.mu * (s.g[AXIS][k] + s.g[k][AXIS]) : 0.0;
            uf[k] = 0.5 * (s.wl[1 + k] + s.wr[1 + k]);
            if (G.active[k]) div = div + s.g[k][k];
        }
        tau[AXIS] = tau[AXIS] + mu2 * div;
        double tb[3], gb[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            tb[k] = G.active[k] ? -m[1 + k] - m[4] * uf[k] : 0.0;
            const double ufb = G.active[k] ? -m[4] * tau[k] : 0.0;
            gl[1 + k] += 0.5 * ufb;
            gr[1 + k] += 0.5 * ufb;
#pragma unroll
            for (int i = 0; i < 3; ++i) gb[k][i] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (!G.active[k]) continue;
            mub += tb[k] * (s.g[AXIS][k] + s.g[k][AXIS]);
            gb[AXIS][k] += tb[k] * s.mu;
            gb[k][AXIS] += tb[k] * s.mu;
            gb[k][k] += tb[AXIS] * mu2;
        }
        mub += -2.0 / 3.0 * (tb[AXIS] * div);
        // g[k][AXIS] = (u_R,k - u_L,k) / dx_AXIS
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            gr[1 + k] += gb[k][AXIS] * G.one_dx[AXIS];
            gl[1 + k] -= gb[k][AXIS] * G.one_dx[AXIS];
        }
        // g[k][i] = (dL + dR) / 2, dX = (u(X + e_i) - u(X - e_i)) / (2 dx_i)
        int j = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (i == AXIS) continue;
            if (G.active[i])
#pragma unroll
                for (int k = 0; k < 3; ++k) Gt[j][k] = gb[k][i] * (0.25 * G.one_dx[i]);
            ++j;
        }
    }
    if (prm.dissip & 2) {
        // q = -kappa(T_f) (T_R - T_L) / dx
        const double kappa = conductivity(s.Tf, s.mu, prm);
        const double qb = m[4];
        const double kb = -qb * ((s.TR - s.TL) * G.one_dx[AXIS]);
        TRb += -qb * kappa * G.one_dx[AXIS];
        TLb += qb * kappa * G.one_dx[AXIS];
        if (prm.kappa.kind == JFA_LAW_PRANDTL) mub += prm.kappa_fac * kb;
        else Tfb += kb * law_slope(s.Tf, kappa, prm.kappa);
    }
    Tfb += mub * law_slope(s.Tf, s.mu, prm.mu);
    if (!s.clamped) {
        TLb += 0.5 * Tfb;
        TRb += 0.5 * Tfb;
    }
    // T = p / (rho R)
    gl[4] += TLb / (s.wl[0] * prm.R);
    gl[0] -= TLb * s.TL / s.wl[0];
    gr[4] += TRb / (s.wr[0] * prm.R);
    gr[0] -= TRb * s.TR / s.wr[0];
}

// ------------------------------------------------------------------ kernels
constexpr int kAdjBlock = 256;

// cotangent of the face between cells f-1 and f along an axis of n cells: lam outside the
// interior is zero (those cells are not updated by the stage)
__device__ __forceinline__ void face_mu(const double* __restrict__ lam, int64_t b_right, int64_t sa, int64_t fs,
                                        int f, int n, double scale, double mu[5]) {
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const double lr = f < n ? ld(lam, i * fs + b_right) : 0.0;
        const double ll = f > 0 ? ld(lam, i * fs + b_right - sa) : 0.0;
        mu[i] = scale * (lr - ll);
    }
}

// Flux-splitting scheme (FluxSplittingScheme.compute_flux_xi, flux_splitting_scheme.py:65-111;
// flux_split_face of the forward): F = R (rec_L(f+) + rec_R(f-)), f+- = (L F_k +- g L U_k)/2 over
// the window cells k, L, R the conservative eigenvectors at the frozen state of the face's two
// cells, g the eigenvalue magnitudes (ROE at the frozen state, CLLF / LLF from the two cells).
// Writes to gW6 the cotangents of the six window cells' primitives for face cotangent mu.
template <int AXIS, int STENCIL>
__device__ __forceinline__ void flux_split_vjp(const double* __restrict__ W, int64_t br, int64_t sa, int64_t fs,
                                               const double mu[5], const Params& prm, double (&gW6)[6][5]) {
    constexpr int v = Ax<AXIS>::v;
    double win[6][5];
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            win[k][i] = ld(W, i * fs + br + (int64_t)(k - 3) * sa);
            gW6[k][i] = 0.0;
        }
    Frozen f;
    frozen_state<true>(win[2], win[3], prm, f);
    double Lm[5][5], Rm[5][5];
    conservative_eigenvectors<AXIS>(f, Lm, Rm);
    const double uL = win[2][v], uR = win[3][v];
    const double cL = sqrt(prm.gamma * win[2][4] / win[2][0]), cR = sqrt(prm.gamma * win[3][4] / win[3][0]);
    double g[5];
    if (prm.riemann == JFA_FS_ROE) {
        g[0] = fabs(f.P[v] - f.c);
        g[1] = g[2] = g[3] = fabs(f.P[v]);
        g[4] = fabs(f.P[v] + f.c);
    } else if (prm.riemann == JFA_FS_CLLF) {
        g[0] = fmax(fabs(uL - cL), fabs(uR - cR));
        g[1] = g[2] = g[3] = fmax(fabs(uL), fabs(uR));
        g[4] = fmax(fabs(uL + cL), fabs(uR + cR));
    } else {
        g[0] = g[1] = g[2] = g[3] = g[4] = fmax(fabs(uL) + cL, fabs(uR) + cR);
    }
    double Uk[6][5], Fk[6][5], gU[6][5], gF[6][5];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        cons_from_prim(win[k], prm, Uk[k]);
        physical_flux<AXIS>(win[k], Uk[k], Fk[k]);
#pragma unroll
        for (int i = 0; i < 5; ++i) gU[k][i] = gF[k][i] = 0.0;
    }
    double fz[FZ_N], gL[5][5], gg[5], fcv[5];
#pragma unroll
    for (int i = 0; i < FZ_N; ++i) fz[i] = 0.0;
#pragma unroll
    for (int a = 0; a < 5; ++a) {
        double gfc = 0.0;
#pragma unroll
        for (int i = 0; i < 5; ++i) gfc += Rm[i][a] * mu[i];
        double cu[6], cf[6], pos[6], neg[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            cu[k] = cf[k] = 0.0;
#pragma unroll
            for (int b = 0; b < 5; ++b) {
                cu[k] += Lm[a][b] * Uk[k][b];
                cf[k] += Lm[a][b] * Fk[k][b];
            }
            pos[k] = 0.5 * (cf[k] + g[a] * cu[k]);
            neg[k] = 0.5 * (cf[k] - g[a] * cu[k]);
        }
        fcv[a] = recon_side<STENCIL>(pos[0], pos[1], pos[2], pos[3], pos[4], pos[5], prm, prm.dx2[AXIS]) +
                 recon_side<STENCIL>(neg[5], neg[4], neg[3], neg[2], neg[1], neg[0], prm, prm.dx2[AXIS]);
        double gp[6], gn[6];
        window_vjp_var<STENCIL>(pos, gfc, 0.0, prm, gp);  // f+ from the left
        window_vjp_var<STENCIL>(neg, 0.0, gfc, prm, gn);  // f- from the right
        gg[a] = 0.0;
#pragma unroll
        for (int b = 0; b < 5; ++b) gL[a][b] = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double gcf = 0.5 * (gp[k] + gn[k]), gcu = 0.5 * g[a] * (gp[k] - gn[k]);
            gg[a] += 0.5 * cu[k] * (gp[k] - gn[k]);
#pragma unroll
            for (int b = 0; b < 5; ++b) {
                gU[k][b] += Lm[a][b] * gcu;
                gF[k][b] += Lm[a][b] * gcf;
                gL[a][b] += gcu * Uk[k][b] + gcf * Fk[k][b];
            }
        }
    }
    eigen_l_vjp<AXIS>(f, gL, fz);
    const double z5[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    eigen_r_vjp<AXIS>(f, mu, fcv, z5, z5, fz);  // gR[i][a] = mu_i fc_a
    if (prm.riemann == JFA_FS_ROE) {
        const double s0 = sgn(f.P[v] - f.c), s4 = sgn(f.P[v] + f.c), sv = sgn(f.P[v]);
        fz[v] += gg[0] * s0 + (gg[1] + gg[2] + gg[3]) * sv + gg[4] * s4;
        fz[FZ_C] += gg[4] * s4 - gg[0] * s0;
    } else {
        double guL = 0.0, guR = 0.0, gcL = 0.0, gcR = 0.0, x1 = 0.0, x2 = 0.0;
        if (prm.riemann == JFA_FS_CLLF) {
            minmax_vjp(false, fabs(uL - cL), fabs(uR - cR), gg[0], x1, x2);
            guL += x1 * sgn(uL - cL);
            gcL -= x1 * sgn(uL - cL);
            guR += x2 * sgn(uR - cR);
            gcR -= x2 * sgn(uR - cR);
            x1 = x2 = 0.0;
            minmax_vjp(false, fabs(uL), fabs(uR), gg[1] + gg[2] + gg[3], x1, x2);
            guL += x1 * sgn(uL);
            guR += x2 * sgn(uR);
            x1 = x2 = 0.0;
            minmax_vjp(false, fabs(uL + cL), fabs(uR + cR), gg[4], x1, x2);
            guL += x1 * sgn(uL + cL);
            gcL += x1 * sgn(uL + cL);
            guR += x2 * sgn(uR + cR);
            gcR += x2 * sgn(uR + cR);
        } else {
            minmax_vjp(false, fabs(uL) + cL, fabs(uR) + cR, gg[0] + gg[1] + gg[2] + gg[3] + gg[4], x1, x2);
            guL += x1 * sgn(uL);
            gcL += x1;
            guR += x2 * sgn(uR);
            gcR += x2;
        }
        gW6[2][v] += guL;
        gW6[3][v] += guR;
        sound_speed_vjp(cL, win[2][0], prm.gamma, gcL, gW6[2][0], gW6[2][4]);
        sound_speed_vjp(cR, win[3][0], prm.gamma, gcR, gW6[3][0], gW6[3][4]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        physical_flux_vjp<AXIS>(win[k], Uk[k], gF[k], gW6[k], gU[k]);
        cons_vjp(win[k], gU[k], prm, gW6[k]);
    }
    double gwl[5], gwr[5];
    char_cons_frozen_vjp(win[2], win[3], f, fz, prm, gwl, gwr);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        gW6[2][i] += gwl[i];
        gW6[3][i] += gwr[i];
    }
}

__device__ __forceinline__ void add_cell(double* __restrict__ gW, int64_t fs, int64_t b, const double g[5]) {
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        JFA_ACCESS(gW + i * fs + b, sizeof(double));
        gW[i * fs + b] += g[i];
    }
}

// Marching transpose sweep along A (not the innermost axis IN).  Chunk z of length L owns the
// cells [-3 + zL, min(-3 + (z+1)L, N+3)) of its column (halo cells included) and evaluates
// the faces f in [o0-2, o1+2] (clipped to 0..N) whose windows touch them; each face adds its
// window cotangents to the owned cells in place (the six contributions a cell receives come
// from consecutive faces of the same thread, so they hit cache and never race).
template <int A, int IN, int STENCIL, int RIEMANN, int VAR>
__global__ __launch_bounds__(kAdjBlock) void adj_march_kernel(const double* __restrict__ W,
                                                              const double* __restrict__ lam, Geo G, Params prm,
                                                              double ts, int L, double* __restrict__ gW) {
    constexpr int O = 3 - A - IN;
    const int iin = blockIdx.x * blockDim.x + threadIdx.x;
    if (iin >= G.n[IN]) return;  // no barriers in this kernel
    const int io = blockIdx.y;
    const int N = G.n[A];
    const int o0 = -3 + (int)blockIdx.z * L;
    const int o1 = min(o0 + L, N + 3);
    const int fa = max(o0 - 2, 0), fb = min(o1 + 2, N);
    const int64_t sa = G.s[A], fs = G.fs;
    const int64_t bcol = (int64_t)(io + G.off[O]) * G.s[O] + (int64_t)(iin + G.off[IN]) * G.s[IN] +
                         (int64_t)G.off[A] * sa;  // cell 0 of the column along A
    const double scale = ts * G.one_dx[A];
    for (int f = fa; f <= fb; ++f) {
        int64_t br = bcol + (int64_t)f * sa;
        JFA_OPAQUE(br);  // keeps per-tap addresses from becoming loop-carried registers
        double mu[5], gl[5], gr[5], fz[FZ_N];
        face_mu(lam, br, sa, fs, f, N, scale, mu);
        if constexpr (RIEMANN == kFluxSplit) {
            double gW6[6][5];
            flux_split_vjp<A, STENCIL>(W, br, sa, fs, mu, prm, gW6);
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int cell = f - 3 + k;
                if (cell >= o0 && cell < o1) {
#pragma unroll
                    for (int t = 0; t < 5; ++t) {
                        double* p = gW + t * fs + br + (int64_t)(k - 3) * sa;
                        JFA_ACCESS(p, sizeof(double));
                        *p += gW6[k][t];
                    }
                }
            }
            continue;
        }
        face_state_vjp_at<A, STENCIL, RIEMANN, VAR>(W, br, sa, fs, mu, prm, gl, gr, fz);
        if constexpr (VAR == JFA_PRIMITIVE) {
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                double x6[6], o6[6];
                load_var6(W, br, sa, fs, i, x6);
                window_vjp_var<STENCIL>(x6, gl[i], gr[i], prm, o6);
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const int cell = f - 3 + k;
                    if (cell >= o0 && cell < o1) {
                        double* p = gW + i * fs + br + (int64_t)(k - 3) * sa;
                        JFA_ACCESS(p, sizeof(double));
                        *p += o6[k];
                    }
                }
            }
        } else if constexpr (VAR == JFA_CHAR_CONSERVATIVE) {
            double wl[5], wr[5];
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                wl[i] = ld(W, i * fs + br - sa);
                wr[i] = ld(W, i * fs + br);
            }
            Frozen fr;
            frozen_state<true>(wl, wr, prm, fr);
            double Lm[5][5], Rm[5][5];
            conservative_eigenvectors<A>(fr, Lm, Rm);
            double gU6[6][5];
            char_cons_window_vjp<A, STENCIL>(W, br, sa, fs, fr, Lm, gl, gr, prm, gU6, fz);
            double gwf[2][5];
            char_cons_frozen_vjp(wl, wr, fr, fz, prm, gwf[0], gwf[1]);
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int cell = f - 3 + k;
                if (cell >= o0 && cell < o1) {
                    double w[5], g[5];
#pragma unroll
                    for (int t = 0; t < 5; ++t) {
                        w[t] = ld(W, t * fs + br + (int64_t)(k - 3) * sa);
                        g[t] = (k == 2 || k == 3) ? gwf[k - 2][t] : 0.0;
                    }
                    cons_vjp(w, gU6[k], prm, g);
#pragma unroll
                    for (int t = 0; t < 5; ++t) {
                        double* p = gW + t * fs + br + (int64_t)(k - 3) * sa;
                        JFA_ACCESS(p, sizeof(double));
                        *p += g[t];
                    }
                }
            }
        } else if constexpr (VAR == JFA_CONSERVATIVE) {
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                double g6[6], dU[6][5];
                cons_window_vjp<STENCIL>(W, br, sa, fs, i, prm, gl[i], gr[i], g6, dU);
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const int cell = f - 3 + k;
                    if (cell >= o0 && cell < o1) {
#pragma unroll
                        for (int t = 0; t < 5; ++t) {
                            if (!cons_dep(i, t)) continue;
                            double* p = gW + t * fs + br + (int64_t)(k - 3) * sa;
                            JFA_ACCESS(p, sizeof(double));
                            *p += g6[k] * dU[k][t];
                        }
                    }
                }
            }
        } else {
            double wl[5], wr[5];
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                wl[i] = ld(W, i * fs + br - sa);
                wr[i] = ld(W, i * fs + br);
            }
            const CharK K = char_frozen(wl, wr, prm);
            double gk[3] = {0.0, 0.0, 0.0};
            for (int c = 0; c < 5; ++c) {
                double g6[6];
                char_window_vjp<A, STENCIL>(W, br, sa, fs, c, K, gl[c], gr[c], prm, g6, gk);
                int i0, i1;
                double c0, c1;
                const int nt = char_terms<A>(c, K, i0, c0, i1, c1);
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const int cell = f - 3 + k;
                    if (cell >= o0 && cell < o1) {
                        double* p = gW + i0 * fs + br + (int64_t)(k - 3) * sa;
                        JFA_ACCESS(p, sizeof(double));
                        *p += c0 * g6[k];
                        if (nt == 2) {
                            double* q = gW + i1 * fs + br + (int64_t)(k - 3) * sa;
                            JFA_ACCESS(q, sizeof(double));
                            *q += c1 * g6[k];
                        }
                    }
                }
            }
            double ga, gq, gwf[2][5];
            char_frozen_vjp(K, gk, fz, ga, gq);
            frozen_cells_vjp(wl, wr, K, ga, gq, prm, gwf[0], gwf[1]);
#pragma unroll
            for (int k = 2; k < 4; ++k) {  // the frozen state reads the face's two cells
                const int cell = f - 3 + k;
                if (cell >= o0 && cell < o1) {
#pragma unroll
                    for (int i = 0; i < 5; ++i) {
                        double* p = gW + i * fs + br + (int64_t)(k - 3) * sa;
                        JFA_ACCESS(p, sizeof(double));
                        *p += gwf[k - 2][i];
                    }
                }
            }
        }
    }
}

// Lane-parallel transpose sweep along axis A (stride s[A]; unit stride on the innermost axis).
// Wave w of a row evaluates the faces F0 .. F0+63, F0 = -5 + 59 w (lane l: face F0 + l, zero
// outside 0..N) and lanes 5..63 own the cells F0 + l - 3, i.e. [-3 + 59 w, -3 + 59 w + 59)
// clipped to the row's N + 6 cells.  Rows run over the two other axes P, Q.
template <int A, int STENCIL, int RIEMANN, int VAR>
__global__ __launch_bounds__(kAdjBlock) void adj_lane_kernel(const double* __restrict__ W,
                                                             const double* __restrict__ lam, Geo G, Params prm,
                                                             double ts, int waves_per_row, int64_t nwaves,
                                                             double* __restrict__ gW) {
    constexpr int P = A == 0 ? 1 : 0;
    constexpr int Q = A == 2 ? 1 : 2;
    const int64_t gwave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const bool vwave = gwave < nwaves;  // whole waves; shuffles below run in every lane of a wave
    const int64_t row = vwave ? gwave / waves_per_row : 0;
    const int w = vwave ? (int)(gwave - row * waves_per_row) : 0;
    const int N = G.n[A];
    const int cP = (int)(row / G.n[Q]);
    const int cQ = (int)(row - (int64_t)cP * G.n[Q]);
    const int64_t fs = G.fs, sa = G.s[A];
    const int64_t brow = (int64_t)(cP + G.off[P]) * G.s[P] + (int64_t)(cQ + G.off[Q]) * G.s[Q] +
                         (int64_t)G.off[A] * sa;  // cell 0 of the row along A
    const int f = -5 + 59 * w + lane;
    const bool vface = vwave && f >= 0 && f <= N;
    const int64_t br = brow + (vface ? f : 0) * sa;
    if constexpr (RIEMANN == kFluxSplit) {
        // lanes without a face evaluate face 0 of the row with a zero cotangent
        double mu[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gW6[6][5], s6[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (vface) face_mu(lam, br, sa, fs, f, N, ts * G.one_dx[A], mu);
        flux_split_vjp<A, STENCIL>(W, br, sa, fs, mu, prm, gW6);
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int t = 0; t < 5; ++t) s6[t] += __shfl_up(gW6[k][t], (unsigned)k, 64);
        const int cell6 = f - 3;
        if (vwave && lane >= 5 && cell6 <= N + 2) add_cell(gW, fs, brow + (int64_t)cell6 * sa, s6);
        return;
    }
    double gl[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gr[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, fz[FZ_N];
#pragma unroll
    for (int i = 0; i < FZ_N; ++i) fz[i] = 0.0;
    if (vface) {
        double mu[5];
        face_mu(lam, br, sa, fs, f, N, ts * G.one_dx[A], mu);
        face_state_vjp_at<A, STENCIL, RIEMANN, VAR>(W, br, sa, fs, mu, prm, gl, gr, fz);
    }
    // cell c = f - 3 gathers window slot k of face c + 3 - k, held by lane - k
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if constexpr (VAR == JFA_PRIMITIVE) {
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            double x6[6] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0}, o6[6];  // lanes without a face: zero cotangents
            if (vface) load_var6(W, br, sa, fs, i, x6);
            window_vjp_var<STENCIL>(x6, gl[i], gr[i], prm, o6);
#pragma unroll
            for (int k = 0; k < 6; ++k) s[i] += __shfl_up(o6[k], (unsigned)k, 64);
        }
    } else if constexpr (VAR == JFA_CHAR_CONSERVATIVE) {
        const int64_t bq = vface ? br : brow;  // lanes without a face: face 0, zero cotangents
        double wl[5], wr[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            wl[i] = ld(W, i * fs + bq - sa);
            wr[i] = ld(W, i * fs + bq);
        }
        Frozen fr;
        frozen_state<true>(wl, wr, prm, fr);
        double Lm[5][5], Rm[5][5];
        conservative_eigenvectors<A>(fr, Lm, Rm);
        double gU6[6][5];
        char_cons_window_vjp<A, STENCIL>(W, bq, sa, fs, fr, Lm, gl, gr, prm, gU6, fz);
        double gwf[2][5];
        char_cons_frozen_vjp(wl, wr, fr, fz, prm, gwf[0], gwf[1]);
        if (!vface)
#pragma unroll
            for (int t = 0; t < 5; ++t) gwf[0][t] = gwf[1][t] = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double w[5], g[5];
#pragma unroll
            for (int t = 0; t < 5; ++t) {
                w[t] = ld(W, t * fs + bq + (int64_t)(k - 3) * sa);
                g[t] = (k == 2 || k == 3) ? gwf[k - 2][t] : 0.0;
            }
            cons_vjp(w, gU6[k], prm, g);
#pragma unroll
            for (int t = 0; t < 5; ++t) s[t] += __shfl_up(g[t], (unsigned)k, 64);
        }
    } else if constexpr (VAR == JFA_CONSERVATIVE) {
        const int64_t bq = vface ? br : brow;  // lanes without a face: face 0, zero cotangents
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            double g6[6], dU[6][5];
            cons_window_vjp<STENCIL>(W, bq, sa, fs, i, prm, gl[i], gr[i], g6, dU);
#pragma unroll
            for (int t = 0; t < 5; ++t) {
                if (!cons_dep(i, t)) continue;
#pragma unroll
                for (int k = 0; k < 6; ++k) s[t] += __shfl_up(g6[k] * dU[k][t], (unsigned)k, 64);
            }
        }
    } else {
        // lanes without a face evaluate face 0 of the row (its window lies in the interior and
        // the face halos) with zero cotangents, so they add nothing
        const int64_t bq = vface ? br : brow;
        double wl[5], wr[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            wl[i] = ld(W, i * fs + bq - sa);
            wr[i] = ld(W, i * fs + bq);
        }
        const CharK K = char_frozen(wl, wr, prm);
        double gk[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            double g6[6];
            char_window_vjp<A, STENCIL>(W, bq, sa, fs, c, K, gl[c], gr[c], prm, g6, gk);
            int i0, i1;
            double c0, c1;
            const int nt = char_terms<A>(c, K, i0, c0, i1, c1);
            double t0 = 0.0, t1 = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                t0 += __shfl_up(c0 * g6[k], (unsigned)k, 64);
                t1 += __shfl_up(c1 * g6[k], (unsigned)k, 64);
            }
            // i0, i1 are uniform across the wave (they depend on c and A only)
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                if (i == i0) s[i] += t0;
                if (nt == 2 && i == i1) s[i] += t1;
            }
        }
        double ga, gq, gwl[5], gwr[5];
        char_frozen_vjp(K, gk, fz, ga, gq);
        if (!vface) ga = gq = 0.0;
        frozen_cells_vjp(wl, wr, K, ga, gq, prm, gwl, gwr);
        // the frozen state reads the face's two cells (window slots 2 and 3)
#pragma unroll
        for (int i = 0; i < 5; ++i) s[i] += __shfl_up(gwl[i], 2u, 64) + __shfl_up(gwr[i], 3u, 64);
    }
    const int cell = f - 3;
    if (vwave && lane >= 5 && cell <= N + 2) add_cell(gW, fs, brow + (int64_t)cell * sa, s);
}

__device__ __forceinline__ int64_t interior_flat(const Geo& G, const int c[3]) {
    int w[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        w[i] = c[i];
        if (G.active[i]) w[i] = w[i] < 0 ? w[i] + G.n[i] : (w[i] >= G.n[i] ? w[i] - G.n[i] : w[i]);
    }
    return ((int64_t)w[0] * G.n[1] + w[1]) * G.n[2] + w[2];
}

// Dissipative part, one thread per interior cell c: the VJPs of both faces of c along every
// active axis (periodic: the forward's faces 0 and n evaluate the same wrapped cells, so their
// cotangents combine into ts (lam_R - lam_L) / dx with wrapped lam), so c owns its whole
// dissipative cotangent, added to gW, and the sum X[3 i + k](c) of the transverse-derivative
// cotangents Gt of its faces (interior layout, ncell per field), gathered by adj_update_kernel.
template <int UNUSED = 0>
__global__ __launch_bounds__(kBlock) void adj_dissip_cell_kernel(const double* __restrict__ W,
                                                                 const double* __restrict__ lam, Geo G, Params prm,
                                                                 double ts, double* __restrict__ gW,
                                                                 double* __restrict__ X) {
    int c[3];
    const int64_t ncell = (int64_t)G.n[0] * G.n[1] * G.n[2];
    if (!flat_index(ncell, G.n[1], G.n[2], c[0], c[1], c[2])) return;
    const int64_t idx = ((int64_t)c[0] * G.n[1] + c[1]) * G.n[2] + c[2];
    const int64_t b = wrapped_index(G, c, 0), fs = G.fs;
    double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, x[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double lc[5] = {ld(lam, b), ld(lam, fs + b), ld(lam, 2 * fs + b), ld(lam, 3 * fs + b), ld(lam, 4 * fs + b)};
#pragma unroll
    for (int A = 0; A < 3; ++A) {
        if (!G.active[A]) continue;
#pragma unroll
        for (int side = 0; side < 2; ++side) {  // 0: lower face (c is its right cell), 1: upper face
            int cr[3] = {c[0], c[1], c[2]}, co[3] = {c[0], c[1], c[2]};
            cr[A] += side;
            co[A] += side ? 1 : -1;  // the face's other cell
            const int64_t bo = wrapped_index(G, co, 0);
            double mu[5], gl[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gr[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, Gt[2][3];
            const double sc = ts * G.one_dx[A];
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                const double lo = ld(lam, i * fs + bo);
                mu[i] = side ? sc * (lo - lc[i]) : sc * (lc[i] - lo);
            }
            if (A == 0) dissipative_face_vjp<0>(W, G, prm, cr, mu, gl, gr, Gt);
            else if (A == 1) dissipative_face_vjp<1>(W, G, prm, cr, mu, gl, gr, Gt);
            else dissipative_face_vjp<2>(W, G, prm, cr, mu, gl, gr, Gt);
#pragma unroll
            for (int i = 0; i < 5; ++i) g[i] += side ? gl[i] : gr[i];
            int j = 0;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                if (t == A) continue;
#pragma unroll
                for (int k = 0; k < 3; ++k) x[3 * t + k] += Gt[j][k];
                ++j;
            }
        }
    }
    add_cell(gW, fs, b, g);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        JFA_ACCESS(X + q * ncell + idx, sizeof(double));
        X[q * ncell + idx] = x[q];
    }
}

// Transverse-derivative cotangents of cell c's velocity: a face adds +Gt to u at L + e_i, R + e_i
// and -Gt at L - e_i, R - e_i, so u_k(c) gathers X[3 i + k](c - e_i) - X[3 i + k](c + e_i)
// (periodic neighbours; X in interior layout).
__device__ __forceinline__ void xbar_gather(const double* __restrict__ X, const Geo& G, const int c[3], double g[5]) {
    const int64_t ncell = (int64_t)G.n[0] * G.n[1] * G.n[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (!G.active[i]) continue;
        int ym[3] = {c[0], c[1], c[2]}, yp[3] = {c[0], c[1], c[2]};
        ym[i] -= 1;
        yp[i] += 1;
        const int64_t jm = interior_flat(G, ym), jp = interior_flat(G, yp);
#pragma unroll
        for (int k = 0; k < 3; ++k) g[1 + k] += ld(X, (3 * i + k) * ncell + jm) - ld(X, (3 * i + k) * ncell + jp);
    }
}

// Per interior cell: halo^T, (dW/dU)^T and the stage blend.
__device__ __forceinline__ void halo_gather(const double* __restrict__ gW, const Geo& G, const BC6& bc,
                                            const int c[3], int64_t b, double g[5]) {
    const int nh = G.nh;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!G.active[a]) continue;
        const int N = G.n[a];
        const int ci = c[a];
        const int64_t sa = G.s[a];
        const int64_t bc0 = b - (int64_t)(ci + nh) * sa;
        const int tl = bc.t[2 * a + 1], th = bc.t[2 * a];
        // low halo (west/south/bottom) -- transpose of halo_scatter's low side
        if (tl == JFA_BC_PERIODIC) {
            if (ci >= N - nh)
                for (int i = 0; i < 5; ++i) g[i] += ld(gW, i * G.fs + bc0 + (int64_t)(ci - N + nh) * sa);
        } else if (tl == JFA_BC_ZEROGRADIENT) {
            if (ci == 0)
                for (int m = 0; m < nh; ++m)
                    for (int i = 0; i < 5; ++i) g[i] += ld(gW, i * G.fs + bc0 + (int64_t)m * sa);
        } else if (ci < nh) {
            for (int i = 0; i < 5; ++i) {
                const double h = ld(gW, i * G.fs + bc0 + (int64_t)(nh - 1 - ci) * sa);
                g[i] += i == a + 1 ? -h : h;
            }
        }
        // high halo (east/north/top)
        if (th == JFA_BC_PERIODIC) {
            if (ci < nh)
                for (int i = 0; i < 5; ++i) g[i] += ld(gW, i * G.fs + bc0 + (int64_t)(N + nh + ci) * sa);
        } else if (th == JFA_BC_ZEROGRADIENT) {
            if (ci == N - 1)
                for (int m = 0; m < nh; ++m)
                    for (int i = 0; i < 5; ++i) g[i] += ld(gW, i * G.fs + bc0 + (int64_t)(N + nh + m) * sa);
        } else if (ci >= N - nh) {
            for (int i = 0; i < 5; ++i) {
                const double h = ld(gW, i * G.fs + bc0 + (int64_t)(N + nh + (N - 1 - ci)) * sa);
                g[i] += i == a + 1 ? -h : h;
            }
        }
    }
}

template <int UNUSED = 0>  // a template, so the translation units that include this header share it
__global__ __launch_bounds__(kBlock) void adj_update_kernel(const double* __restrict__ W,
                                                            const double* __restrict__ lam,
                                                            const double* __restrict__ gW, Geo G, Params prm,
                                                            BC6 bc, double c0, double c1, double* __restrict__ lam_in,
                                                            double* __restrict__ lam_n,
                                                            const double* __restrict__ X) {
    int c[3];
    if (!flat_index((int64_t)G.n[0] * G.n[1] * G.n[2], G.n[1], G.n[2], c[0], c[1], c[2])) return;
    const int64_t b = (int64_t)(c[0] + G.off[0]) * G.s[0] + (int64_t)(c[1] + G.off[1]) * G.s[1] + (c[2] + G.off[2]);
    const int64_t fs = G.fs;
    double g[5], w[5], l[5], gu[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        g[i] = ld(gW, i * fs + b);
        w[i] = ld(W, i * fs + b);
        l[i] = ld(lam, i * fs + b);
    }
    halo_gather(gW, G, bc, c, b, g);
    if (X) xbar_gather(X, G, c, g);
    prim_vjp(w, g, prm, gu);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        JFA_ACCESS(lam_in + i * fs + b, sizeof(double));
        lam_in[i * fs + b] += c0 * l[i] + gu[i];
    }
    if (lam_n) {
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            JFA_ACCESS(lam_n + i * fs + b, sizeof(double));
            lam_n[i * fs + b] += c1 * l[i];
        }
    }
}

template <int UNUSED = 0>
__global__ __launch_bounds__(kBlock) void adj_zero_kernel(double* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// ------------------------------------------------------------------ host side
template <int A, int IN, int STENCIL, int RIEMANN, int VAR>
static void launch_adj_march(const double* W, const double* lam, const Geo& G, const Params& prm, double ts,
                             double* gW, hipStream_t st) {
    constexpr int O = 3 - A - IN;
    int bx = ((G.n[IN] + 63) / 64) * 64;
    if (bx > kAdjBlock) bx = kAdjBlock;
    const int L = 16;
    const int ncell = G.n[A] + 6;
    dim3 grid((unsigned)((G.n[IN] + bx - 1) / bx), (unsigned)G.n[O], (unsigned)((ncell + L - 1) / L));
    hipLaunchKernelGGL((adj_march_kernel<A, IN, STENCIL, RIEMANN, VAR>), grid, dim3(bx), 0, st, W, lam, G, prm, ts,
                       L, gW);
}

template <int A, int STENCIL, int RIEMANN, int VAR>
static void launch_adj_lanes(const double* W, const double* lam, const Geo& G, const Params& prm, double ts,
                             double* gW, hipStream_t st) {
    constexpr int P = A == 0 ? 1 : 0;
    constexpr int Q = A == 2 ? 1 : 2;
    const int waves_per_row = (G.n[A] + 6 + 58) / 59;
    const int64_t nwaves = (int64_t)waves_per_row * G.n[P] * G.n[Q];
    const int64_t blocks = (nwaves + kAdjBlock / 64 - 1) / (kAdjBlock / 64);
    hipLaunchKernelGGL((adj_lane_kernel<A, STENCIL, RIEMANN, VAR>), dim3((unsigned)blocks), dim3(kAdjBlock), 0, st,
                       W, lam, G, prm, ts, waves_per_row, nwaves, gW);
}

// Waves of the marching sweep along A (columns along IN, chunks of 16 cells along A).
static inline int64_t march_waves(const Geo& G, int A, int IN) {
    const int O = 3 - A - IN;
    return (int64_t)((G.n[IN] + 63) / 64) * G.n[O] * ((G.n[A] + 6 + 15) / 16);
}

// Below this many waves the marching sweep cannot fill the device and the lane-parallel sweep
// (strided loads along A, every face its own lane) is used for that axis instead.
int64_t adjoint_march_min_waves();

// Transpose sweeps of the convective flux along every active axis into gW.
template <int STENCIL, int RIEMANN, int VAR>
static void adj_sweeps(const double* W, const double* lam, const Geo& G, const Params& prm, double ts, double* gW,
                       hipStream_t st) {
    const int IN = innermost_axis(G);
    const int64_t min_waves = adjoint_march_min_waves();
    for (int A = 0; A < IN; ++A) {
        if (!G.active[A]) continue;
        if (march_waves(G, A, IN) < min_waves) {
            if (A == 0) launch_adj_lanes<0, STENCIL, RIEMANN, VAR>(W, lam, G, prm, ts, gW, st);
            else launch_adj_lanes<1, STENCIL, RIEMANN, VAR>(W, lam, G, prm, ts, gW, st);
        } else if (A == 0 && IN == 1) {
            launch_adj_march<0, 1, STENCIL, RIEMANN, VAR>(W, lam, G, prm, ts, gW, st);
        } else if (A == 0) {
            launch_adj_march<0, 2, STENCIL, RIEMANN, VAR>(W, lam, G, prm, ts, gW, st);
        } else {
            launch_adj_march<1, 2, STENCIL, RIEMANN, VAR>(W, lam, G, prm, ts, gW, st);
        }
    }
    if (IN == 0) launch_adj_lanes<0, STENCIL, RIEMANN, VAR>(W, lam, G, prm, ts, gW, st);
    else if (IN == 1) launch_adj_lanes<1, STENCIL, RIEMANN, VAR>(W, lam, G, prm, ts, gW, st);
    else launch_adj_lanes<2, STENCIL, RIEMANN, VAR>(W, lam, G, prm, ts, gW, st);
}

// adj_sweeps for a Riemann solver chosen at run time.
template <int STENCIL, int VAR>
inline void adj_sweeps_r(int riemann, const double* W, const double* lam, const Geo& G, const Params& prm, double ts,
                         double* gW, hipStream_t st) {
    if (riemann == JFA_FS_ROE || riemann == JFA_FS_CLLF || riemann == JFA_FS_LLF) {
        // the flux-splitting face flux reads the primitive window (the scheme's variable is unused)
        if constexpr (VAR == JFA_PRIMITIVE) adj_sweeps<STENCIL, kFluxSplit, JFA_PRIMITIVE>(W, lam, G, prm, ts, gW, st);
        return;
    }
    if (riemann == JFA_HLLC) adj_sweeps<STENCIL, JFA_HLLC, VAR>(W, lam, G, prm, ts, gW, st);
    else if (riemann == JFA_HLL) adj_sweeps<STENCIL, JFA_HLL, VAR>(W, lam, G, prm, ts, gW, st);
    else if (riemann == JFA_HLLC_LM) adj_sweeps<STENCIL, JFA_HLLC_LM, VAR>(W, lam, G, prm, ts, gW, st);
    else if (riemann == JFA_AUSMP) adj_sweeps<STENCIL, JFA_AUSMP, VAR>(W, lam, G, prm, ts, gW, st);
    else adj_sweeps<STENCIL, JFA_RUSANOV, VAR>(W, lam, G, prm, ts, gW, st);
}

// The sweeps of the run-time-selected stencils (WENO1, WENO3-JS/-Z, TENO5, TENO5-A;
// jfa_adjoint_ext.hip, a translation unit of its own).
void adj_sweeps_ext(int var, int riemann, const double* W, const double* lam, const Geo& G, const Params& prm,
                    double ts, double* gW, hipStream_t st);

}  // namespace jfa
