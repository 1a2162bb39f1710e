"""PyTorch (CPU, float64) restatement of one RK3 stage, for reverse-mode checks.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this module; only ``tests/``
use it, as the checker of the hand-written stage VJP (jfa_rk3_stage_vjp): torch.autograd of
this restatement stands in for jax.grad of the reference, which cannot be imported here.

The stage map is  f(U_in) = c0 U_in + c1 U_n + dt m rhs(prim(U_in))  on the interior, with the
face halos filled from the interior as the reference fills them (the same formulas, operands
and association order as ``jxf_oracle``; each function cites the reference file:line).
Covered: WENO5-JS / WENO5-Z (and WENO1, WENO3-JS / -Z, TENO5, TENO5-A) on the primitives, on the conservatives (CONSERVATIVE) or, with
the ARITHMETIC or ROE frozen state, on the characteristic projection of the primitives
(CHAR-PRIMITIVE) or of the conservatives (CHAR-CONSERVATIVE), HLLC with the ARITHMETIC / RUSANOV / DAVIS /
EINFELDT / TORO signal speeds, HLL and HLLC-LM with the same speeds, AUSM+, Rusanov, the flux-splitting scheme (ROE / CLLF / LLF), PERIODIC / ZEROGRADIENT / SYMMETRY face halos, and on periodic
domains the second-order central viscous and heat fluxes with ``jxf_oracle.Transport``'s laws.

Derivative conventions at non-smooth points follow torch's, which are jax's for the operations
used: d|x|/dx = sign(x) (0 at 0), d sign(x)/dx = 0, and maximum / minimum give half of the
cotangent to each operand at a tie (jax's balanced max); ``tests/test_torch_stage.py`` pins
them.
"""
from __future__ import annotations

import torch

from .jxf_oracle import (EPS64, FACE_AXIS, FACE_LOCATIONS, RK3_CONS_MULTIPLIER, RK3_TIMESTEP_MULTIPLIER,
                         SPATIAL_STENCIL_EPS, VEL_ID_EIGS, VELOCITY_MINOR_AXES, WENO5_SHIFTS, WENO5_WINDOW_IDX,
                         Domain, Transport, _face_fill_slice, shift_slice)

_DR = (1 / 10, 6 / 10, 3 / 10)   # weno5_base.py:24
_CR = ((1 / 3, -7 / 6, 11 / 6),  # weno5_base.py:25-27
       (-1 / 6, 5 / 6, 1 / 3),
       (1 / 3, 5 / 6, -1 / 6))


# ------------------------------------------------------------------ WENO5
def weno5_smoothness(a, b, c, d, e):
    """weno5_base.py:34-41."""
    beta_0 = 13.0 / 12.0 * torch.square(a - 2 * b + c) + 1.0 / 4.0 * torch.square(a - 4 * b + 3 * c)
    beta_1 = 13.0 / 12.0 * torch.square(b - 2 * c + d) + 1.0 / 4.0 * torch.square(b - d)
    beta_2 = 13.0 / 12.0 * torch.square(c - 2 * d + e) + 1.0 / 4.0 * torch.square(3 * c - 4 * d + e)
    return beta_0, beta_1, beta_2


def weno5_polynomials(a, b, c, d, e):
    """weno5_base.py:44-50."""
    return (_CR[0][0] * a + _CR[0][1] * b + _CR[0][2] * c,
            _CR[1][0] * b + _CR[1][1] * c + _CR[1][2] * d,
            _CR[2][0] * c + _CR[2][1] * d + _CR[2][2] * e)


def weno5_js(a, b, c, d, e, eps=SPATIAL_STENCIL_EPS):
    """weno/weno5_js.py:39-58."""
    b0, b1, b2 = weno5_smoothness(a, b, c, d, e)
    al0 = _DR[0] * (1.0 / (b0 * b0 + eps))
    al1 = _DR[1] * (1.0 / (b1 * b1 + eps))
    al2 = _DR[2] * (1.0 / (b2 * b2 + eps))
    one_alpha = 1.0 / (al0 + al1 + al2)
    p0, p1, p2 = weno5_polynomials(a, b, c, d, e)
    return (al0 * one_alpha) * p0 + (al1 * one_alpha) * p1 + (al2 * one_alpha) * p2


def weno5_z(a, b, c, d, e, eps=SPATIAL_STENCIL_EPS):
    """weno/weno5_z.py:38-59."""
    b0, b1, b2 = weno5_smoothness(a, b, c, d, e)
    tau_5 = torch.abs(b0 - b2)
    al0 = _DR[0] * (1.0 + tau_5 / (b0 + eps))
    al1 = _DR[1] * (1.0 + tau_5 / (b1 + eps))
    al2 = _DR[2] * (1.0 + tau_5 / (b2 + eps))
    one_alpha = 1.0 / (al0 + al1 + al2)
    p0, p1, p2 = weno5_polynomials(a, b, c, d, e)
    return (al0 * one_alpha) * p0 + (al1 * one_alpha) * p1 + (al2 * one_alpha) * p2


_DR3 = (1 / 3, 2 / 3)              # weno3_base.py:24
_CR3 = ((-0.5, 1.5), (0.5, 0.5))   # weno3_base.py:25


def weno1(a, b, c, d, e, eps=SPATIAL_STENCIL_EPS):
    """weno/weno1_js.py:25-30: the adjacent cell's value."""
    return c


def _weno3_parts(b, c, d):
    """weno3_base.py:32-45: smoothness and candidates of the stencil (u_im, u_i, u_ip)."""
    return (torch.square(c - b), torch.square(d - c)), (_CR3[0][0] * b + _CR3[0][1] * c, _CR3[1][0] * c + _CR3[1][1] * d)


def weno3_js(a, b, c, d, e, eps=SPATIAL_STENCIL_EPS):
    """weno/weno3_js.py:27-40."""
    (b0, b1), (p0, p1) = _weno3_parts(b, c, d)
    al0 = _DR3[0] * (1.0 / (b0 * b0 + eps))
    al1 = _DR3[1] * (1.0 / (b1 * b1 + eps))
    one_alpha = 1.0 / (al0 + al1)
    return al0 * one_alpha * p0 + al1 * one_alpha * p1


def weno3_z(a, b, c, d, e, eps=SPATIAL_STENCIL_EPS):
    """weno/weno3_z.py:27-39."""
    (b0, b1), (p0, p1) = _weno3_parts(b, c, d)
    tau_3 = torch.abs(b0 - b1)
    al0 = _DR3[0] * (1.0 + tau_3 / (b0 + eps))
    al1 = _DR3[1] * (1.0 + tau_3 / (b1 + eps))
    one_alpha = 1.0 / (al0 + al1)
    return al0 * one_alpha * p0 + al1 * one_alpha * p1


def _teno5(a, b, c, d, e, CT, dr, eps):
    """teno/teno5.py:44-73: gamma_k = (1 + tau5/(beta_k + eps))^6, sharp cutoff gamma_k / sum < CT,
    omega_k = d_k delta_k / (sum d delta + eps), value sum omega_k p_k."""
    beta = weno5_smoothness(a, b, c, d, e)
    tau_5 = torch.abs(beta[0] - beta[2])
    gamma = [torch.pow(1.0 + tau_5 / (bk + eps), 6.0) for bk in beta]
    one_gamma_sum = 1.0 / (gamma[0] + gamma[1] + gamma[2])
    w = [dr[k] * torch.where(gamma[k] * one_gamma_sum < CT, torch.zeros_like(tau_5), torch.ones_like(tau_5))
         for k in range(3)]
    one_dk = 1.0 / (w[0] + w[1] + w[2] + eps)
    p = weno5_polynomials(a, b, c, d, e)
    return (w[0] * one_dk) * p[0] + (w[1] * one_dk) * p[1] + (w[2] * one_dk) * p[2]


def teno5(a, b, c, d, e, eps=SPATIAL_STENCIL_EPS):
    """teno/teno5.py:32-73: d = (0.05, 0.55, 0.40), CT = 1e-5."""
    return _teno5(a, b, c, d, e, 1e-5, (0.05, 0.55, 0.40), eps)


def teno5_a(a, b, c, d, e, eps=SPATIAL_STENCIL_EPS):
    """teno/teno5_a.py:42-99: the adaptive cutoff from eta of neighbouring differences (xi = 1e-3,
    Cr = 0.24), WENO5's d."""
    ed = 0.9 * 0.24 / (1 - 0.24) * 1e-3 ** 2
    eta0 = (torch.abs(2 * (c - b) * (b - a)) + ed) / ((c - b) ** 2 + (b - a) ** 2 + ed)
    eta1 = (torch.abs(2 * (d - c) * (c - b)) + ed) / ((d - c) ** 2 + (c - b) ** 2 + ed)
    eta2 = (torch.abs(2 * (e - d) * (d - c)) + ed) / ((e - d) ** 2 + (d - c) ** 2 + ed)
    eta = torch.minimum(eta0, torch.minimum(eta1, eta2))
    m = 1 - torch.minimum(torch.ones_like(eta), eta / 0.24)
    g = torch.pow(1 - m, 4) * (1 + 4 * m)
    CT = torch.pow(10.0, -(torch.ceil(10.0 - 5.0 * (1 - g)) - 1.0))
    return _teno5(a, b, c, d, e, CT, _DR, eps)


WENO5 = {"WENO5-JS": weno5_js, "WENO5-Z": weno5_z, "WENO1": weno1, "WENO3-JS": weno3_js, "WENO3-Z": weno3_z,
         "TENO5": teno5, "TENO5-A": teno5_a}


# ------------------------------------------------------------------ ideal gas
def cons_from_prim(W, gamma):
    """equation_manager.py:93-101, ideal_gas.py:85-88."""
    rho = W[0]
    e = W[4] / (rho * (gamma - 1.0))
    E = rho * (0.5 * (torch.square(W[1]) + torch.square(W[2]) + torch.square(W[3])) + e)
    return torch.stack([rho, rho * W[1], rho * W[2], rho * W[3], E])


def prim_from_cons(U, gamma):
    """equation_manager.py:164-171, ideal_gas.py:73-75."""
    rho = U[0]
    one_rho = 1.0 / rho
    u, v, w = U[1] * one_rho, U[2] * one_rho, U[3] * one_rho
    e = U[4] * one_rho - 0.5 * (torch.square(u) + torch.square(v) + torch.square(w))
    return torch.stack([rho, u, v, w, (gamma - 1.0) * e * rho])


def physical_flux_xi(W, U, axis):
    """equation_manager.py:237-252."""
    f = [U[axis + 1], U[axis + 1] * W[1], U[axis + 1] * W[2], U[axis + 1] * W[3], None]
    f[axis + 1] = f[axis + 1] + W[4]
    f[4] = W[axis + 1] * (U[4] + W[4])
    return torch.stack(f)


# ------------------------------------------------------------------ signal speeds, Riemann solvers
def _speeds(kind, u_L, u_R, a_L, a_R, rho_L, rho_R, p_L=None, p_R=None, gamma=None):
    """signal_speeds.py:10-157 (ARITHMETIC :10-29, RUSANOV :31-49, DAVIS :51-69, EINFELDT :109-133,
    TORO :135-157 with estimate_pressure :201-231)."""
    if kind == "TORO":
        rho_bar, a_bar = 0.5 * (rho_L + rho_R), 0.5 * (a_L + a_R)
        p_pvrs = 0.5 * (p_L + p_R) - 0.5 * (u_R - u_L) * rho_bar * a_bar
        p_star = torch.maximum(torch.zeros_like(p_pvrs), p_pvrs)
        g_ = (gamma + 1) * 0.5 / gamma

        def q(p_K):  # 1.0 * (p* <= p_K) + sqrt(1 + g_ (p*/p_K - 1)) * (p* > p_K); the shock branch
            # is evaluated where it is taken only, so its NaN elsewhere does not reach the gradient
            shock = p_star > p_K
            arg = torch.where(shock, 1 + g_ * (p_star / p_K - 1), torch.ones_like(p_K))
            return torch.where(shock, torch.sqrt(arg), torch.ones_like(p_K))

        return u_L - a_L * q(p_L), u_R + a_R * q(p_R)
    if kind == "ARITHMETIC":
        u_mean, a_mean = 0.5 * (u_L + u_R), 0.5 * (a_L + a_R)
        return torch.minimum(u_mean - a_mean, u_L - a_L), torch.maximum(u_mean + a_mean, u_R + a_R)
    if kind == "RUSANOV":
        S_plus = torch.maximum(torch.abs(u_L) + a_L, torch.abs(u_R) + a_R)
        return -S_plus, S_plus
    if kind == "DAVIS":
        return torch.minimum(u_L - a_L, u_R - a_R), torch.maximum(u_L + a_L, u_R + a_R)
    if kind == "EINFELDT":
        sL, sR = torch.sqrt(rho_L), torch.sqrt(rho_R)
        one_dens = 1.0 / (sL + sR)
        eta2 = 0.5 * sL * sR * one_dens * one_dens
        u_bar = (sL * u_L + sR * u_R) * one_dens
        d_bar = torch.sqrt((sL * a_L * a_L + sR * a_R * a_R) * one_dens + eta2 * torch.square(u_R - u_L))
        return torch.minimum(u_bar - d_bar, u_L - a_L), torch.maximum(u_bar + d_bar, u_R + a_R)
    raise NotImplementedError(kind)


def _flux_star(W, U, S_K, S_star, axis, left):
    """HLLC._compute_flux_star_K_single_phase (HLLC.py:41-78)."""
    v = axis + 1
    m0, m1 = VELOCITY_MINOR_AXES[axis]
    pre = (S_K - W[v]) / (S_K - S_star) * W[0]
    us = [pre, pre, pre, pre, pre * (U[4] / U[0] + (S_star - W[v]) * (S_star + W[4] / W[0] / (S_K - W[v])))]
    us[v] = us[v] * S_star
    us[m0] = us[m0] * W[m0]
    us[m1] = us[m1] * W[m1]
    S = torch.minimum(S_K, torch.zeros_like(S_K)) if left else torch.maximum(S_K, torch.zeros_like(S_K))
    return physical_flux_xi(W, U, axis) + S * (torch.stack(us) - U)


def hllc(W_L, W_R, U_L, U_R, axis, gamma, signal_speed):
    """HLLC._solve_riemann_problem_xi_single_phase (HLLC.py:80-126), S* from signal_speeds.py:159-199."""
    v = axis + 1
    c_L = torch.sqrt(gamma * W_L[4] / W_L[0])
    c_R = torch.sqrt(gamma * W_R[4] / W_R[0])
    S_L, S_R = _speeds(signal_speed, W_L[v], W_R[v], c_L, c_R, W_L[0], W_R[0], W_L[4], W_R[4], gamma)
    rdu_L = W_L[0] * (S_L - W_L[v])
    rdu_R = W_R[0] * (S_R - W_R[v])
    S_star = ((W_R[4] - W_L[4]) + (W_L[v] * rdu_L - W_R[v] * rdu_R)) / (rdu_L - rdu_R)
    F_L = _flux_star(W_L, U_L, S_L, S_star, axis, True)
    F_R = _flux_star(W_R, U_R, S_R, S_star, axis, False)
    sgn = torch.sign(S_star)
    return 0.5 * (1 + sgn) * F_L + 0.5 * (1 - sgn) * F_R


def rusanov(W_L, W_R, U_L, U_R, axis, gamma, signal_speed=None):
    """Rusanov._solve_riemann_problem_xi_single_phase (Rusanov.py:25-47)."""
    v = axis + 1
    c_L = torch.sqrt(gamma * W_L[4] / W_L[0])
    c_R = torch.sqrt(gamma * W_R[4] / W_R[0])
    alpha = torch.maximum(torch.abs(W_L[v]) + c_L, torch.abs(W_R[v]) + c_R)
    return 0.5 * (physical_flux_xi(W_L, U_L, axis) + physical_flux_xi(W_R, U_R, axis)) - 0.5 * alpha * (U_R - U_L)


def hll(W_L, W_R, U_L, U_R, axis, gamma, signal_speed, eps=EPS64):
    """HLL._solve_riemann_problem_xi_single_phase (HLL.py:29-61); eps the working precision's
    machine epsilon (riemann_solver.py:28)."""
    v = axis + 1
    c_L = torch.sqrt(gamma * W_L[4] / W_L[0])
    c_R = torch.sqrt(gamma * W_R[4] / W_R[0])
    S_L, S_R = _speeds(signal_speed, W_L[v], W_R[v], c_L, c_R, W_L[0], W_R[0], W_L[4], W_R[4], gamma)
    s_L = torch.minimum(S_L, torch.zeros_like(S_L))
    s_R = torch.maximum(S_R, torch.zeros_like(S_R))
    F_L = physical_flux_xi(W_L, U_L, axis)
    F_R = physical_flux_xi(W_R, U_R, axis)
    return (s_R * F_L - s_L * F_R + s_L * s_R * (U_R - U_L)) / (s_R - s_L + eps)


def hllc_lm(W_L, W_R, U_L, U_R, axis, gamma, signal_speed):
    """HLLCLM._solve_riemann_problem_xi_single_phase (HLLCLM.py:59-136): HLLC star states, wave
    speeds scaled by phi = sin(min(1, Ma / 0.1) pi / 2), Ma = max(|u_L / c_L|, |u_R / c_R|)
    (:37-57), the central star flux and the sign blend of the outer waves."""
    v = axis + 1
    m0, m1 = VELOCITY_MINOR_AXES[axis]
    c_L = torch.sqrt(gamma * W_L[4] / W_L[0])
    c_R = torch.sqrt(gamma * W_R[4] / W_R[0])
    S_L, S_R = _speeds(signal_speed, W_L[v], W_R[v], c_L, c_R, W_L[0], W_R[0], W_L[4], W_R[4], gamma)
    rdu_L = W_L[0] * (S_L - W_L[v])
    rdu_R = W_R[0] * (S_R - W_R[v])
    S_star = ((W_R[4] - W_L[4]) + (W_L[v] * rdu_L - W_R[v] * rdu_R)) / (rdu_L - rdu_R)

    def star(W, U, S_K):  # :95-108
        pre = (S_K - W[v]) / (S_K - S_star) * W[0]
        u = [pre, pre, pre, pre, pre * (U[4] / U[0] + (S_star - W[v]) * (S_star + W[4] / W[0] / (S_K - W[v])))]
        u[v] = u[v] * S_star
        u[m0] = u[m0] * W[m0]
        u[m1] = u[m1] * W[m1]
        return torch.stack(u)

    u_star_L, u_star_R = star(W_L, U_L, S_L), star(W_R, U_R, S_R)
    Ma = torch.maximum(torch.abs(W_L[v] / c_L), torch.abs(W_R[v] / c_R))
    phi = torch.sin(torch.minimum(torch.ones_like(Ma), Ma / 0.1) * torch.pi * 0.5)
    F_L = physical_flux_xi(W_L, U_L, axis)
    F_R = physical_flux_xi(W_R, U_R, axis)
    flux_star = 0.5 * (F_L + F_R) + 0.5 * ((phi * S_L) * (u_star_L - U_L) + torch.abs(S_star) * (u_star_L - u_star_R)
                                           + (phi * S_R) * (u_star_R - U_R))
    return (0.5 * (1 + torch.sign(S_L)) * F_L + 0.5 * (1 - torch.sign(S_R)) * F_R
            + 0.25 * (1 - torch.sign(S_L)) * (1 + torch.sign(S_R)) * flux_star)


def ausmp(W_L, W_R, U_L, U_R, axis, gamma, signal_speed=None):
    """AUSMP._solve_riemann_problem_xi_single_phase (AUSMP.py:30-103): the arithmetic interface
    speed of sound, alpha = 3/16, beta = 1/8, phi = (rho, rho u, rho v, rho w, E + p)."""
    v = axis + 1
    alpha, beta = 3.0 / 16.0, 1.0 / 8.0
    phi_L = torch.stack([U_L[0], U_L[1], U_L[2], U_L[3], U_L[4] + W_L[4]])
    phi_R = torch.stack([U_R[0], U_R[1], U_R[2], U_R[3], U_R[4] + W_R[4]])
    a = 0.5 * (torch.sqrt(gamma * W_L[4] / W_L[0]) + torch.sqrt(gamma * W_R[4] / W_R[0]))
    M_l = W_L[v] / a
    M_r = W_R[v] / a
    M_plus = torch.where(torch.abs(M_l) >= 1, 0.5 * (M_l + torch.abs(M_l)),
                         0.25 * torch.square(M_l + 1.0) + beta * torch.square(M_l * M_l - 1.0))
    M_minus = torch.where(torch.abs(M_r) >= 1, 0.5 * (M_r - torch.abs(M_r)),
                          -0.25 * torch.square(M_r - 1.0) - beta * torch.square(M_r * M_r - 1.0))
    M = M_plus + M_minus
    P_plus = torch.where(torch.abs(M_l) >= 1.0, 0.5 * (1 + torch.sign(M_l)),
                         0.25 * torch.square(M_l + 1.0) * (2.0 - M_l) + alpha * M_l * torch.square(M_l * M_l - 1.0))
    P_minus = torch.where(torch.abs(M_r) >= 1.0, 0.5 * (1 - torch.sign(M_r)),
                          0.25 * torch.square(M_r - 1.0) * (2.0 + M_r) - alpha * M_r * torch.square(M_r * M_r - 1.0))
    F = a * (0.5 * (M + torch.abs(M)) * phi_L + 0.5 * (M - torch.abs(M)) * phi_R)
    e = torch.zeros_like(F)
    e[v] = 1.0
    return F + e * (P_plus * W_L[4] + P_minus * W_R[4])


RIEMANN = {"HLLC": hllc, "RUSANOV": rusanov, "HLL": hll, "HLLC-LM": hllc_lm, "AUSMP": ausmp}


# ------------------------------------------------------------------ characteristic projection
def _char_from_prim(x, P, c, cc, axis):
    """get_characteristics_from_primitives, single phase (eigendecomposition.py:421-433)."""
    v = axis + 1
    m0, m1 = VELOCITY_MINOR_AXES[axis]
    e0, e1 = VEL_ID_EIGS[axis]
    out = [None] * 5
    out[0] = -0.5 / c * x[v] + 0.5 / (cc * P[0]) * x[4]
    out[1] = x[0] - 1.0 / cc * x[4]
    out[e0] = x[m0]
    out[e1] = x[m1]
    out[4] = 0.5 / c * x[v] + 0.5 / (cc * P[0]) * x[4]
    return torch.stack(out)


def _prim_from_char(x, P, c, cc, axis):
    """get_primitives_from_characteristics, single phase (eigendecomposition.py:512-523)."""
    v = axis + 1
    m0, m1 = VELOCITY_MINOR_AXES[axis]
    e0, e1 = VEL_ID_EIGS[axis]
    out = [None] * 5
    out[0] = P[0] * (x[0] + x[4]) + x[1]
    out[v] = c * (-x[0] + x[4])
    out[m0] = x[e0]
    out[m1] = x[e1]
    out[4] = cc * P[0] * (x[0] + x[4])
    return torch.stack(out)


def _frozen(W_L, W_R, gamma, frozen):
    """compute_frozen_state, single phase (eigendecomposition.py:139-281): the density average
    P[0] and the sound speed c, c^2 the projection uses; ARITHMETIC means (:148,223-230) or ROE
    averages (:233-275)."""
    if frozen == "ARITHMETIC":
        P = 0.5 * (W_L + W_R)
        c = torch.sqrt(gamma * P[4] / P[0])
        return P, c, c * c
    sL, sR = torch.sqrt(W_L[0]), torch.sqrt(W_R[0])
    rho_div = 1.0 / (sL + sR)                                              # :240
    P0 = torch.sqrt(W_L[0] * W_R[0])                                       # :237
    psi_ave = (sL * (W_L[4] / W_L[0]) + sR * (W_R[4] / W_R[0])) * rho_div  # :258
    g = gamma - 1
    grueneisen_ave = (sL * g + sR * g) * rho_div                           # :260-262
    dvel = W_R[1:4] - W_L[1:4]
    sq_dvel = torch.square(dvel[0]) + torch.square(dvel[1]) + torch.square(dvel[2])  # :265
    p_over_rho_ave = (sL * W_L[4] / W_L[0] + sR * W_R[4] / W_R[0]) * rho_div \
        + 0.5 * P0 * rho_div * rho_div * sq_dvel                           # :267-268
    cc = psi_ave + grueneisen_ave * p_over_rho_ave                         # :274
    return torch.stack([P0]), torch.sqrt(cc), cc                           # :275


def char_primitive_states(W, dom: Domain, axis: int, stencil: str, gamma: float, frozen: str = "ARITHMETIC"):
    """HighOrderGodunov.reconstruct_xi, CHAR-PRIMITIVE (high_order_godunov.py:298-316,401-402) with
    the ARITHMETIC or ROE frozen state of the face's two cells: the 6-cell window of every face
    (eigendecomposition.py:75-118) projected with it, WENO5 on the characteristics (window slices
    weno5_base.py:31), projected back."""
    win = [W[shift_slice(dom, axis, k)] for k in range(-3, 3)]
    P, c, cc = _frozen(win[2], win[3], gamma, frozen)
    ch = [_char_from_prim(w, P, c, cc, axis) for w in win]
    XL = WENO5[stencil](*[ch[i] for i in WENO5_WINDOW_IDX[0]])
    XR = WENO5[stencil](*[ch[i] for i in WENO5_WINDOW_IDX[1]])
    return _prim_from_char(XL, P, c, cc, axis), _prim_from_char(XR, P, c, cc, axis)


def _frozen_full(W_L, W_R, gamma, frozen):
    """compute_frozen_state, single phase (eigendecomposition.py:139-281), everything the
    conservative eigenvectors read: primes_ave, enthalpy_ave, grueneisen_ave, c, c^2, |u_ave|^2."""
    def enthalpy(p, rho, u):  # ideal_gas.py:90-110
        return (p / (gamma - 1) + 0.5 * rho * (torch.square(u[0]) + torch.square(u[1]) + torch.square(u[2])) + p) / rho

    if frozen == "ARITHMETIC":
        P = 0.5 * (W_L + W_R)                                                   # :148
        g = (gamma - 1) * torch.ones_like(P[0])                                 # :216-218
        H = enthalpy(P[4], P[0], P[1:4])                                        # :219-222
        c = torch.sqrt(gamma * P[4] / P[0])                                     # :223-225
        cc = c * c                                                              # :230
    else:
        sL, sR = torch.sqrt(W_L[0]), torch.sqrt(W_R[0])
        P = (sL * W_L + sR * W_R) / (sL + sR)                                   # :293
        P = torch.cat([torch.sqrt(W_L[0] * W_R[0])[None], P[1:]])               # :237
        rho_div = 1.0 / (sL + sR)                                               # :240
        H = (sL * enthalpy(W_L[4], W_L[0], W_L[1:4]) + sR * enthalpy(W_R[4], W_R[0], W_R[1:4])) * rho_div  # :250
        psi_ave = (sL * (W_L[4] / W_L[0]) + sR * (W_R[4] / W_R[0])) * rho_div  # :258
        g = (sL * (gamma - 1) + sR * (gamma - 1)) * rho_div                     # :260-262
        dvel = W_R[1:4] - W_L[1:4]
        sq_dvel = torch.square(dvel[0]) + torch.square(dvel[1]) + torch.square(dvel[2])  # :265
        p_over_rho_ave = (sL * W_L[4] / W_L[0] + sR * W_R[4] / W_R[0]) * rho_div \
            + 0.5 * P[0] * rho_div * rho_div * sq_dvel                          # :267-268
        cc = psi_ave + g * p_over_rho_ave                                       # :274
        c = torch.sqrt(cc)                                                      # :275
    q2 = torch.square(P[1]) + torch.square(P[2]) + torch.square(P[3])          # :231, :271
    return P, H, g, c, cc, q2


def _eigen_conservatives(P, H, g, c, cc, q2, axis):
    """eigendecomposition_conservatives, single phase (eigendecomposition.py:595-666): R, L as
    5 x 5 nested lists indexed [row][col]."""
    one_cc = 1.0 / cc
    one_rho = 1.0 / P[0]
    z = torch.zeros_like(P[0])
    o = torch.ones_like(P[0])
    R = [[z] * 5 for _ in range(5)]
    L = [[z] * 5 for _ in range(5)]
    v = axis + 1
    m0, m1 = VELOCITY_MINOR_AXES[axis]
    R[0][0] = o                                                   # :608-631
    R[v][0] = P[v] - c
    R[m0][0] = P[m0]
    R[m1][0] = P[m1]
    R[4][0] = H - P[v] * c
    R[0][v] = g
    R[1][v] = g * P[1]
    R[2][v] = g * P[2]
    R[3][v] = g * P[3]
    R[4][v] = g * H - cc
    R[m0][m0] = -P[0]
    R[4][m0] = -P[0] * P[m0]
    R[m1][m1] = P[0]
    R[4][m1] = P[0] * P[m1]
    R[0][4] = o
    R[v][4] = P[v] + c
    R[m0][4] = P[m0]
    R[m1][4] = P[m1]
    R[4][4] = H + P[v] * c
    L[0][0] = 0.5 * one_cc * (g * q2 - g * H + (P[v] + c) * c)  # :635-658
    L[0][v] = 0.5 * one_cc * (-P[v] * g - c)
    L[0][m0] = 0.5 * one_cc * (-P[m0] * g)
    L[0][m1] = 0.5 * one_cc * (-P[m1] * g)
    L[0][4] = 0.5 * one_cc * g
    L[v][0] = one_cc * (H - q2)
    L[v][1] = P[1] * one_cc
    L[v][2] = P[2] * one_cc
    L[v][3] = P[3] * one_cc
    L[v][4] = -one_cc
    L[m0][0] = P[m0] * one_rho
    L[m0][m0] = -one_rho
    L[m1][0] = -P[m1] * one_rho
    L[m1][m1] = one_rho
    L[4][0] = 0.5 * one_cc * (g * q2 - g * H - (P[v] - c) * c)
    L[4][v] = 0.5 * one_cc * (-P[v] * g + c)
    L[4][m0] = 0.5 * one_cc * (-P[m0] * g)
    L[4][m1] = 0.5 * one_cc * (-P[m1] * g)
    L[4][4] = 0.5 * one_cc * g
    return R, L


def _transform(M, x):
    """einsum 'ij...,j...->i...' (eigendecomposition.py:717-742), summed in j order."""
    out = []
    for i in range(5):
        acc = M[i][0] * x[0]
        for j in range(1, 5):
            acc = acc + M[i][j] * x[j]
        out.append(acc)
    return torch.stack(out)


def char_conservative_states(W, U, dom: Domain, axis: int, stencil: str, gamma: float, frozen: str = "ARITHMETIC"):
    """HighOrderGodunov.reconstruct_xi, CHAR-CONSERVATIVE (high_order_godunov.py:404-417): the
    conservative 6-cell window projected with the left eigenvectors at the frozen state of the
    face's two cells, WENO5 on the characteristics, back with the right eigenvectors; the face
    primitives are prim(U_L), prim(U_R)."""
    win = [U[shift_slice(dom, axis, k)] for k in range(-3, 3)]
    cl, cr = W[shift_slice(dom, axis, -1)], W[shift_slice(dom, axis, 0)]
    R, L = _eigen_conservatives(*_frozen_full(cl, cr, gamma, frozen), axis)
    ch = [_transform(L, w) for w in win]
    XL = WENO5[stencil](*[ch[i] for i in WENO5_WINDOW_IDX[0]])
    XR = WENO5[stencil](*[ch[i] for i in WENO5_WINDOW_IDX[1]])
    U_L, U_R = _transform(R, XL), _transform(R, XR)
    return U_L, U_R


def flux_splitting_flux_xi(W, U, dom: Domain, axis: int, stencil: str, gamma: float, frozen: str, kind: str):
    """FluxSplittingScheme.compute_flux_xi (flux_splitting_scheme.py:65-111), single phase: the
    window of conservatives and physical fluxes projected on the left eigenvectors of the frozen
    state of the face's two cells, f+- = (L F +- Lambda L U)/2 with the eigenvalue magnitudes
    of eigendecomposition.py:672-696 (ROE / CLLF / LLF), f+ reconstructed with j = 0, f- with
    j = 1, the sum projected back with the right eigenvectors."""
    v = axis + 1
    win_W = [W[shift_slice(dom, axis, k)] for k in range(-3, 3)]
    win_U = [U[shift_slice(dom, axis, k)] for k in range(-3, 3)]
    cl, cr = win_W[2], win_W[3]
    P, H, g, c, cc, q2 = _frozen_full(cl, cr, gamma, frozen)
    R, L = _eigen_conservatives(P, H, g, c, cc, q2, axis)
    if kind == "ROE":                                                        # :674-677
        g1, g234, g5 = torch.abs(P[v] - c), torch.abs(P[v]), torch.abs(P[v] + c)
    else:
        c_L = torch.sqrt(gamma * cl[4] / cl[0])
        c_R = torch.sqrt(gamma * cr[4] / cr[0])
        if kind == "CLLF":                                                   # :680-688
            g1 = torch.maximum(torch.abs(cl[v] - c_L), torch.abs(cr[v] - c_R))
            g234 = torch.maximum(torch.abs(cl[v]), torch.abs(cr[v]))
            g5 = torch.maximum(torch.abs(cl[v] + c_L), torch.abs(cr[v] + c_R))
        else:                                                                # :690-695
            g1 = g234 = g5 = torch.maximum(torch.abs(cl[v]) + c_L, torch.abs(cr[v]) + c_R)
    lam = torch.stack([g1, g234, g234, g234, g5])
    ch = [_transform(L, u) for u in win_U]
    chf = [_transform(L, physical_flux_xi(w, u, axis)) for w, u in zip(win_W, win_U)]
    pos = [0.5 * (cf + lam * cu) for cu, cf in zip(ch, chf)]
    neg = [0.5 * (cf - lam * cu) for cu, cf in zip(ch, chf)]
    f_L = WENO5[stencil](*[pos[i] for i in WENO5_WINDOW_IDX[0]])
    f_R = WENO5[stencil](*[neg[i] for i in WENO5_WINDOW_IDX[1]])
    return _transform(R, f_L + f_R)


# ------------------------------------------------------------------ dissipative fluxes (periodic)
def _law(T, ref, T_ref, exp, C):
    """material.py:82-124: SUTHERLAND ref ((T_ref + C)/(T + C)) (T/T_ref)^1.5 (power3_2 = x sqrt(x),
    math/power_functions.py:29-30), else the power law ref (T/T_ref)^exp (exp = 0: constant)."""
    if C is not None:
        x = T / T_ref
        return ref * ((T_ref + C) / (T + C)) * (x * torch.sqrt(x))
    return ref * (T / T_ref) ** exp if exp != 0.0 else ref * torch.ones_like(T)


def _wrap(W, dom: Domain):
    """The buffer with every halo (faces, edges, vertices) a periodic copy of the interior."""
    Wi = W[dom.interior]
    nh = dom.nh
    for i in dom.active:
        n = Wi.shape[1 + i]
        Wi = torch.cat([Wi.narrow(1 + i, n - nh, nh), Wi, Wi.narrow(1 + i, 0, nh)], dim=1 + i)
    return Wi


def _face_cells(buf, dom: Domain, axis: int, k: int, other=None, ok: int = 0):
    """Cells f + k of every face f along axis (k = -1 left, 0 right), shifted by ok along other."""
    n = dom.nh
    sl = [dom.dom(0), dom.dom(1), dom.dom(2)]
    hi = -n + k + 1
    sl[axis] = slice(n + k, None if hi == 0 else hi)
    if other is not None and ok != 0:
        sl[other] = slice(n + ok, -n + ok)
    return buf[(Ellipsis,) + tuple(sl)]


def dissipative_flux_xi(W, dom: Domain, axis: int, tr: Transport, gamma: float):
    """compute_viscous_flux_xi / compute_tau / compute_heat_flux_xi (source_term_solver.py:188-350,
    414-448,504-580) on a periodic domain, as the amounts added to the convective face flux
    (space_solver.py:565-587): momentum -tau_axis,k, energy -sum_k tau_axis,k u_k + q."""
    Wp = _wrap(W, dom)
    L, R = _face_cells(Wp, dom, axis, -1), _face_cells(Wp, dom, axis, 0)
    T_L, T_R = L[4] / (L[0] * tr.R), R[4] / (R[0] * tr.R)                     # ideal_gas.py:81-83
    T_f = 0.5 * (T_L + T_R)                                                      # central_2
    T_f = torch.where(T_f <= 0.0, torch.full_like(T_f, EPS64), T_f)              # :301
    mu1 = _law(T_f, tr.mu_ref, tr.T_ref, tr.mu_exp, tr.mu_sutherland)
    mu2 = tr.bulk_viscosity - 2.0 / 3.0 * mu1                                    # :527
    zero = torch.zeros_like(T_f)
    g = [[zero] * 3 for _ in range(3)]  # g[k][i] = d u_k / d x_i at the faces (:414-448,538-580)
    for i in dom.active:
        for k in range(3):
            if i == axis:
                g[k][i] = (R[1 + k] - L[1 + k]) * dom.one_dx[i]                  # deriv_face_2
            else:
                h = 0.5 * dom.one_dx[i]                                          # deriv_center_2
                dL = (_face_cells(Wp, dom, axis, -1, i, 1)[1 + k] - _face_cells(Wp, dom, axis, -1, i, -1)[1 + k]) * h
                dR = (_face_cells(Wp, dom, axis, 0, i, 1)[1 + k] - _face_cells(Wp, dom, axis, 0, i, -1)[1 + k]) * h
                g[k][i] = 0.5 * (dL + dR)                                        # central_2
    tau = [zero, zero, zero]
    if tr.viscous:
        for k in dom.active:
            tau[k] = mu1 * (g[axis][k] + g[k][axis])
        div = g[dom.active[0]][dom.active[0]]
        for k in dom.active[1:]:
            div = div + g[k][k]
        tau[axis] = tau[axis] + mu2 * div
    vel_tau = 0.0
    for k in dom.active:
        vel_tau = vel_tau + tau[k] * (0.5 * (L[1 + k] + R[1 + k]))
    E = -vel_tau if tr.viscous else zero
    if tr.heat:
        if tr.prandtl is not None:
            kappa = (gamma / (gamma - 1.0) * tr.R) * mu1 / tr.prandtl           # material.py:115-116
        else:
            kappa = _law(T_f, tr.kappa_ref, tr.kappa_T_ref, tr.kappa_exp, tr.kappa_sutherland)
        E = E + (-kappa * ((T_R - T_L) * dom.one_dx[axis]))                      # deriv_face_2
    return torch.stack([zero, -tau[0], -tau[1], -tau[2], E])


# ------------------------------------------------------------------ RHS and halos
def compute_rhs(W, dom: Domain, stencil: str, riemann: str, signal_speed, gamma: float, transport=None,
                variable: str = "PRIMITIVE", U=None, frozen: str = "ARITHMETIC"):
    """SpaceSolver.compute_rhs (space_solver.py:266,289-314,489,597-599) with PRIMITIVE
    (high_order_godunov.py:267-280) or CONSERVATIVE (:282-296: WENO on the conservative buffer U,
    then W_L = prim(U_L)) reconstruction: 0.0 + sum over axes of (1/dx)(F[:-1] - F[1:]),
    the dissipative face flux added to the convective one (:565-587) with a transport."""
    rhs = 0.0
    for axis in dom.active:
        if riemann.startswith("FLUX-SPLITTING-"):
            F = flux_splitting_flux_xi(W, U, dom, axis, stencil, gamma, frozen, riemann[len("FLUX-SPLITTING-"):])
            if transport is not None:
                F = F + dissipative_flux_xi(W, dom, axis, transport, gamma)
            n = F.shape[axis + 1]
            rhs = rhs + (0.0 + dom.one_dx[axis] * (F.narrow(axis + 1, 0, n - 1) - F.narrow(axis + 1, 1, n - 1)))
            continue
        if variable == "CHAR-PRIMITIVE":
            W_L, W_R = char_primitive_states(W, dom, axis, stencil, gamma, frozen)
        elif variable == "CONSERVATIVE":
            U_L = WENO5[stencil](*[U[shift_slice(dom, axis, k)] for k in WENO5_SHIFTS[0]])
            U_R = WENO5[stencil](*[U[shift_slice(dom, axis, k)] for k in WENO5_SHIFTS[1]])
        elif variable == "CHAR-CONSERVATIVE":
            U_L, U_R = char_conservative_states(W, U, dom, axis, stencil, gamma, frozen)
        else:
            W_L = WENO5[stencil](*[W[shift_slice(dom, axis, k)] for k in WENO5_SHIFTS[0]])
            W_R = WENO5[stencil](*[W[shift_slice(dom, axis, k)] for k in WENO5_SHIFTS[1]])
        if variable in ("CONSERVATIVE", "CHAR-CONSERVATIVE"):
            W_L, W_R = prim_from_cons(U_L, gamma), prim_from_cons(U_R, gamma)
        else:
            U_L, U_R = cons_from_prim(W_L, gamma), cons_from_prim(W_R, gamma)
        F = RIEMANN[riemann](W_L, W_R, U_L, U_R, axis, gamma, signal_speed)
        if transport is not None:
            F = F + dissipative_flux_xi(W, dom, axis, transport, gamma)
        n = F.shape[axis + 1]
        rhs = rhs + (0.0 + dom.one_dx[axis] * (F.narrow(axis + 1, 0, n - 1) - F.narrow(axis + 1, 1, n - 1)))
    return rhs


def face_halos(W, dom: Domain, bcs: dict):
    """face_halo_update (halos/outer/material.py:94-287; retrieve slices
    boundary_condition.py:564-595, fill slices halo_slices.py:66-73) of a buffer whose interior is
    set: PERIODIC copies the opposite interior layer, ZEROGRADIENT the first/last interior cell,
    SYMMETRY the mirrored layer with the normal velocity negated (:721-731)."""
    nh = dom.nh
    for loc in FACE_LOCATIONS:
        a = FACE_AXIS[loc]
        if a not in dom.active:
            continue
        hi = loc in ("east", "north", "top")
        bc = bcs[loc]
        s = [dom.dom(0), dom.dom(1), dom.dom(2)]
        if bc == "PERIODIC":
            s[a] = slice(nh, 2 * nh) if hi else slice(-2 * nh, -nh)
        elif bc == "ZEROGRADIENT":
            s[a] = slice(-nh - 1, -nh) if hi else slice(nh, nh + 1)
        elif bc == "SYMMETRY":  # the mirrored layer: the nh cells next to the face, reversed
            s[a] = slice(-2 * nh, -nh) if hi else slice(nh, 2 * nh)
        else:
            raise NotImplementedError(bc)
        halo = W[(Ellipsis,) + tuple(s)]
        if bc == "SYMMETRY":
            halo = torch.flip(halo, dims=(a + 1,))
            sign = torch.ones(5, dtype=W.dtype).reshape(5, 1, 1, 1)
            sign[a + 1] = -1.0
            halo = halo * sign
        W = W.clone()
        W[_face_fill_slice(dom, loc)] = halo  # ZEROGRADIENT's single layer broadcasts over the nh halo cells
    return W


def stage_map(U_int, Un_int, dom: Domain, stencil: str, riemann: str, signal_speed, gamma: float, bcs: dict,
              stage: int, dt: float, eps: float, transport=None, variable: str = "PRIMITIVE",
              frozen: str = "ARITHMETIC"):
    """f(U_in interior) -> U_out interior: W = prim(U_in) on the interior of an eps-filled buffer,
    face halos of W, rhs, then c0 U_in + c1 U_n + dt m rhs (RK3.py:27-62, time_integrator.py:57).
    The conservative buffer is U_in on the interior and cons(W) in the halos (outer/material.py:248-250)."""
    m = RK3_TIMESTEP_MULTIPLIER[stage]
    c0, c1 = (1.0, 0.0) if stage == 0 else RK3_CONS_MULTIPLIER[stage - 1]
    W = torch.full((5,) + dom.shape, eps, dtype=torch.float64)
    W[dom.interior] = prim_from_cons(U_int, gamma)
    W = face_halos(W, dom, bcs)
    if transport is not None:
        assert all(bcs[f] == "PERIODIC" for f in FACE_LOCATIONS if FACE_AXIS[f] in dom.active)
    U = None
    if variable in ("CONSERVATIVE", "CHAR-CONSERVATIVE") or riemann.startswith("FLUX-SPLITTING-"):
        U = cons_from_prim(W, gamma).clone()
        U[dom.interior] = U_int
    rhs = compute_rhs(W, dom, stencil, riemann, signal_speed, gamma, transport, variable, U, frozen)
    if stage == 0:
        return U_int + (dt * m) * rhs
    return (c0 * U_int + c1 * Un_int) + (dt * m) * rhs
