"""Exact solution of the 1-D Riemann problem for an ideal gas.

Textbook restatement of Toro, "Riemann Solvers and Numerical Methods for Fluid
Dynamics", 3rd ed., chapter 4 (pressure function 4.6-4.8, Newton iteration 4.44,
sampling 4.4-4.5).  Used only by the tests as a physics check of a full
simulation, not as a parity oracle.
"""
import numpy as np


def _f_K(p, rho_K, p_K, c_K, gamma):
    """Toro 4.6/4.7 and its derivative."""
    if p > p_K:  # shock
        A = 2.0 / ((gamma + 1.0) * rho_K)
        B = (gamma - 1.0) / (gamma + 1.0) * p_K
        sq = np.sqrt(A / (p + B))
        f = (p - p_K) * sq
        df = sq * (1.0 - 0.5 * (p - p_K) / (B + p))
    else:  # rarefaction
        pr = p / p_K
        f = 2.0 * c_K / (gamma - 1.0) * (pr ** ((gamma - 1.0) / (2.0 * gamma)) - 1.0)
        df = 1.0 / (rho_K * c_K) * pr ** (-(gamma + 1.0) / (2.0 * gamma))
    return f, df


def star_state(rho_L, u_L, p_L, rho_R, u_R, p_R, gamma):
    c_L = np.sqrt(gamma * p_L / rho_L)
    c_R = np.sqrt(gamma * p_R / rho_R)
    p = max(1e-8, 0.5 * (p_L + p_R) - 0.125 * (u_R - u_L) * (rho_L + rho_R) * (c_L + c_R))
    for _ in range(100):
        fL, dL = _f_K(p, rho_L, p_L, c_L, gamma)
        fR, dR = _f_K(p, rho_R, p_R, c_R, gamma)
        dp = (fL + fR + (u_R - u_L)) / (dL + dR)
        p_new = max(1e-12, p - dp)
        if abs(p_new - p) / (0.5 * (p_new + p)) < 1e-14:
            p = p_new
            break
        p = p_new
    fL, _ = _f_K(p, rho_L, p_L, c_L, gamma)
    fR, _ = _f_K(p, rho_R, p_R, c_R, gamma)
    u = 0.5 * (u_L + u_R) + 0.5 * (fR - fL)
    return p, u


def sample(x_over_t, rho_L, u_L, p_L, rho_R, u_R, p_R, gamma):
    """Toro section 4.5: (rho, u, p) at similarity coordinate S = x/t."""
    p_s, u_s = star_state(rho_L, u_L, p_L, rho_R, u_R, p_R, gamma)
    g = gamma
    out = np.zeros((3,) + np.shape(x_over_t))
    S = np.asarray(x_over_t, dtype=float)
    c_L = np.sqrt(g * p_L / rho_L)
    c_R = np.sqrt(g * p_R / rho_R)
    for idx, s in np.ndenumerate(S):
        if s <= u_s:  # left of contact
            if p_s > p_L:  # left shock
                S_L = u_L - c_L * np.sqrt((g + 1) / (2 * g) * p_s / p_L + (g - 1) / (2 * g))
                if s <= S_L:
                    r, u, p = rho_L, u_L, p_L
                else:
                    r = rho_L * ((p_s / p_L + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p_s / p_L + 1))
                    u, p = u_s, p_s
            else:  # left rarefaction
                S_HL = u_L - c_L
                c_sL = c_L * (p_s / p_L) ** ((g - 1) / (2 * g))
                S_TL = u_s - c_sL
                if s <= S_HL:
                    r, u, p = rho_L, u_L, p_L
                elif s >= S_TL:
                    r = rho_L * (p_s / p_L) ** (1 / g)
                    u, p = u_s, p_s
                else:
                    fac = 2 / (g + 1) + (g - 1) / ((g + 1) * c_L) * (u_L - s)
                    r = rho_L * fac ** (2 / (g - 1))
                    u = 2 / (g + 1) * (c_L + (g - 1) / 2 * u_L + s)
                    p = p_L * fac ** (2 * g / (g - 1))
        else:  # right of contact
            if p_s > p_R:  # right shock
                S_R = u_R + c_R * np.sqrt((g + 1) / (2 * g) * p_s / p_R + (g - 1) / (2 * g))
                if s >= S_R:
                    r, u, p = rho_R, u_R, p_R
                else:
                    r = rho_R * ((p_s / p_R + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p_s / p_R + 1))
                    u, p = u_s, p_s
            else:  # right rarefaction
                S_HR = u_R + c_R
                c_sR = c_R * (p_s / p_R) ** ((g - 1) / (2 * g))
                S_TR = u_s + c_sR
                if s >= S_HR:
                    r, u, p = rho_R, u_R, p_R
                elif s <= S_TR:
                    r = rho_R * (p_s / p_R) ** (1 / g)
                    u, p = u_s, p_s
                else:
                    fac = 2 / (g + 1) - (g - 1) / ((g + 1) * c_R) * (u_R - s)
                    r = rho_R * fac ** (2 / (g - 1))
                    u = 2 / (g + 1) * (-c_R + (g - 1) / 2 * u_R + s)
                    p = p_R * fac ** (2 * g / (g - 1))
        out[(0,) + idx] = r
        out[(1,) + idx] = u
        out[(2,) + idx] = p
    return out
