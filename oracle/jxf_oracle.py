"""CPU oracle: a NumPy float64 restatement of JAX-Fluids' convective-flux hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this module;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker.

Every function follows the reference expression by expression (same operands,
same association order) and cites the reference file:line it restates.  All
citations are relative to the JAX-Fluids v0.2.1 source tree
(``src/jaxfluids/...``).

Layout (same as the reference): a field buffer has shape
``(nvar, Nx+2nh, Ny+2nh, Nz+2nh)`` where an inactive axis has extent 1 and no
halo cells (``initialization/helper_functions.py:49``).  A face array along
``axis`` has extent ``N_axis+1`` along that axis and the interior extent along
the other active axes (``solvers/space_solver.py:142-146``).

Pinning: every public function here is checked in ``tests/test_oracle_*.py``
against closed forms / invariants (see the header of each test module).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Dict, Optional, Sequence, Tuple

import numpy as np

# config/precision.py:39,53 -- spatial stencil epsilon (both precisions)
SPATIAL_STENCIL_EPS = 1e-30
# config/precision.py:36,50 -- general epsilon = machine eps of the dtype
EPS64 = float(np.finfo(np.float64).eps)
EPS32 = float(np.finfo(np.float32).eps)

# equation_information.py:110
VELOCITY_MINOR_AXES = ((2, 3), (3, 1), (1, 2))
# solvers/riemann_solvers/eigendecomposition.py:69-73
VEL_ID_EIGS = ((2, 3), (3, 2), (2, 3))

FACE_LOCATIONS = ("east", "west", "north", "south", "top", "bottom")
FACE_AXIS = {"east": 0, "west": 0, "north": 1, "south": 1, "top": 2, "bottom": 2}


# --------------------------------------------------------------------------
# Domain / slices
# --------------------------------------------------------------------------
@dataclass
class Domain:
    """Uniform grid (domain/domain_information.py:68-300, homogenous.py:7-21)."""

    cells: Tuple[int, int, int]
    ranges: Tuple[Tuple[float, float], ...] = ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0))
    nh: int = 3

    def __post_init__(self):
        self.cells = tuple(int(c) for c in self.cells)
        self.active = tuple(i for i in range(3) if self.cells[i] > 1)
        # homogenous.py:13  cell size = (b - a) / n
        self.dx = tuple((float(r[1]) - float(r[0])) / n for r, n in zip(self.ranges, self.cells))
        # domain_information.py:290  1/dx
        self.one_dx = tuple(1.0 / d for d in self.dx)
        # domain_information.py:698-701 smallest cell size over active axes
        self.dx_min = min(self.dx[i] for i in self.active)

    @property
    def shape(self):
        # initialization/helper_functions.py:49
        return tuple(n + 2 * self.nh if n > 1 else 1 for n in self.cells)

    def dom(self, i):
        """Interior slice along dim i (spatial_stencil.py:24-26)."""
        return slice(self.nh, -self.nh) if i in self.active else slice(None)

    @property
    def interior(self):
        return (Ellipsis, self.dom(0), self.dom(1), self.dom(2))

    def cell_centers(self):
        """homogenous.py:13-14 (linspace of centers)."""
        out = []
        for i in range(3):
            a, b = float(self.ranges[i][0]), float(self.ranges[i][1])
            d = self.dx[i]
            out.append(np.linspace(a + d / 2, b - d / 2, self.cells[i]))
        return out

    def mesh_grid(self):
        c = self.cells_centers_active()
        return np.meshgrid(*c, indexing="ij")

    def cells_centers_active(self):
        cc = self.cell_centers()
        return [cc[i] for i in range(3)]


def shift_slice(dom: Domain, axis: int, k: int, at_cell_center: bool = False):
    """Slice of the buffer shifted by k along axis (spatial_stencil.py:84-100).

    nlo = n + k ; nhi = -n + k + 1 - at_cell_center (None if 0); other axes take
    the domain slice.  For faces the result has N_axis+1 entries along axis.
    """
    n = dom.nh
    nlo = n + k
    nhi = -n + k + 1 - int(at_cell_center)
    assert nlo >= 0
    if nhi == 0:
        nhi = None
    else:
        assert nhi < 0
    s = [dom.dom(0), dom.dom(1), dom.dom(2)]
    s[axis] = slice(nlo, nhi)
    return (Ellipsis,) + tuple(s)


# --------------------------------------------------------------------------
# WENO5 (stencils/reconstruction/shock_capturing/weno5_base.py, weno/*.py)
# --------------------------------------------------------------------------
_DR = (1 / 10, 6 / 10, 3 / 10)                  # weno5_base.py:24
_CR = ((1 / 3, -7 / 6, 11 / 6),                 # weno5_base.py:25-27
       (-1 / 6, 5 / 6, 1 / 3),
       (1 / 3, 5 / 6, -1 / 6))
# weno5_base.py:30 array_slices([range(-3,2,1), range(2,-3,-1)])
WENO5_SHIFTS = ((-3, -2, -1, 0, 1), (2, 1, 0, -1, -2))
# weno5_base.py:31 stencil_slices([range(0,5,1), range(5,0,-1)]) (window indices)
WENO5_WINDOW_IDX = ((0, 1, 2, 3, 4), (5, 4, 3, 2, 1))


def weno5_smoothness(u_imm, u_im, u_i, u_ip, u_ipp):
    """weno5_base.py:34-41 (Jiang-Shu beta_0..beta_2)."""
    beta_0 = 13.0 / 12.0 * np.square(u_imm - 2 * u_im + u_i) + 1.0 / 4.0 * np.square(u_imm - 4 * u_im + 3 * u_i)
    beta_1 = 13.0 / 12.0 * np.square(u_im - 2 * u_i + u_ip) + 1.0 / 4.0 * np.square(u_im - u_ip)
    beta_2 = 13.0 / 12.0 * np.square(u_i - 2 * u_ip + u_ipp) + 1.0 / 4.0 * np.square(3 * u_i - 4 * u_ip + u_ipp)
    return beta_0, beta_1, beta_2


def weno5_polynomials(u_imm, u_im, u_i, u_ip, u_ipp):
    """weno5_base.py:44-50 (three 3-point candidate values)."""
    p_0 = _CR[0][0] * u_imm + _CR[0][1] * u_im + _CR[0][2] * u_i
    p_1 = _CR[1][0] * u_im + _CR[1][1] * u_i + _CR[1][2] * u_ip
    p_2 = _CR[2][0] * u_i + _CR[2][1] * u_ip + _CR[2][2] * u_ipp
    return p_0, p_1, p_2


def weno5_js_point(u_imm, u_im, u_i, u_ip, u_ipp, eps=SPATIAL_STENCIL_EPS):
    """weno/weno5_js.py:39-58."""
    beta_0, beta_1, beta_2 = weno5_smoothness(u_imm, u_im, u_i, u_ip, u_ipp)
    one_beta_0_sq = 1.0 / (beta_0 * beta_0 + eps)
    one_beta_1_sq = 1.0 / (beta_1 * beta_1 + eps)
    one_beta_2_sq = 1.0 / (beta_2 * beta_2 + eps)
    alpha_0 = _DR[0] * one_beta_0_sq
    alpha_1 = _DR[1] * one_beta_1_sq
    alpha_2 = _DR[2] * one_beta_2_sq
    one_alpha = 1.0 / (alpha_0 + alpha_1 + alpha_2)
    omega_0 = alpha_0 * one_alpha
    omega_1 = alpha_1 * one_alpha
    omega_2 = alpha_2 * one_alpha
    p_0, p_1, p_2 = weno5_polynomials(u_imm, u_im, u_i, u_ip, u_ipp)
    return omega_0 * p_0 + omega_1 * p_1 + omega_2 * p_2


def weno5_z_point(u_imm, u_im, u_i, u_ip, u_ipp, eps=SPATIAL_STENCIL_EPS):
    """weno/weno5_z.py:38-59 (Borges et al. 2008)."""
    beta_0, beta_1, beta_2 = weno5_smoothness(u_imm, u_im, u_i, u_ip, u_ipp)
    tau_5 = np.abs(beta_0 - beta_2)
    alpha_z_0 = _DR[0] * (1.0 + tau_5 / (beta_0 + eps))
    alpha_z_1 = _DR[1] * (1.0 + tau_5 / (beta_1 + eps))
    alpha_z_2 = _DR[2] * (1.0 + tau_5 / (beta_2 + eps))
    one_alpha_z = 1.0 / (alpha_z_0 + alpha_z_1 + alpha_z_2)
    omega_z_0 = alpha_z_0 * one_alpha_z
    omega_z_1 = alpha_z_1 * one_alpha_z
    omega_z_2 = alpha_z_2 * one_alpha_z
    p_0, p_1, p_2 = weno5_polynomials(u_imm, u_im, u_i, u_ip, u_ipp)
    return omega_z_0 * p_0 + omega_z_1 * p_1 + omega_z_2 * p_2


# --------------------------------------------------------------------------
# Narrower and TENO stencils on the same operands.  Each takes WENO5's five operands
# (WENO5_SHIFTS / WENO5_WINDOW_IDX) and reads the ones of its own slice table: WENO3's
# array_slices([range(-2,1,1), range(1,-2,-1)]) (weno3_base.py:28) are the middle three
# (u_im, u_i, u_ip), WENO1's ([range(-1,0,1), range(0,-1,-1)]) (weno/weno1_js.py:22) is u_i;
# the characteristic windows of sizes 4 and 2 (eigendecomposition.py:75-95, half size 2 and 1)
# are the middle cells of the 6-cell one.
# --------------------------------------------------------------------------
_DR3 = (1 / 3, 2 / 3)                           # weno3_base.py:24
_CR3 = ((-0.5, 1.5), (0.5, 0.5))                # weno3_base.py:25


def weno1_point(u_imm, u_im, u_i, u_ip, u_ipp, eps=SPATIAL_STENCIL_EPS):
    """weno/weno1_js.py:25-30: the adjacent cell's value."""
    return np.array(u_i, copy=True)


def weno3_smoothness(u_im, u_i, u_ip):
    """weno3_base.py:32-37."""
    return np.square(u_i - u_im), np.square(u_ip - u_i)


def weno3_polynomials(u_im, u_i, u_ip):
    """weno3_base.py:40-45."""
    return _CR3[0][0] * u_im + _CR3[0][1] * u_i, _CR3[1][0] * u_i + _CR3[1][1] * u_ip


def weno3_js_point(u_imm, u_im, u_i, u_ip, u_ipp, eps=SPATIAL_STENCIL_EPS):
    """weno/weno3_js.py:27-40."""
    beta_0, beta_1 = weno3_smoothness(u_im, u_i, u_ip)
    one_beta_0_sq = 1.0 / (beta_0 * beta_0 + eps)
    one_beta_1_sq = 1.0 / (beta_1 * beta_1 + eps)
    alpha_0 = _DR3[0] * one_beta_0_sq
    alpha_1 = _DR3[1] * one_beta_1_sq
    one_alpha = 1.0 / (alpha_0 + alpha_1)
    p_0, p_1 = weno3_polynomials(u_im, u_i, u_ip)
    return alpha_0 * one_alpha * p_0 + alpha_1 * one_alpha * p_1


def weno3_z_point(u_imm, u_im, u_i, u_ip, u_ipp, eps=SPATIAL_STENCIL_EPS):
    """weno/weno3_z.py:27-39 (Don & Borges 2013)."""
    beta_0, beta_1 = weno3_smoothness(u_im, u_i, u_ip)
    tau_3 = np.abs(beta_0 - beta_1)
    alpha_z_0 = _DR3[0] * (1.0 + tau_3 / (beta_0 + eps))
    alpha_z_1 = _DR3[1] * (1.0 + tau_3 / (beta_1 + eps))
    one_alpha_z = 1.0 / (alpha_z_0 + alpha_z_1)
    p_0, p_1 = weno3_polynomials(u_im, u_i, u_ip)
    return alpha_z_0 * one_alpha_z * p_0 + alpha_z_1 * one_alpha_z * p_1


def _teno5_weights(beta, CT, dr, eps):
    """teno/teno5.py:44-68: gamma_k = (C + tau5/(beta_k + eps))^q, C = 1, q = 6; sharp cutoff
    gamma_k / sum gamma < CT; omega_k = d_k delta_k / (sum d delta + eps)."""
    beta_0, beta_1, beta_2 = beta
    tau_5 = np.abs(beta_0 - beta_2)
    gamma = [np.power(1.0 + tau_5 / (b + eps), 6) for b in beta]
    one_gamma_sum = 1.0 / (gamma[0] + gamma[1] + gamma[2])
    w = [dr[k] * np.where(gamma[k] * one_gamma_sum < CT, 0.0, 1.0) for k in range(3)]
    one_dk = 1.0 / (w[0] + w[1] + w[2] + eps)
    return [wk * one_dk for wk in w]


def teno5_point(u_imm, u_im, u_i, u_ip, u_ipp, eps=SPATIAL_STENCIL_EPS):
    """teno/teno5.py:32-73 (Fu et al. 2016): d = (0.05, 0.55, 0.40) (:26), CT = 1e-5 (:30)."""
    beta = weno5_smoothness(u_imm, u_im, u_i, u_ip, u_ipp)
    om = _teno5_weights(beta, 1e-5, (0.05, 0.55, 0.40), eps)
    p = weno5_polynomials(u_imm, u_im, u_i, u_ip, u_ipp)
    return om[0] * p[0] + om[1] * p[1] + om[2] * p[2]


def teno5_a_point(u_imm, u_im, u_i, u_ip, u_ipp, eps=SPATIAL_STENCIL_EPS):
    """teno/teno5_a.py:42-99: adaptive cutoff (:66-78) with xi = 1e-3, Cr = 0.24, alpha = (10, 5)
    (:29-33); the weights are WENO5Base's _dr (the class sets an unused dr_, :23)."""
    beta = weno5_smoothness(u_imm, u_im, u_i, u_ip, u_ipp)
    xi, Cr = 1e-3, 0.24
    ed = 0.9 * Cr / (1 - Cr) * xi ** 2
    eta_0 = (np.abs(2 * (u_i - u_im) * (u_im - u_imm)) + ed) / ((u_i - u_im) ** 2 + (u_im - u_imm) ** 2 + ed)
    eta_1 = (np.abs(2 * (u_ip - u_i) * (u_i - u_im)) + ed) / ((u_ip - u_i) ** 2 + (u_i - u_im) ** 2 + ed)
    eta_2 = (np.abs(2 * (u_ipp - u_ip) * (u_ip - u_i)) + ed) / ((u_ipp - u_ip) ** 2 + (u_ip - u_i) ** 2 + ed)
    eta = np.minimum(eta_0, np.minimum(eta_1, eta_2))
    m = 1 - np.minimum(1.0, eta / Cr)
    g = np.power(1 - m, 4) * (1 + 4 * m)
    beta_bar = np.ceil(10.0 - 5.0 * (1 - g)) - 1.0
    CT = np.power(10.0, -beta_bar)
    om = _teno5_weights(beta, CT, _DR, eps)
    p = weno5_polynomials(u_imm, u_im, u_i, u_ip, u_ipp)
    return om[0] * p[0] + om[1] * p[1] + om[2] * p[2]


WENO5_POINT = {"WENO5-JS": weno5_js_point, "WENO5-Z": weno5_z_point, "WENO1": weno1_point,
               "WENO3-JS": weno3_js_point, "WENO3-Z": weno3_z_point, "TENO5": teno5_point,
               "TENO5-A": teno5_a_point}


# --------------------------------------------------------------------------
# WENO7-JS (weno/weno7_js.py:10-90; Balsara & Shu 2000): array_slices([range(-4,3,1),
# range(3,-4,-1)]) -- seven of the eight window cells -4..3, mirrored for the right state; four
# 4-point candidates; indicators with integer coefficients (no common factor), evaluated in the
# reference's operation order.
# --------------------------------------------------------------------------
WENO7_SHIFTS = ((-4, -3, -2, -1, 0, 1, 2), (3, 2, 1, 0, -1, -2, -3))
WENO7_WINDOW_IDX = ((0, 1, 2, 3, 4, 5, 6), (7, 6, 5, 4, 3, 2, 1))   # stencil_slices (:31)
_DR7 = (1 / 35, 12 / 35, 18 / 35, 4 / 35)                          # :20
_CR7 = ((-1 / 4, 13 / 12, -23 / 12, 25 / 12), (1 / 12, -5 / 12, 13 / 12, 1 / 4),
        (-1 / 12, 7 / 12, 7 / 12, -1 / 12), (1 / 4, 13 / 12, -5 / 12, 1 / 12))   # :22-27


def weno7_js_point(b0_, b1_, b2_, b3_, b4_, b5_, b6_, eps=SPATIAL_STENCIL_EPS):
    """weno/weno7_js.py:33-90 on the seven operands buffer[s1_[0..6]]."""
    u = (b0_, b1_, b2_, b3_, b4_, b5_, b6_)
    beta_0 = u[0] * (547 * u[0] - 3882 * u[1] + 4642 * u[2] - 1854 * u[3]) \
        + u[1] * (7043 * u[1] - 17246 * u[2] + 7042 * u[3]) \
        + u[2] * (11003 * u[2] - 9402 * u[3]) \
        + u[3] * (2107 * u[3])
    beta_1 = u[1] * (267 * u[1] - 1642 * u[2] + 1602 * u[3] - 494 * u[4]) \
        + u[2] * (2843 * u[2] - 5966 * u[3] + 1922 * u[4]) \
        + u[3] * (3443 * u[3] - 2522 * u[4]) \
        + u[4] * (547 * u[4])
    beta_2 = u[2] * (547 * u[2] - 2522 * u[3] + 1922 * u[4] - 494 * u[5]) \
        + u[3] * (3443 * u[3] - 5966 * u[4] + 1602 * u[5]) \
        + u[4] * (2843 * u[4] - 1642 * u[5]) \
        + u[5] * (267 * u[5])
    beta_3 = u[3] * (2107 * u[3] - 9402 * u[4] + 7042 * u[5] - 1854 * u[6]) \
        + u[4] * (11003 * u[4] - 17246 * u[5] + 4642 * u[6]) \
        + u[5] * (7043 * u[5] - 3882 * u[6]) \
        + u[6] * (547 * u[6])
    one = [1.0 / (b * b + eps) for b in (beta_0, beta_1, beta_2, beta_3)]
    alpha = [_DR7[k] * one[k] for k in range(4)]
    one_alpha = 1.0 / (alpha[0] + alpha[1] + alpha[2] + alpha[3])
    omega = [a * one_alpha for a in alpha]
    p = [_CR7[k][0] * u[k] + _CR7[k][1] * u[k + 1] + _CR7[k][2] * u[k + 2] + _CR7[k][3] * u[k + 3]
         for k in range(4)]
    return omega[0] * p[0] + omega[1] * p[1] + omega[2] * p[2] + omega[3] * p[3]


# --------------------------------------------------------------------------
# Six-point stencils (weno6_base.py, teno6_base.py): array_slices([range(-3,3,1),
# range(2,-4,-1)]) -- the whole 6-cell window, mirrored for the right state.  The smoothness
# indicator of the 6-point stencil is a quadratic form with 8-digit integer coefficients whose
# terms cancel; it is evaluated in the reference's operation order.
# --------------------------------------------------------------------------
WENO6_SHIFTS = ((-3, -2, -1, 0, 1, 2), (2, 1, 0, -1, -2, -3))
WENO6_WINDOW_IDX = ((0, 1, 2, 3, 4, 5), (5, 4, 3, 2, 1, 0))
_CR6 = ((1 / 3, -7 / 6, 11 / 6), (-1 / 6, 5 / 6, 1 / 3), (1 / 3, 5 / 6, -1 / 6), (11 / 6, -7 / 6, 1 / 3))  # weno6_base.py:23-26


def _beta_six_point(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp):
    """weno6_base.py:37-45 (beta_3) = teno6_base.py:45-53 (beta_6)."""
    return 1.0 / 10080 / 12 * (
        271779 * u_imm * u_imm +
        u_imm * (-2380800 * u_im + 4086352 * u_i - 3462252 * u_ip + 1458762 * u_ipp - 245620 * u_ippp) +
        u_im * (5653317 * u_im - 20427884 * u_i + 17905032 * u_ip - 7727988 * u_ipp + 1325006 * u_ippp) +
        u_i * (19510972 * u_i - 35817664 * u_ip + 15929912 * u_ipp - 2792660 * u_ippp) +
        u_ip * (17195652 * u_ip - 15880404 * u_ipp + 2863984 * u_ippp) +
        u_ipp * (3824847 * u_ipp - 1429976 * u_ippp) +
        139633 * u_ippp * u_ippp)


def _beta_three_point(u_imm, u_im, u_i, u_ip, u_ipp):
    """weno6_base.py:34-36 (the WENO5 indicators in the reference's form)."""
    beta_0 = 13.0 / 12.0 * np.square(u_imm - 2 * u_im + u_i) + 1.0 / 4.0 * np.square(u_imm - 4 * u_im + 3 * u_i)
    beta_1 = 13.0 / 12.0 * np.square(u_im - 2 * u_i + u_ip) + 1.0 / 4.0 * np.square(u_im - u_ip)
    beta_2 = 13.0 / 12.0 * np.square(u_i - 2 * u_ip + u_ipp) + 1.0 / 4.0 * np.square(3 * u_i - 4 * u_ip + u_ipp)
    return beta_0, beta_1, beta_2


def weno6_cu_point(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp, eps=SPATIAL_STENCIL_EPS, dx=None):
    """weno/weno6_cu.py:24-61 (Hu, Wang & Adams 2010): d = (1/20, 9/20, 9/20, 1/20)
    (weno6_base.py:22), C = 20 (:16), tau_6 = beta_3 - (beta_0 + 4 beta_1 + beta_2)/6."""
    beta_0, beta_1, beta_2 = _beta_three_point(u_imm, u_im, u_i, u_ip, u_ipp)
    beta_3 = _beta_six_point(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp)
    tau_6 = beta_3 - 1 / 6 * (beta_0 + 4 * beta_1 + beta_2)
    dr, C = (1 / 20, 9 / 20, 9 / 20, 1 / 20), 20
    alpha = [dr[k] * (C + tau_6 / (b + eps)) for k, b in enumerate((beta_0, beta_1, beta_2, beta_3))]
    one_alpha = 1.0 / (alpha[0] + alpha[1] + alpha[2] + alpha[3])
    p_0 = _CR6[0][0] * u_imm + _CR6[0][1] * u_im + _CR6[0][2] * u_i
    p_1 = _CR6[1][0] * u_im + _CR6[1][1] * u_i + _CR6[1][2] * u_ip
    p_2 = _CR6[2][0] * u_i + _CR6[2][1] * u_ip + _CR6[2][2] * u_ipp
    p_3 = _CR6[3][0] * u_ip + _CR6[3][1] * u_ipp + _CR6[3][2] * u_ippp
    return (alpha[0] * one_alpha * p_0 + alpha[1] * one_alpha * p_1 + alpha[2] * one_alpha * p_2
            + alpha[3] * one_alpha * p_3)


def teno6_point(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp, eps=SPATIAL_STENCIL_EPS, dx=None):
    """teno/teno6.py:27-67 with teno6_base.py:20-64: three 3-point candidates and one 4-point
    candidate (3/12, 13/12, -5/12, 1/12) with its own indicator beta_3 (:40-44), the 6-point
    indicator beta_6 for tau_6 = |beta_6 - (beta_0 + 4 beta_1 + beta_2)/6|, gamma_k =
    (1 + tau_6/(beta_k + eps))^6, cutoff CT = 1e-7, d = (0.05, 0.45, 0.30, 0.20)."""
    beta_0, beta_1, beta_2 = _beta_three_point(u_imm, u_im, u_i, u_ip, u_ipp)
    beta_3 = 1.0 / 240.0 * (
        u_i * (2107 * u_i - 9402 * u_ip + 7042 * u_ipp - 1854 * u_ippp)
        + u_ip * (11003 * u_ip - 17246 * u_ipp + 4642 * u_ippp)
        + u_ipp * (7043 * u_ipp - 3882 * u_ippp)
        + 547 * u_ippp * u_ippp)
    beta_6 = _beta_six_point(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp)
    tau_6 = np.abs(beta_6 - 1 / 6 * (beta_0 + 4 * beta_1 + beta_2))
    gamma = [_integer_pow6(1.0 + tau_6 / (b + eps)) for b in (beta_0, beta_1, beta_2, beta_3)]
    one_gamma_sum = 1.0 / (gamma[0] + gamma[1] + gamma[2] + gamma[3])
    dr = (0.050, 0.450, 0.300, 0.200)
    w = [dr[k] * np.where(gamma[k] * one_gamma_sum < 1e-7, 0.0, 1.0) for k in range(4)]
    one_dk = 1.0 / (w[0] + w[1] + w[2] + w[3] + eps)
    p_0 = _CR6[0][0] * u_imm + _CR6[0][1] * u_im + _CR6[0][2] * u_i
    p_1 = _CR6[1][0] * u_im + _CR6[1][1] * u_i + _CR6[1][2] * u_ip
    p_2 = _CR6[2][0] * u_i + _CR6[2][1] * u_ip + _CR6[2][2] * u_ipp
    p_3 = 3 / 12 * u_i + 13 / 12 * u_ip + -5 / 12 * u_ipp + 1 / 12 * u_ippp
    return (w[0] * one_dk * p_0 + w[1] * one_dk * p_1 + w[2] * one_dk * p_2 + w[3] * one_dk * p_3)


def _integer_pow4(x):
    """jnp.power(x, 4) with an integer exponent: lax.integer_pow, by squaring."""
    x2 = x * x
    return x2 * x2


def _integer_pow6(x):
    """x ** 6 with the Python int q = 6 (teno6.py:52-55, teno6_a.py:71-74): lax.integer_pow by
    binary powering, x^2 * (x^2 * x^2)."""
    x2 = x * x
    return x2 * (x2 * x2)


def _weno6_cum(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp, alpha_of):
    beta_0, beta_1, beta_2 = _beta_three_point(u_imm, u_im, u_i, u_ip, u_ipp)
    beta_3 = _beta_six_point(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp)
    dr = (1 / 20, 9 / 20, 9 / 20, 1 / 20)                                  # weno6_base.py:22
    alpha = [dr[k] * _integer_pow4(alpha_of(beta_0, beta_1, beta_2, beta_3, b))
             for k, b in enumerate((beta_0, beta_1, beta_2, beta_3))]
    one_alpha = 1.0 / (alpha[0] + alpha[1] + alpha[2] + alpha[3])
    p_0 = _CR6[0][0] * u_imm + _CR6[0][1] * u_im + _CR6[0][2] * u_i
    p_1 = _CR6[1][0] * u_im + _CR6[1][1] * u_i + _CR6[1][2] * u_ip
    p_2 = _CR6[2][0] * u_i + _CR6[2][1] * u_ip + _CR6[2][2] * u_ipp
    p_3 = _CR6[3][0] * u_ip + _CR6[3][1] * u_ipp + _CR6[3][2] * u_ippp
    return (alpha[0] * one_alpha * p_0 + alpha[1] * one_alpha * p_1 + alpha[2] * one_alpha * p_2
            + alpha[3] * one_alpha * p_3)


def weno6_cum1_point(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp, eps=SPATIAL_STENCIL_EPS, dx=1.0):
    """weno/weno6_cum1.py:28-67 (Hu & Adams 2011): C_q = 1000, q = 4, tau_6 as WENO6-CU,
    alpha_k = d_k (C_q + tau_6 / (beta_k + eps dx^2))^4 with the stencil epsilon and the cell size
    along the axis (its eps_ = 1e-8 is not read)."""
    dx2 = dx * dx

    def alpha_of(beta_0, beta_1, beta_2, beta_3, b):
        tau_6 = beta_3 - 1 / 6 * (beta_0 + 4 * beta_1 + beta_2)
        return 1000 + tau_6 / (b + eps * dx2)
    return _weno6_cum(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp, alpha_of)


def weno6_cum2_point(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp, eps=SPATIAL_STENCIL_EPS, dx=1.0):
    """weno/weno6_cum2.py:29-72: C_q = 1000, q = 4, eps_ = 1e-8, chi = 1/eps_,
    beta_ave = (beta_0 + beta_2 + 4 beta_1)/6, tau_6 = beta_3 - beta_ave,
    alpha_k = d_k (C_q + tau_6/(beta_k + eps_ dx^2) (beta_ave + chi dx^2)/(beta_k + chi dx^2))^4."""
    dx2 = dx * dx
    eps_ = 1e-8
    chi = 1.0 / eps_

    def alpha_of(beta_0, beta_1, beta_2, beta_3, b):
        beta_ave = 1 / 6 * (beta_0 + beta_2 + 4 * beta_1)
        tau_6 = beta_3 - beta_ave
        return 1000 + tau_6 / (b + eps_ * dx2) * (beta_ave + chi * dx2) / (b + chi * dx2)
    return _weno6_cum(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp, alpha_of)


def teno6_a_point(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp, eps=SPATIAL_STENCIL_EPS, dx=None):
    """teno/teno6_a.py:20-131: TENO6's candidates and indicators with |beta_3|, |beta_6|
    (is_positivity_limiter_smoothness, :36,58-60), C = 1, q = 6, the spectrally optimised
    d = (0.0855682281039113, 0.4294317718960898, 0.1727270875843552, 0.3122729124156450) (:25-28),
    and the adaptive cutoff CT = 10^-(ceil(10.5 - 4.5 (1 - g)) - 1) from the minimum eta of the four
    neighbouring difference pairs (:96-131) with xi = 1e-3, Cr = 0.17 (:31-35); no eps in the
    weight normalisation (:81)."""
    beta_0, beta_1, beta_2 = _beta_three_point(u_imm, u_im, u_i, u_ip, u_ipp)
    beta_3 = np.abs(1.0 / 240.0 * (
        u_i * (2107 * u_i - 9402 * u_ip + 7042 * u_ipp - 1854 * u_ippp)
        + u_ip * (11003 * u_ip - 17246 * u_ipp + 4642 * u_ippp)
        + u_ipp * (7043 * u_ipp - 3882 * u_ippp)
        + 547 * u_ippp * u_ippp))
    beta_6 = np.abs(_beta_six_point(u_imm, u_im, u_i, u_ip, u_ipp, u_ippp))
    tau_6 = np.abs(beta_6 - 1 / 6 * (beta_0 + 4 * beta_1 + beta_2))
    gamma = [_integer_pow6(1.0 + tau_6 / (b + eps)) for b in (beta_0, beta_1, beta_2, beta_3)]
    one_gamma_sum = 1.0 / (gamma[0] + gamma[1] + gamma[2] + gamma[3])
    xi, Cr = 1e-3, 0.17
    ed = 0.9 * Cr / (1 - Cr) * xi ** 2
    d = (u_im - u_imm, u_i - u_im, u_ip - u_i, u_ipp - u_ip, u_ippp - u_ipp)
    eta = None
    for k in range(4):
        e_k = (np.abs(2 * d[k + 1] * d[k]) + ed) / (d[k + 1] ** 2 + d[k] ** 2 + ed)
        eta = e_k if eta is None else np.minimum(eta, e_k)
    m = 1 - np.minimum(1.0, eta / Cr)
    g = np.power(1 - m, 4) * (1 + 4 * m)
    beta_bar = np.ceil(10.5 - 4.5 * (1 - g)) - 1.0
    CT = np.power(10.0, -beta_bar)
    dr = (0.0855682281039113, 0.4294317718960898, 0.1727270875843552, 0.3122729124156450)
    w = [dr[k] * np.where(gamma[k] * one_gamma_sum < CT, 0.0, 1.0) for k in range(4)]
    one_dk = 1.0 / (w[0] + w[1] + w[2] + w[3])
    p_0 = _CR6[0][0] * u_imm + _CR6[0][1] * u_im + _CR6[0][2] * u_i
    p_1 = _CR6[1][0] * u_im + _CR6[1][1] * u_i + _CR6[1][2] * u_ip
    p_2 = _CR6[2][0] * u_i + _CR6[2][1] * u_ip + _CR6[2][2] * u_ipp
    p_3 = 3 / 12 * u_i + 13 / 12 * u_ip + -5 / 12 * u_ipp + 1 / 12 * u_ippp
    return (w[0] * one_dk * p_0 + w[1] * one_dk * p_1 + w[2] * one_dk * p_2 + w[3] * one_dk * p_3)


WENO6_POINT = {"WENO6-CU": weno6_cu_point, "TENO6": teno6_point, "WENO6-CUM1": weno6_cum1_point,
               "WENO6-CUM2": weno6_cum2_point, "TENO6-A": teno6_a_point}


def weno5_reconstruct_xi(buffer, dom: Domain, axis: int, j: int, stencil: str = "WENO5-JS",
                         eps=SPATIAL_STENCIL_EPS):
    """SpatialReconstruction.reconstruct_xi for WENO5 on a cell buffer.

    weno/weno5_js.py:23-37 / weno5_z.py:24-36: the five operands are the buffer
    shifted by WENO5_SHIFTS[j] along axis (spatial_stencil.py:45-112).
    """
    if stencil == "WENO7-JS":
        return weno7_js_point(*[buffer[shift_slice(dom, axis, k)] for k in WENO7_SHIFTS[j]], eps=eps)
    if stencil in WENO6_POINT:
        ops = [buffer[shift_slice(dom, axis, k)] for k in WENO6_SHIFTS[j]]
        return WENO6_POINT[stencil](*ops, eps=eps, dx=dom.dx[axis])  # dx=cell_sizes[axis]
    ops = [buffer[shift_slice(dom, axis, k)] for k in WENO5_SHIFTS[j]]
    return WENO5_POINT[stencil](*ops, eps=eps)


def weno5_reconstruct_window(window, j: int, stencil: str, eps=SPATIAL_STENCIL_EPS, dx=1.0):
    """reconstruct_xi after set_slices_stencil (spatial_reconstruction.py:133-153).

    ``window`` has the 6-cell stencil window on axis 1: shape (nvar, 6, *face); dx is the cell
    size along the axis (read by WENO6-CUM1/-CUM2).
    """
    if stencil == "WENO7-JS":  # the 8-cell window
        return weno7_js_point(*[window[:, i] for i in WENO7_WINDOW_IDX[j]], eps=eps)
    if stencil in WENO6_POINT:
        ops = [window[:, i] for i in WENO6_WINDOW_IDX[j]]
        return WENO6_POINT[stencil](*ops, eps=eps, dx=dx)
    ops = [window[:, i] for i in WENO5_WINDOW_IDX[j]]
    return WENO5_POINT[stencil](*ops, eps=eps)


# --------------------------------------------------------------------------
# Ideal gas + equation manager
# --------------------------------------------------------------------------
def speed_of_sound(p, rho, gamma):
    """materials/single_materials/ideal_gas.py:69-71."""
    return np.sqrt(gamma * p / rho)


def cons_from_prim(W, gamma):
    """equation_manager.py:93-101 with ideal_gas.py:85-88 (e = p/(rho(gamma-1)))."""
    rho = W[0]
    vel = W[1:4]
    e = W[4] / (rho * (gamma - 1.0))
    mom = rho * vel
    # math/sum_consistent.py:22-24 (plain a+b+c)
    E = rho * (0.5 * (np.square(vel[0]) + np.square(vel[1]) + np.square(vel[2])) + e)
    return np.stack([rho, mom[0], mom[1], mom[2], E], axis=0)


def prim_from_cons(U, gamma):
    """equation_manager.py:164-171 with ideal_gas.py:73-75 (p=(gamma-1) e rho)."""
    rho = U[0]
    one_rho = 1.0 / rho
    vel = U[1:4] * one_rho
    e = U[4] * one_rho - 0.5 * (np.square(vel[0]) + np.square(vel[1]) + np.square(vel[2]))
    p = (gamma - 1.0) * e * rho
    return np.stack([rho, vel[0], vel[1], vel[2], p], axis=0)


def physical_flux_xi(W, U, axis):
    """equation_manager.py:237-252 (single phase)."""
    rho_ui = U[axis + 1]
    f = [rho_ui, U[axis + 1] * W[1], U[axis + 1] * W[2], U[axis + 1] * W[3], None]
    f[axis + 1] = f[axis + 1] + W[4]
    f[4] = W[axis + 1] * (U[4] + W[4])
    return np.stack(f, axis=0)


# --------------------------------------------------------------------------
# Eigendecomposition / characteristic projection
# (solvers/riemann_solvers/eigendecomposition.py)
# --------------------------------------------------------------------------
def total_enthalpy(p, rho, vel, gamma):
    """ideal_gas.py:90-110."""
    E = p / (gamma - 1) + 0.5 * rho * (np.square(vel[0]) + np.square(vel[1]) + np.square(vel[2]))
    return (E + p) / rho


def frozen_state(W_L, W_R, gamma, frozen: str = "ARITHMETIC"):
    """compute_frozen_state, single phase (eigendecomposition.py:139-281).

    Returns primes_ave, enthalpy_ave, grueneisen_ave, c_ave, cc_ave, velocity_square.
    """
    if frozen == "ARITHMETIC":
        primes_ave = 0.5 * (W_L + W_R)                                         # :148
        grueneisen_ave = (gamma - 1) * np.ones_like(primes_ave[0])             # :216-218
        enthalpy_ave = total_enthalpy(primes_ave[4], primes_ave[0], primes_ave[1:4], gamma)  # :219-222
        c_ave = speed_of_sound(primes_ave[4], primes_ave[0], gamma)             # :223-225
        cc_ave = c_ave * c_ave                                                  # :230
        velocity_square = np.square(primes_ave[1]) + np.square(primes_ave[2]) + np.square(primes_ave[3])  # :231
    elif frozen == "ROE":
        sL, sR = np.sqrt(W_L[0]), np.sqrt(W_R[0])
        primes_ave = (np.sqrt(W_L[0]) * W_L + np.sqrt(W_R[0]) * W_R) / (np.sqrt(W_L[0]) + np.sqrt(W_R[0]))  # :293
        primes_ave = primes_ave.copy()
        primes_ave[0] = np.sqrt(W_L[0] * W_R[0])                                # :237
        rho_div = 1.0 / (sL + sR)                                               # :240
        H_L = total_enthalpy(W_L[4], W_L[0], W_L[1:4], gamma)
        H_R = total_enthalpy(W_R[4], W_R[0], W_R[1:4], gamma)
        enthalpy_ave = (sL * H_L + sR * H_R) * rho_div                          # :250
        psi_L, psi_R = W_L[4] / W_L[0], W_R[4] / W_R[0]                         # ideal_gas.py:61-63
        psi_ave = (sL * psi_L + sR * psi_R) * rho_div                           # :258
        g = gamma - 1
        grueneisen_ave = (sL * g + sR * g) * rho_div                            # :260-262
        dvel = W_R[1:4] - W_L[1:4]
        sq_dvel = np.square(dvel[0]) + np.square(dvel[1]) + np.square(dvel[2])  # :265
        p_over_rho_ave = (sL * W_L[4] / W_L[0] + sR * W_R[4] / W_R[0]) * rho_div \
            + 0.5 * primes_ave[0] * rho_div * rho_div * sq_dvel                 # :267-268
        velocity_square = np.square(primes_ave[1]) + np.square(primes_ave[2]) + np.square(primes_ave[3])  # :271
        cc_ave = psi_ave + grueneisen_ave * p_over_rho_ave                      # :274
        c_ave = np.sqrt(cc_ave)                                                 # :275
    else:
        raise NotImplementedError(frozen)
    return primes_ave, enthalpy_ave, grueneisen_ave, c_ave, cc_ave, velocity_square


def char_from_prim(xin, primes_ave, c_ave, cc_ave, axis):
    """get_characteristics_from_primitives, single phase (eigendecomposition.py:421-433).

    ``xin`` is (5, ...); the frozen state broadcasts against it.
    """
    vel = axis + 1
    m0, m1 = VELOCITY_MINOR_AXES[axis]
    e0, e1 = VEL_ID_EIGS[axis]
    xout = [None] * 5
    xout[0] = -0.5 / c_ave * xin[vel] + 0.5 / (cc_ave * primes_ave[0]) * xin[4]
    xout[1] = xin[0] - 1.0 / cc_ave * xin[4]
    xout[e0] = xin[m0]
    xout[e1] = xin[m1]
    xout[4] = 0.5 / c_ave * xin[vel] + 0.5 / (cc_ave * primes_ave[0]) * xin[4]
    return np.stack(xout, axis=0)


def prim_from_char(xin, primes_ave, c_ave, cc_ave, axis):
    """get_primitives_from_characteristics, single phase (eigendecomposition.py:512-523)."""
    vel = axis + 1
    m0, m1 = VELOCITY_MINOR_AXES[axis]
    e0, e1 = VEL_ID_EIGS[axis]
    xout = [None] * 5
    xout[0] = primes_ave[0] * (xin[0] + xin[4]) + xin[1]
    xout[vel] = c_ave * (-xin[0] + xin[4])
    xout[m0] = xin[e0]
    xout[m1] = xin[e1]
    xout[4] = cc_ave * primes_ave[0] * (xin[0] + xin[4])
    return np.stack(xout, axis=0)


def eigen_conservatives(primes_ave, enthalpy_ave, grueneisen_ave, c_ave, cc_ave, velocity_square, axis):
    """eigendecomposition_conservatives, single phase (eigendecomposition.py:595-666).

    Returns (R, L), each of shape (5, 5, ...) indexed [row][col].
    """
    one_cc_ave = 1.0 / cc_ave
    one_rho_ave = 1.0 / primes_ave[0]
    z = np.zeros_like(primes_ave[0])
    o = np.ones_like(primes_ave[0])
    R = [[z for _ in range(5)] for _ in range(5)]
    L = [[z for _ in range(5)] for _ in range(5)]
    v = axis + 1
    m0, m1 = VELOCITY_MINOR_AXES[axis]
    P = primes_ave
    g = grueneisen_ave
    H = enthalpy_ave
    # RIGHT (:608-631)
    R[0][0] = o
    R[v][0] = P[v] - c_ave
    R[m0][0] = P[m0]
    R[m1][0] = P[m1]
    R[4][0] = H - P[v] * c_ave
    R[0][v] = g
    R[1][v] = g * P[1]
    R[2][v] = g * P[2]
    R[3][v] = g * P[3]
    R[4][v] = g * H - cc_ave
    R[m0][m0] = -P[0]
    R[4][m0] = -P[0] * P[m0]
    R[m1][m1] = P[0]
    R[4][m1] = P[0] * P[m1]
    R[0][4] = o
    R[v][4] = P[v] + c_ave
    R[m0][4] = P[m0]
    R[m1][4] = P[m1]
    R[4][4] = H + P[v] * c_ave
    # LEFT (:635-658)
    L[0][0] = 0.5 * one_cc_ave * (g * velocity_square - g * H + (P[v] + c_ave) * c_ave)
    L[0][v] = 0.5 * one_cc_ave * (-P[v] * g - c_ave)
    L[0][m0] = 0.5 * one_cc_ave * (-P[m0] * g)
    L[0][m1] = 0.5 * one_cc_ave * (-P[m1] * g)
    L[0][4] = 0.5 * one_cc_ave * g
    L[v][0] = one_cc_ave * (H - velocity_square)
    L[v][1] = P[1] * one_cc_ave
    L[v][2] = P[2] * one_cc_ave
    L[v][3] = P[3] * one_cc_ave
    L[v][4] = -one_cc_ave
    L[m0][0] = P[m0] * one_rho_ave
    L[m0][m0] = -one_rho_ave
    L[m1][0] = -P[m1] * one_rho_ave
    L[m1][m1] = one_rho_ave
    L[4][0] = 0.5 * one_cc_ave * (g * velocity_square - g * H - (P[v] - c_ave) * c_ave)
    L[4][v] = 0.5 * one_cc_ave * (-P[v] * g + c_ave)
    L[4][m0] = 0.5 * one_cc_ave * (-P[m0] * g)
    L[4][m1] = 0.5 * one_cc_ave * (-P[m1] * g)
    L[4][4] = 0.5 * one_cc_ave * g
    return np.array(R), np.array(L)


def transform(M, x):
    """einsum 'ij...,j...->i...' (eigendecomposition.py:717-742), summed in j order."""
    out = []
    for i in range(5):
        acc = M[i][0] * x[0]
        for jj in range(1, 5):
            acc = acc + M[i][jj] * x[jj]
        out.append(acc)
    return np.stack(out, axis=0)


def stencil_window(buffer, dom: Domain, axis: int, size: int = 6):
    """get_stencil_window (eigendecomposition.py:75-118): cells f-size/2..f+size/2-1 of face f
    (half_stencil_size = stencil_size // 2; 6 cells, 8 for WENO7-JS).

    Returned with the window on axis 1: (nvar, size, *face_shape).
    """
    h = size // 2
    return np.stack([buffer[shift_slice(dom, axis, k)] for k in range(-h, h)], axis=1)


def window_size(stencil: str) -> int:
    """The reconstruction stencil's _stencil_size as the eigendecomposition window uses it: 8 for
    WENO7-JS (weno7_js.py:29); the narrower stencils read the middle cells of the 6-cell window."""
    return 8 if stencil == "WENO7-JS" else 6


def face_LR_cells(buffer, dom: Domain, axis: int):
    """slice_LR (eigendecomposition.py:98-102): cells left/right of each face."""
    return buffer[shift_slice(dom, axis, -1)], buffer[shift_slice(dom, axis, 0)]


# --------------------------------------------------------------------------
# Signal speeds and Riemann solvers (solvers/riemann_solvers/)
# --------------------------------------------------------------------------
def signal_speed_arithmetic(u_L, u_R, a_L, a_R, **kw):
    """signal_speeds.py:10-29."""
    u_mean = 0.5 * (u_L + u_R)
    a_mean = 0.5 * (a_L + a_R)
    return np.minimum(u_mean - a_mean, u_L - a_L), np.maximum(u_mean + a_mean, u_R + a_R)


def signal_speed_rusanov(u_L, u_R, a_L, a_R, **kw):
    """signal_speeds.py:31-49."""
    S_plus = np.maximum(np.abs(u_L) + a_L, np.abs(u_R) + a_R)
    return -S_plus, S_plus


def signal_speed_davis(u_L, u_R, a_L, a_R, **kw):
    """signal_speeds.py:51-69 (Toro 10.48)."""
    return np.minimum(u_L - a_L, u_R - a_R), np.maximum(u_L + a_L, u_R + a_R)


def signal_speed_davis_2(u_L, u_R, a_L, a_R, rho_L, rho_R, H_L, H_R, gamma, **kw):
    """signal_speeds.py:71-107: u_Roe -/+ a_Roe from the Roe averages of u and the total
    enthalpy H.  HLLC/HLL pass p_L/p_R, not H_L/H_R, so the reference cannot use it inside a
    Riemann solver (its H_L/H_R are missing there); it is a plain function only."""
    sqrt_rho_L = np.sqrt(rho_L)
    sqrt_rho_R = np.sqrt(rho_R)
    one_dens = 1.0 / (sqrt_rho_L + sqrt_rho_R)
    u_Roe = (sqrt_rho_L * u_L + sqrt_rho_R * u_R) * one_dens
    H_Roe = (sqrt_rho_L * H_L + sqrt_rho_R * H_R) * one_dens
    a_Roe = np.sqrt((gamma - 1) * (H_Roe - 0.5 * u_Roe * u_Roe))
    return u_Roe - a_Roe, u_Roe + a_Roe


def signal_speed_einfeldt(u_L, u_R, a_L, a_R, rho_L, rho_R, **kw):
    """signal_speeds.py:109-133 (Toro 10.52-10.54)."""
    sqrt_rho_L = np.sqrt(rho_L)
    sqrt_rho_R = np.sqrt(rho_R)
    one_dens = 1.0 / (sqrt_rho_L + sqrt_rho_R)
    eta2 = 0.5 * sqrt_rho_L * sqrt_rho_R * one_dens * one_dens
    u_bar = (sqrt_rho_L * u_L + sqrt_rho_R * u_R) * one_dens
    d_bar = np.sqrt((sqrt_rho_L * a_L * a_L + sqrt_rho_R * a_R * a_R) * one_dens + eta2 * np.square(u_R - u_L))
    return np.minimum(u_bar - d_bar, u_L - a_L), np.maximum(u_bar + d_bar, u_R + a_R)


def estimate_pressure(u_L, u_R, a_L, a_R, rho_L, rho_R, p_L, p_R):
    """signal_speeds.py:201-231 (Toro 9.28)."""
    rho_bar = 0.5 * (rho_L + rho_R)
    a_bar = 0.5 * (a_L + a_R)
    p_pvrs = 0.5 * (p_L + p_R) - 0.5 * (u_R - u_L) * rho_bar * a_bar
    return np.maximum(0.0, p_pvrs)


def signal_speed_toro(u_L, u_R, a_L, a_R, rho_L, rho_R, p_L, p_R, gamma, **kw):
    """signal_speeds.py:135-157 (Toro 10.59-10.60)."""
    p_star = estimate_pressure(u_L, u_R, a_L, a_R, rho_L, rho_R, p_L, p_R)
    q_L = toro_q(p_star, p_L, gamma)
    q_R = toro_q(p_star, p_R, gamma)
    return u_L - a_L * q_L, u_R + a_R * q_R


def toro_q(p_star, p_K, gamma):
    """signal_speeds.py:152-154: Toro's wave-speed factor q_K (10.61): 1 for a rarefaction
    (p* <= p_K), sqrt(1 + (gamma+1)/(2 gamma) (p*/p_K - 1)) for a shock."""
    gamma_ = (gamma + 1) * 0.5 / gamma
    with np.errstate(invalid="ignore"):
        return 1.0 * (p_star <= p_K) + np.sqrt(1 + gamma_ * (p_star / p_K - 1)) * (p_star > p_K)


SIGNAL_SPEEDS = {
    "ARITHMETIC": signal_speed_arithmetic,
    "RUSANOV": signal_speed_rusanov,
    "DAVIS": signal_speed_davis,
    "DAVIS2": signal_speed_davis_2,
    "EINFELDT": signal_speed_einfeldt,
    "TORO": signal_speed_toro,
}


def compute_sstar(u_L, u_R, p_L, p_R, rho_L, rho_R, S_L, S_R):
    """signal_speeds.py:159-199 (Toro 10.70)."""
    rho_delta_u_L = rho_L * (S_L - u_L)
    rho_delta_u_R = rho_R * (S_R - u_R)
    return ((p_R - p_L) + (u_L * rho_delta_u_L - u_R * rho_delta_u_R)) / (rho_delta_u_L - rho_delta_u_R)


def _hllc_flux_star(W, U, S_K, S_star, axis, side):
    """HLLC._compute_flux_star_K_single_phase (HLLC.py:41-78)."""
    v = axis + 1
    m0, m1 = VELOCITY_MINOR_AXES[axis]
    pre = (S_K - W[v]) / (S_K - S_star) * W[0]                                    # :52
    u_star = [pre, pre, pre, pre,
              pre * (U[4] / U[0] + (S_star - W[v]) * (S_star + W[4] / W[0] / (S_K - W[v])))]  # :54-60
    u_star[v] = u_star[v] * S_star                                                # :61
    u_star[m0] = u_star[m0] * W[m0]                                               # :62
    u_star[m1] = u_star[m1] * W[m1]                                               # :63
    u_star = np.stack(u_star)
    F = physical_flux_xi(W, U, axis)                                              # :66
    S = np.minimum(S_K, 0.0) if side == "L" else np.maximum(S_K, 0.0)            # :69-72
    return F + S * (u_star - U)                                                   # :76


def hllc(W_L, W_R, U_L, U_R, axis, gamma, signal_speed="EINFELDT"):
    """HLLC._solve_riemann_problem_xi_single_phase (HLLC.py:80-126)."""
    v = axis + 1
    c_L = speed_of_sound(W_L[4], W_L[0], gamma)
    c_R = speed_of_sound(W_R[4], W_R[0], gamma)
    S_L, S_R = SIGNAL_SPEEDS[signal_speed](W_L[v], W_R[v], c_L, c_R, rho_L=W_L[0], rho_R=W_R[0],
                                            p_L=W_L[4], p_R=W_R[4], gamma=gamma)
    S_star = compute_sstar(W_L[v], W_R[v], W_L[4], W_R[4], W_L[0], W_R[0], S_L, S_R)
    Fs_L = _hllc_flux_star(W_L, U_L, S_L, S_star, axis, "L")
    Fs_R = _hllc_flux_star(W_R, U_R, S_R, S_star, axis, "R")
    sgn = np.sign(S_star)
    return 0.5 * (1 + sgn) * Fs_L + 0.5 * (1 - sgn) * Fs_R                        # :123-124


def rusanov(W_L, W_R, U_L, U_R, axis, gamma, signal_speed=None):
    """Rusanov._solve_riemann_problem_xi_single_phase (Rusanov.py:25-47)."""
    v = axis + 1
    c_L = speed_of_sound(W_L[4], W_L[0], gamma)
    c_R = speed_of_sound(W_R[4], W_R[0], gamma)
    alpha = np.maximum(np.abs(W_L[v]) + c_L, np.abs(W_R[v]) + c_R)
    F_L = physical_flux_xi(W_L, U_L, axis)
    F_R = physical_flux_xi(W_R, U_R, axis)
    return 0.5 * (F_L + F_R) - 0.5 * alpha * (U_R - U_L)


def hll(W_L, W_R, U_L, U_R, axis, gamma, signal_speed="EINFELDT", eps=EPS64):
    """HLL._solve_riemann_problem_xi_single_phase (HLL.py:29-61); eps = precision.get_eps()
    (riemann_solver.py:28)."""
    v = axis + 1
    c_L = speed_of_sound(W_L[4], W_L[0], gamma)
    c_R = speed_of_sound(W_R[4], W_R[0], gamma)
    S_L, S_R = SIGNAL_SPEEDS[signal_speed](W_L[v], W_R[v], c_L, c_R, rho_L=W_L[0], rho_R=W_R[0],
                                            p_L=W_L[4], p_R=W_R[4], gamma=gamma)
    wave_speed_L = np.minimum(S_L, 0.0)
    wave_speed_R = np.maximum(S_R, 0.0)
    F_L = physical_flux_xi(W_L, U_L, axis)
    F_R = physical_flux_xi(W_R, U_R, axis)
    return (wave_speed_R * F_L - wave_speed_L * F_R + wave_speed_L * wave_speed_R * (U_R - U_L)) \
        / (wave_speed_R - wave_speed_L + eps)


def hllc_lm(W_L, W_R, U_L, U_R, axis, gamma, signal_speed="EINFELDT"):
    """HLLCLM._solve_riemann_problem_xi_single_phase (HLLCLM.py:59-136; Fleischmann et al. 2020):
    HLLC star states, wave speeds scaled by phi = sin(min(1, Ma / 0.1) pi / 2) with
    Ma = max(|u_L / c_L|, |u_R / c_R|) (:37-57), the central star flux of Eq. (18) and the
    sign blend of the outer waves."""
    v = axis + 1
    m0, m1 = VELOCITY_MINOR_AXES[axis]
    c_L = speed_of_sound(W_L[4], W_L[0], gamma)
    c_R = speed_of_sound(W_R[4], W_R[0], gamma)
    S_L, S_R = SIGNAL_SPEEDS[signal_speed](W_L[v], W_R[v], c_L, c_R, rho_L=W_L[0], rho_R=W_R[0],
                                            p_L=W_L[4], p_R=W_R[4], gamma=gamma)
    S_star = compute_sstar(W_L[v], W_R[v], W_L[4], W_R[4], W_L[0], W_R[0], S_L, S_R)

    def star(W, U, S_K):  # :95-108
        pre = (S_K - W[v]) / (S_K - S_star) * W[0]
        u = [pre, pre, pre, pre, pre * (U[4] / U[0] + (S_star - W[v]) * (S_star + W[4] / W[0] / (S_K - W[v])))]
        u[v] = u[v] * S_star
        u[m0] = u[m0] * W[m0]
        u[m1] = u[m1] * W[m1]
        return np.stack(u)

    u_star_L, u_star_R = star(W_L, U_L, S_L), star(W_R, U_R, S_R)
    Ma_local = np.maximum(np.abs(W_L[v] / c_L), np.abs(W_R[v] / c_R))
    phi = np.sin(np.minimum(1.0, Ma_local / 0.1) * np.pi * 0.5)
    ws_L, ws_R = phi * S_L, phi * S_R
    F_L = physical_flux_xi(W_L, U_L, axis)
    F_R = physical_flux_xi(W_R, U_R, axis)
    flux_star = 0.5 * (F_L + F_R) + 0.5 * (ws_L * (u_star_L - U_L) + np.abs(S_star) * (u_star_L - u_star_R)
                                           + ws_R * (u_star_R - U_R))
    return (0.5 * (1 + np.sign(S_L)) * F_L + 0.5 * (1 - np.sign(S_R)) * F_R
            + 0.25 * (1 - np.sign(S_L)) * (1 + np.sign(S_R)) * flux_star)


def ausmp(W_L, W_R, U_L, U_R, axis, gamma, signal_speed=None):
    """AUSMP._solve_riemann_problem_xi_single_phase (AUSMP.py:30-83; Liou 1996) with the
    arithmetic interface speed of sound (:26,54-55), alpha = 3/16, beta = 1/8 (:27-28) and
    phi = (rho, rho u, rho v, rho w, E + p) (get_phi, :85-103); signal speeds are not used."""
    v = axis + 1
    alpha, beta = 3.0 / 16.0, 1.0 / 8.0
    phi_L = np.stack([U_L[0], U_L[1], U_L[2], U_L[3], U_L[4] + W_L[4]])
    phi_R = np.stack([U_R[0], U_R[1], U_R[2], U_R[3], U_R[4] + W_R[4]])
    c_L = speed_of_sound(W_L[4], W_L[0], gamma)
    c_R = speed_of_sound(W_R[4], W_R[0], gamma)
    a = 0.5 * (c_L + c_R)
    M_l = W_L[v] / a
    M_r = W_R[v] / a
    M_plus = np.where(np.abs(M_l) >= 1, 0.5 * (M_l + np.abs(M_l)),
                      0.25 * np.square(M_l + 1.0) + beta * np.square(M_l * M_l - 1.0))
    M_minus = np.where(np.abs(M_r) >= 1, 0.5 * (M_r - np.abs(M_r)),
                       -0.25 * np.square(M_r - 1.0) - beta * np.square(M_r * M_r - 1.0))
    M_ausm = M_plus + M_minus
    M_ausm_plus = 0.5 * (M_ausm + np.abs(M_ausm))
    M_ausm_minus = 0.5 * (M_ausm - np.abs(M_ausm))
    P_plus = np.where(np.abs(M_l) >= 1.0, 0.5 * (1 + np.sign(M_l)),
                      0.25 * np.square(M_l + 1.0) * (2.0 - M_l) + alpha * M_l * np.square(M_l * M_l - 1.0))
    P_minus = np.where(np.abs(M_r) >= 1.0, 0.5 * (1 - np.sign(M_r)),
                       0.25 * np.square(M_r - 1.0) * (2.0 + M_r) - alpha * M_r * np.square(M_r * M_r - 1.0))
    pressure_ausm = P_plus * W_L[4] + P_minus * W_R[4]
    F = a * (M_ausm_plus * phi_L + M_ausm_minus * phi_R)
    F[v] = F[v] + pressure_ausm
    return F


RIEMANN = {"HLLC": hllc, "RUSANOV": rusanov, "HLL": hll, "HLLC-LM": hllc_lm, "AUSMP": ausmp}


# --------------------------------------------------------------------------
# Setup + high-order Godunov flux
# --------------------------------------------------------------------------
@dataclass
class Transport:
    """Transport properties of the dissipative fluxes (materials/single_materials/material.py:46-130).

    Dynamic viscosity (get_dynamic_viscosity, material.py:82-101): with mu_sutherland None, the
    power law mu_ref (T / T_ref)^mu_exp that a CUSTOM value or lambda expresses (mu_exp = 0: the
    constant mu_ref); with mu_sutherland = C, SUTHERLAND mu_ref ((T_ref + C)/(T + C)) (T/T_ref)^1.5
    (:94-96, power3_2 = x sqrt(x), math/power_functions.py:29-30).
    Thermal conductivity (get_thermal_conductivity, :103-124): with prandtl set, PRANDTL
    kappa = cp mu / Pr (:115-116), cp = gamma R / (gamma - 1) (ideal_gas.py:32); else with kappa_ref
    set, the same two laws from (kappa_ref, kappa_T_ref, kappa_exp | kappa_sutherland); neither:
    no heat flux.  T = p / (rho R) (ideal_gas.py:81-83).
    """

    R: float = 1.0
    mu_ref: float = 0.0
    T_ref: float = 1.0
    mu_exp: float = 0.0
    bulk_viscosity: float = 0.0
    prandtl: Optional[float] = None
    viscous: bool = True                  # active_physics.is_viscous_flux
    mu_sutherland: Optional[float] = None
    kappa_ref: Optional[float] = None
    kappa_T_ref: float = 1.0
    kappa_exp: float = 0.0
    kappa_sutherland: Optional[float] = None

    @property
    def heat(self) -> bool:               # active_physics.is_heat_flux
        return self.prandtl is not None or self.kappa_ref is not None

    @staticmethod
    def _law(T, ref, T_ref, exp, C):
        if C is not None:
            x = T / T_ref
            return ref * ((T_ref + C) / (T + C)) * (x * np.sqrt(x))
        return ref * (T / T_ref) ** exp if exp != 0.0 else ref * np.ones_like(T)

    def mu(self, T):
        return self._law(T, self.mu_ref, self.T_ref, self.mu_exp, self.mu_sutherland)

    def cp(self, gamma):
        return gamma / (gamma - 1.0) * self.R

    def kappa(self, T, gamma):
        if self.prandtl is not None:
            return self.cp(gamma) * self.mu(T) / self.prandtl
        return self._law(T, self.kappa_ref, self.kappa_T_ref, self.kappa_exp, self.kappa_sutherland)


@dataclass
class Setup:
    """Subset of NumericalSetup / CaseSetup on the hot path."""

    gamma: float = 1.4
    stencil: str = "WENO5-JS"                 # reconstruction_stencil
    variable: str = "PRIMITIVE"               # reconstruction_variable
    riemann: str = "HLLC"                     # riemann_solver
    signal_speed: str = "EINFELDT"            # signal_speed
    frozen_state: str = "ARITHMETIC"          # frozen_state
    convective_solver: str = "GODUNOV"        # convective_solver: "GODUNOV" / "FLUX-SPLITTING"
    flux_splitting: str = "ROE"               # flux_splitting/flux_splitting: "ROE" / "CLLF" / "LLF"
    cfl: float = 0.5
    fixed_dt: Optional[float] = None
    bcs: Dict[str, str] = field(default_factory=lambda: {f: "PERIODIC" for f in FACE_LOCATIONS})
    eps_stencil: float = SPATIAL_STENCIL_EPS
    transport: Optional[Transport] = None     # viscous (and heat) fluxes when set
    interpolation_limiter: bool = False       # positivity.is_interpolation_limiter
    limit_velocity: bool = False              # positivity.limit_velocity
    limiter_eps: Tuple[float, float] = (1e-12, 1e-10)  # (density, pressure), config/precision.py:54
    flux_limiter: Optional[str] = None        # positivity.flux_limiter: "SIMPLE" / "NASA" / "HAS"
    flux_partition: str = "UNIFORM"           # positivity.flux_partition: "UNIFORM" / "CELLSIZE" / "WAVESPEED"
    flux_limiter_eps: Tuple[float, float] = (1e-12, 1e-10)  # config/precision.py:55

    def __post_init__(self):
        # riemann "FLUX-SPLITTING-<eigenvalues>" (the native library's face-flux name) selects the
        # flux-splitting convective solver
        if self.riemann.startswith("FLUX-SPLITTING-"):
            self.convective_solver = "FLUX-SPLITTING"
            self.flux_splitting = self.riemann[len("FLUX-SPLITTING-"):]


def reconstruct_xi(W, U, dom: Domain, axis: int, s: Setup):
    """HighOrderGodunov.reconstruct_xi (high_order_godunov.py:233-419).

    Returns U_L, U_R, W_L, W_R on the faces along axis.
    """
    g = s.gamma
    if s.variable == "PRIMITIVE":                                             # :267-280
        W_L = weno5_reconstruct_xi(W, dom, axis, 0, s.stencil, s.eps_stencil)
        W_R = weno5_reconstruct_xi(W, dom, axis, 1, s.stencil, s.eps_stencil)
        U_L, U_R = cons_from_prim(W_L, g), cons_from_prim(W_R, g)
    elif s.variable == "CONSERVATIVE":                                        # :282-296
        U_L = weno5_reconstruct_xi(U, dom, axis, 0, s.stencil, s.eps_stencil)
        U_R = weno5_reconstruct_xi(U, dom, axis, 1, s.stencil, s.eps_stencil)
        W_L, W_R = prim_from_cons(U_L, g), prim_from_cons(U_R, g)
    elif s.variable == "CHAR-PRIMITIVE":                                      # :298-316,401-402
        win = stencil_window(W, dom, axis, window_size(s.stencil))
        cl, cr = face_LR_cells(W, dom, axis)
        P, _, _, c, cc, _ = frozen_state(cl, cr, g, s.frozen_state)
        # expand_dims(axis=axis-3) of the frozen state == broadcast over the window axis
        char = char_from_prim(win, P[:, None], c[None], cc[None], axis)
        cL = weno5_reconstruct_window(char, 0, s.stencil, s.eps_stencil, dom.dx[axis])
        cR = weno5_reconstruct_window(char, 1, s.stencil, s.eps_stencil, dom.dx[axis])
        W_L = prim_from_char(cL, P, c, cc, axis)
        W_R = prim_from_char(cR, P, c, cc, axis)
        U_L, U_R = cons_from_prim(W_L, g), cons_from_prim(W_R, g)
    elif s.variable == "CHAR-CONSERVATIVE":                                   # :404-417
        win = stencil_window(U, dom, axis, window_size(s.stencil))
        cl, cr = face_LR_cells(W, dom, axis)
        fs = frozen_state(cl, cr, g, s.frozen_state)
        R, L = eigen_conservatives(*fs, axis=axis)
        char = transform(L[:, :, None], win)
        cL = weno5_reconstruct_window(char, 0, s.stencil, s.eps_stencil, dom.dx[axis])
        cR = weno5_reconstruct_window(char, 1, s.stencil, s.eps_stencil, dom.dx[axis])
        U_L, U_R = transform(R, cL), transform(R, cR)
        W_L, W_R = prim_from_cons(U_L, g), prim_from_cons(U_R, g)
    else:
        raise NotImplementedError(s.variable)
    return U_L, U_R, W_L, W_R


def limit_interpolation_xi(W, W_j, j: int, dom: Domain, axis: int, s: Setup):
    """PositivityLimiterInterpolation.limit_interpolation_xi, single phase
    (solvers/positivity/limiter_interpolation.py:86-208): face states with density below
    eps.density take the first-order value (WENO1, the cell upwind of the face on side j:
    weno1_js.py:22-29) in the limited fields (density and pressure; all five with
    limit_velocity, :61-69); then, on the updated states, those with p + pb < eps.pressure
    (pb = 0 for an ideal gas) likewise.  Returns (U_j, W_j, count) with U_j = cons(W_j) on every
    face (:206) and count the number of replacements."""
    first = face_LR_cells(W, dom, axis)[j]
    ids = [0, 1, 2, 3, 4] if s.limit_velocity else [0, 4]
    W_j = W_j.copy()
    count = 0
    for field_id, eps in ((0, s.limiter_eps[0]), (4, s.limiter_eps[1])):
        mask = W_j[field_id] < eps
        count += int(np.sum(mask))
        for i in ids:
            W_j[i] = W_j[i] * (1 - mask) + first[i] * mask
    return cons_from_prim(W_j, s.gamma), W_j, count


def flux_splitting_eigenvalues(W, dom: Domain, axis: int, P, c_ave, gamma, flux_splitting: str):
    """Eigenvalue magnitudes of the flux splitting (eigendecomposition.py:672-696): the diagonal
    (gamma_1, gamma_234 x3, gamma_5) on every face."""
    v = axis + 1
    if flux_splitting == "ROE":                                              # :674-677
        g1, g234, g5 = np.abs(P[v] - c_ave), np.abs(P[v]), np.abs(P[v] + c_ave)
    elif flux_splitting in ("CLLF", "LLF"):
        W_L, W_R = face_LR_cells(W, dom, axis)
        c_L = speed_of_sound(W_L[4], W_L[0], gamma)
        c_R = speed_of_sound(W_R[4], W_R[0], gamma)
        if flux_splitting == "CLLF":                                         # :680-688
            g1 = np.maximum(np.abs(W_L[v] - c_L), np.abs(W_R[v] - c_R))
            g234 = np.maximum(np.abs(W_L[v]), np.abs(W_R[v]))
            g5 = np.maximum(np.abs(W_L[v] + c_L), np.abs(W_R[v] + c_R))
        else:                                                                # :690-695
            g1 = g234 = g5 = np.maximum(np.abs(W_L[v]) + c_L, np.abs(W_R[v]) + c_R)
    else:
        raise NotImplementedError(flux_splitting)                            # :698-699
    return np.stack([g1, g234, g234, g234, g5], axis=0)


def flux_splitting_flux_xi(W, U, dom: Domain, axis: int, s: Setup):
    """FluxSplittingScheme.compute_flux_xi (flux_splitting_scheme.py:65-111), single phase.

    The window of conservatives and of physical fluxes is projected on the left eigenvectors of
    the frozen state of each face's two cells; with the diagonal eigenvalue matrix Lambda,
    f+- = (L F +- Lambda L U) / 2 (:91-92); f+ is reconstructed with j = 0, f- with j = 1
    (:94-105), and the sum is projected back with the right eigenvectors (:107-109)."""
    cl, cr = face_LR_cells(W, dom, axis)
    fs = frozen_state(cl, cr, s.gamma, s.frozen_state)
    R, L = eigen_conservatives(*fs, axis=axis)
    lam = flux_splitting_eigenvalues(W, dom, axis, fs[0], fs[3], s.gamma, s.flux_splitting)
    char = transform(L[:, :, None], stencil_window(U, dom, axis))
    F = physical_flux_xi(W, U, axis)                                         # equation_manager.get_fluxes_xi
    char_flux = transform(L[:, :, None], stencil_window(F, dom, axis))
    # transformtochar(char, eigvals): the diagonal matrix times each window cell
    lam_char = lam[:, None] * char
    positive = 0.5 * (char_flux + lam_char)
    negative = 0.5 * (char_flux - lam_char)
    f_L = weno5_reconstruct_window(positive, 0, s.stencil, s.eps_stencil, dom.dx[axis])
    f_R = weno5_reconstruct_window(negative, 1, s.stencil, s.eps_stencil, dom.dx[axis])
    return transform(R, f_L + f_R)


def compute_flux_xi(W, U, dom: Domain, axis: int, s: Setup, return_count: bool = False):
    """HighOrderGodunov.compute_flux_xi (high_order_godunov.py:117-231), single phase:
    reconstruct, apply the interpolation limiter when enabled (:163-176), then solve the
    Riemann problem.  With convective_solver FLUX-SPLITTING, flux_splitting_flux_xi (no
    interpolation limiter there)."""
    if s.convective_solver == "FLUX-SPLITTING":
        F = flux_splitting_flux_xi(W, U, dom, axis, s)
        return (F, None) if return_count else F
    U_L, U_R, W_L, W_R = reconstruct_xi(W, U, dom, axis, s)
    count = None
    if s.interpolation_limiter:
        U_L, W_L, cL = limit_interpolation_xi(W, W_L, 0, dom, axis, s)
        U_R, W_R, cR = limit_interpolation_xi(W, W_R, 1, dom, axis, s)
        count = cL + cR
    F = RIEMANN[s.riemann](W_L, W_R, U_L, U_R, axis, s.gamma, s.signal_speed)
    return (F, count) if return_count else F


def flux_divergence(F, one_dx, axis):
    """space_solver.py:489,597-599: zeros + (1/dx) * (F[:-1] - F[1:]) along axis."""
    lo = [slice(None)] * 3
    hi = [slice(None)] * 3
    lo[axis] = slice(None, -1)
    hi[axis] = slice(1, None)
    lo = (Ellipsis,) + tuple(lo)
    hi = (Ellipsis,) + tuple(hi)
    return 0.0 + one_dx * (F[lo] - F[hi])


# --------------------------------------------------------------------------
# Dissipative fluxes, second-order central (solvers/source_term_solver.py:188-350,414-448,
# 504-580; stencils central_2.py, deriv_face_2.py, deriv_center_2.py).  Periodic domains: the
# primitive buffer is rebuilt from the interior with wrap padding, which fills face, edge and
# vertex halos exactly as the reference's periodic halo update does for viscous runs
# (halo_manager.py:125-129).
# --------------------------------------------------------------------------
def wrap_halos(W, dom: Domain):
    """Buffer whose halos (faces, edges, vertices) are periodic copies of the interior."""
    Wi = W[dom.interior]
    pad = [(0, 0)] + [(dom.nh, dom.nh) if i in dom.active else (0, 0) for i in range(3)]
    return np.pad(Wi, pad, mode="wrap")


def _face_cells(buf, dom: Domain, axis: int, k: int, other: Optional[int] = None, ok: int = 0):
    """Cells f + k (k = -1: left of face f, 0: right) of every face f along axis, shifted by ok
    along `other`; face-array shape (shift_slice with an extra transverse offset)."""
    n = dom.nh
    sl = [dom.dom(0), dom.dom(1), dom.dom(2)]
    hi = -n + k + 1
    sl[axis] = slice(n + k, None if hi == 0 else hi)
    if other is not None and ok != 0:
        sl[other] = slice(n + ok, -n + ok)
    return buf[(Ellipsis,) + tuple(sl)]


def velocity_gradient_at_faces(W, dom: Domain, axis: int):
    """compute_velocity_gradient (source_term_solver.py:414-448) with
    compute_xi_derivatives_at_xj (:538-580): g[k][i] = d u_k / d x_i at the faces along axis;
    i == axis: (u_R - u_L)/dx (deriv_face_2); i != axis: the central derivative at the two
    adjacent cells (deriv_center_2), averaged to the face (central_2); inactive i: 0."""
    L = _face_cells(W, dom, axis, -1)
    g = [[None] * 3 for _ in range(3)]
    for i in range(3):
        for k in range(3):
            if i not in dom.active:
                g[k][i] = np.zeros_like(L[0])
            elif i == axis:
                g[k][i] = (_face_cells(W, dom, axis, 0)[1 + k] - L[1 + k]) * dom.one_dx[i]
            else:
                h = 0.5 * dom.one_dx[i]
                dL = (_face_cells(W, dom, axis, -1, i, 1)[1 + k] - _face_cells(W, dom, axis, -1, i, -1)[1 + k]) * h
                dR = (_face_cells(W, dom, axis, 0, i, 1)[1 + k] - _face_cells(W, dom, axis, 0, i, -1)[1 + k]) * h
                g[k][i] = 0.5 * (dL + dR)
    return g


def dissipative_flux_xi(W, dom: Domain, axis: int, s: Setup):
    """Viscous flux (compute_viscous_flux_xi :264-350, compute_tau :504-536) and heat flux
    (compute_heat_flux_xi :188-262) on the faces along axis, as the amounts added to the
    convective face flux (space_solver.py:565-587): momentum -tau_axis,k, energy
    -sum_k tau_axis,k u_k + q with q = -kappa dT/dx_axis."""
    tr = s.transport
    # the cells next to the faces and their transverse neighbours, face and edge halos included
    # (filled by halo_update with the edge fill); on periodic domains these are the wrapped
    # interior cells
    periodic = all(s.bcs[f] == "PERIODIC" for f in FACE_LOCATIONS if FACE_AXIS[f] in dom.active)
    Wp = wrap_halos(W, dom) if periodic else W
    L, R = _face_cells(Wp, dom, axis, -1), _face_cells(Wp, dom, axis, 0)
    T_L, T_R = L[4] / (L[0] * tr.R), R[4] / (R[0] * tr.R)
    T_f = 0.5 * (T_L + T_R)                                    # central_2.py reconstruct_xi
    T_f = np.where(T_f <= 0.0, EPS64, T_f)                     # source_term_solver.py:301
    mu1 = tr.mu(T_f)
    mu2 = tr.bulk_viscosity - 2.0 / 3.0 * mu1                  # :527
    g = velocity_gradient_at_faces(Wp, dom, axis)
    tau = [np.zeros_like(T_f) for _ in range(3)]
    if tr.viscous:
        for k in range(3):
            if axis in dom.active and k in dom.active:
                tau[k] = mu1 * (g[axis][k] + g[k][axis])
        tau[axis] = tau[axis] + mu2 * sum(g[k][k] for k in dom.active)
    u_f = [0.5 * (L[1 + k] + R[1 + k]) for k in range(3)]     # central_2 of the velocity
    vel_tau = 0.0
    for k in dom.active:
        vel_tau = vel_tau + tau[k] * u_f[k]
    D = np.zeros((5,) + T_f.shape)
    for k in range(3):
        D[1 + k] = -tau[k]
    D[4] = -vel_tau
    if tr.heat:
        q = -tr.kappa(T_f, s.gamma) * ((T_R - T_L) * dom.one_dx[axis])   # deriv_face_2
        D[4] = D[4] + q
    return D


def flux_partition_sigma(W, dom: Domain, axis: int, s: Setup):
    """compute_partition (solvers/positivity/limiter_flux.py:696-727): 1/sigma of the axis, i.e.
    dim (UNIFORM), sum_active(1/dx) / (1/dx_axis) (CELLSIZE), or, with m_a the maximum of
    |u_a| + c over the whole primitive buffer, halos included (WAVESPEED, :717-726),
    sum_active(m_a / dx_a) / (m_axis / dx_axis), the sum started at 0 as Python's sum."""
    one_dx = dom.one_dx[axis]
    if s.flux_partition == "UNIFORM":                                     # :711-712
        return float(len(dom.active))
    if s.flux_partition == "CELLSIZE":                                    # :714-715
        return sum(dom.one_dx[a] for a in dom.active) / one_dx
    if s.flux_partition == "WAVESPEED":                                   # :717-726
        with np.errstate(all="ignore"):
            c = speed_of_sound(W[4], W[0], s.gamma)
            m = [np.max(np.abs(W[1 + a]) + c) for a in range(3)]
        return sum(m[a] * dom.one_dx[a] for a in dom.active) / (m[axis] * one_dx)
    raise NotImplementedError(s.flux_partition)


def positivity_flux_xi(F, W, U, dom: Domain, axis: int, s: Setup, dt: float, return_count: bool = False):
    """PositivityLimiterFlux.compute_positivity_preserving_flux, SIMPLE / NASA / HAS, single
    phase (solvers/positivity/limiter_flux.py:147-316,452-562,563-640,696-774).

    lambda = dt (1/dx_axis) sigma, sigma from flux_partition_sigma.  Face f lies between cells f-1
    and f; the pseudo-updated states are U_minus = U_f + 2 lambda (F - Fs_f) and
    U_plus = U_{f-1} - 2 lambda (F - Fs_{f-1}), Fs the cells' own physical flux (NASA) or 0
    (SIMPLE, HAS).  The first-order flux F_pos is HLLC with Einfeldt speeds on the two cells'
    primitives (a HighOrderGodunov with WENO1, :57-98).
    SIMPLE / NASA: where min(rho_minus, rho_plus) < eps_rho the flux becomes F_pos, then, with the
    updated flux, likewise where min(p_minus, p_plus) < eps_p; count: faces switched, summed over
    the two checks.
    HAS (Hu, Adams & Shu, :200-246): theta = (eps - q_LF) / (q - q_LF) on each side whose density q
    is below eps_rho (1 elsewhere), q_LF the same side's density with F_pos; the flux becomes
    (1 - theta) F_pos + theta F with theta = min over the sides; then likewise with the pressures
    of the updated flux's states and of the F_pos states; count: sum of 1 - min(theta_rho, theta_p)."""
    lam = dt * dom.one_dx[axis] * flux_partition_sigma(W, dom, axis, s)   # :204
    W_L, W_R = face_LR_cells(W, dom, axis)
    U_L, U_R = face_LR_cells(U, dom, axis)
    F_pos = hllc(W_L, W_R, cons_from_prim(W_L, s.gamma), cons_from_prim(W_R, s.gamma), axis, s.gamma,
                 "EINFELDT")
    if s.flux_limiter == "HAS":
        eps_rho, eps_p = s.flux_limiter_eps
        with np.errstate(all="ignore"):
            U_minus = U_R + 2 * lam * F                                   # :220-224
            U_minus_LF = U_R + 2 * lam * F_pos
            theta_minus = np.where(U_minus[0] < eps_rho, (eps_rho - U_minus_LF[0]) / (U_minus[0] - U_minus_LF[0]),
                                   1.0)
            U_plus = U_L - 2 * lam * F                                    # :226-230
            U_plus_LF = U_L - 2 * lam * F_pos
            theta_plus = np.where(U_plus[0] < eps_rho, (eps_rho - U_plus_LF[0]) / (U_plus[0] - U_plus_LF[0]), 1.0)
            theta_rho = np.minimum(theta_minus, theta_plus)
            F = (1 - theta_rho) * F_pos + theta_rho * F                   # :234
            U_minus = U_R + 2 * lam * F                                   # :237-241
            W_minus = prim_from_cons(U_minus, s.gamma)
            W_minus_LF = prim_from_cons(U_minus_LF, s.gamma)
            theta_minus = np.where(W_minus[4] < eps_p, (eps_p - W_minus_LF[4]) / (W_minus[4] - W_minus_LF[4]), 1.0)
            U_plus = U_L - 2 * lam * F                                    # :243-246
            W_plus = prim_from_cons(U_plus, s.gamma)
            W_plus_LF = prim_from_cons(U_plus_LF, s.gamma)
            theta_plus = np.where(W_plus[4] < eps_p, (eps_p - W_plus_LF[4]) / (W_plus[4] - W_plus_LF[4]), 1.0)
            theta_p = np.minimum(theta_minus, theta_plus)
            count = float(np.sum(1 - np.minimum(theta_rho, theta_p)))     # :249
            F = (1 - theta_p) * F_pos + theta_p * F                       # :251
        return (F, count) if return_count else F
    if s.flux_limiter == "NASA":                                          # :528-541
        Fs_L, Fs_R = physical_flux_xi(W_L, U_L, axis), physical_flux_xi(W_R, U_R, axis)
    elif s.flux_limiter == "SIMPLE":                                      # :520-527
        Fs_L = Fs_R = 0.0
    else:
        raise NotImplementedError(s.flux_limiter)

    def integrate(F):
        return U_R + 2.0 * lam * (F - Fs_R), U_L - 2.0 * lam * (F - Fs_L)

    count = 0
    with np.errstate(all="ignore"):
        U_minus, U_plus = integrate(F)
        theta = np.minimum(U_minus[0], U_plus[0]) < s.flux_limiter_eps[0]          # :570-575
        count += int(np.sum(theta))
        F = np.where(theta, F_pos, F)                                              # :764
        U_minus, U_plus = integrate(F)
        p_minus = prim_from_cons(U_minus, s.gamma)[4]
        p_plus = prim_from_cons(U_plus, s.gamma)[4]
        theta = np.minimum(p_minus, p_plus) < s.flux_limiter_eps[1]                # :619-627
        count += int(np.sum(theta))
        F = np.where(theta, F_pos, F)
    return (F, count) if return_count else F


def compute_rhs(W, U, dom: Domain, s: Setup, dt: Optional[float] = None):
    """SpaceSolver.compute_rhs (space_solver.py:266,289-314): 0.0 + rhs_x + rhs_y + rhs_z;
    with the positivity flux limiter, each axis' convective flux is limited with the step's
    time step dt (:530-541); with transport set, each axis' face flux carries the dissipative
    part (:565-587)."""
    rhs = 0.0
    for axis in dom.active:
        F = compute_flux_xi(W, U, dom, axis, s)
        if s.flux_limiter:
            F = positivity_flux_xi(F, W, U, dom, axis, s, dt)
        if s.transport is not None:
            F = F + dissipative_flux_xi(W, dom, axis, s)
        rhs = rhs + flux_divergence(F, dom.one_dx[axis], axis)
    return rhs


# --------------------------------------------------------------------------
# Halos (halos/outer/material.py, boundary_condition.py, halo_slices.py)
# --------------------------------------------------------------------------
def _face_fill_slice(dom: Domain, loc: str):
    """halo_slices.py:66-73."""
    nh = dom.nh
    a = FACE_AXIS[loc]
    s = [dom.dom(0), dom.dom(1), dom.dom(2)]
    s[a] = slice(-nh, None) if loc in ("east", "north", "top") else slice(None, nh)
    return (Ellipsis,) + tuple(s)


def _face_retrieve_slice(dom: Domain, loc: str, bc: str):
    """boundary_condition.py:564-595."""
    nh = dom.nh
    a = FACE_AXIS[loc]
    hi = loc in ("east", "north", "top")
    s = [dom.dom(0), dom.dom(1), dom.dom(2)]
    if bc == "PERIODIC":
        s[a] = slice(nh, 2 * nh) if hi else slice(-2 * nh, -nh)
    elif bc == "SYMMETRY":
        s[a] = slice(-nh - 1, -2 * nh - 1, -1) if hi else slice(2 * nh - 1, nh - 1, -1)
    elif bc == "ZEROGRADIENT":
        s[a] = slice(-nh - 1, -nh) if hi else slice(nh, nh + 1)
    else:
        raise NotImplementedError(bc)
    return (Ellipsis,) + tuple(s)


# domain/__init__.py:9-15; the first face of each name lies on the lower axis
EDGE_LOCATIONS = ("west_south", "west_north", "east_north", "east_south", "south_bottom", "north_bottom",
                  "south_top", "north_top", "east_bottom", "west_bottom", "east_top", "west_top")
_OPPOSITE = {"east": "west", "west": "east", "north": "south", "south": "north", "top": "bottom", "bottom": "top"}


def edge_type(s: Setup, edge: str) -> str:
    """assign_boundary_type_combinations (outer/boundary_condition.py:128-180): the first face of
    the edge (in name order) that is PERIODIC or SYMMETRY gives its type, the other face "ANY";
    with neither, "ANY_ANY"."""
    out, found = [], False
    for face in edge.split("_"):
        b = s.bcs[face]
        if not found and b in ("PERIODIC", "SYMMETRY"):
            out.append(b)
            found = True
        else:
            out.append("ANY")
    return "_".join(out)


def _edge_slice(dom: Domain, edge: str, shift: Optional[Tuple[int, int]] = None):
    """edge_slices of halos/halo_slices.py:77-150: the edge halo block (nh x nh cells at the two
    faces, interior along the third axis); shift (da, db) in {0, 1} moves it by nh cells into the
    interior along the first (da) or second (db) face's axis (the "_10" / "_01" blocks), and
    PERIODIC retrieval uses the opposite face's block."""
    nh = dom.nh
    sl = [dom.dom(0), dom.dom(1), dom.dom(2)]
    for k, face in enumerate(edge.split("_")):
        a = FACE_AXIS[face]
        hi = face in ("east", "north", "top")
        moved = shift is not None and shift[k]
        if not moved:
            sl[a] = slice(-nh, None) if hi else slice(None, nh)
        else:
            sl[a] = slice(-2 * nh, -nh) if hi else slice(nh, 2 * nh)
    return (Ellipsis,) + tuple(sl)


def edge_halo_update(W, U, dom: Domain, s: Setup):
    """BoundaryConditionMaterial.edge_halo_update / compute_edge_halos (outer/material.py:289-380)
    on the active edges (both faces' axes active):
    ANY_ANY: the mean of the face-halo blocks next to the edge, moved nh cells into the interior
    along either face's axis ("_10", "_01"), unmirrored (halo_slices.py:77-150);
    PERIODIC on a face: the block of the opposite face moved along that face's axis ("_10" or
    "_01" of the edge with that face replaced by its opposite, boundary_condition.py:607-654);
    SYMMETRY on a face: the block moved along that face's axis, reversed along it, with that
    axis' velocity negated (:657-760).  U = cons(W) on the edge cells."""
    W = W.copy()
    U = None if U is None else U.copy()
    for edge in EDGE_LOCATIONS:
        faces = edge.split("_")
        if not all(FACE_AXIS[f] in dom.active for f in faces):
            continue
        t = edge_type(s, edge).split("_")
        if t == ["ANY", "ANY"]:
            halos = 0.5 * (W[_edge_slice(dom, edge, (1, 0))] + W[_edge_slice(dom, edge, (0, 1))])
        else:
            k = t.index("PERIODIC") if "PERIODIC" in t else t.index("SYMMETRY")
            shift = (1, 0) if k == 0 else (0, 1)
            if t[k] == "PERIODIC":
                src = list(faces)
                src[k] = _OPPOSITE[faces[k]]
                halos = W[_edge_slice(dom, "_".join(src), shift)]
            else:
                halos = W[_edge_slice(dom, edge, shift)]
                a = FACE_AXIS[faces[k]]
                flip = [slice(None)] * 4
                flip[1 + a] = slice(None, None, -1)
                halos = halos[tuple(flip)].copy()
                halos[1 + a] = -halos[1 + a]
        W[_edge_slice(dom, edge)] = halos
        if U is not None:
            U[_edge_slice(dom, edge)] = cons_from_prim(halos, s.gamma)
    return W, U


def halo_update(W, U, dom: Domain, s: Setup, fill_edge: Optional[bool] = None):
    """BoundaryConditionMaterial.face_halo_update for PERIODIC / ZEROGRADIENT /
    SYMMETRY (outer/material.py:94-287, miscellaneous :868-894).  U halo is
    recomputed from the W halo (:248-250).  Edge halos follow with viscous or heat fluxes
    (fill_edge_halos_material, halo_manager.py:125-129,183-186; 2-D and 3-D), vertex halos never
    on this path."""
    if fill_edge is None:
        fill_edge = s.transport is not None
    W, U = _face_halo_update(W, U, dom, s)
    if fill_edge and len(dom.active) > 1:
        W, U = edge_halo_update(W, U, dom, s)
    return W, U


def _face_halo_update(W, U, dom: Domain, s: Setup):
    W = W.copy()
    U = None if U is None else U.copy()
    for loc in FACE_LOCATIONS:
        if FACE_AXIS[loc] not in dom.active:
            continue
        bc = s.bcs[loc]
        halos = W[_face_retrieve_slice(dom, loc, bc)]
        if bc == "SYMMETRY":
            sign = np.ones((5, 1, 1, 1))
            sign[FACE_AXIS[loc] + 1] = -1.0                  # boundary_condition.py:721-731
            halos = halos * sign
        fill = _face_fill_slice(dom, loc)
        # mask = 1: buffer*(1-1) + halos*1 == halos for finite values (:278-282)
        W[fill] = halos
        if U is not None:
            U[fill] = cons_from_prim(halos, s.gamma)
    return W, U


# --------------------------------------------------------------------------
# Time integration (time_integration/RK3.py, time_integrator.py, time_step_size.py)
# --------------------------------------------------------------------------
RK3_TIMESTEP_MULTIPLIER = (1.0, 0.25, 2.0 / 3.0)            # RK3.py:27
RK3_INCREMENT_FACTOR = (1.0, 0.5, 1.0)                      # RK3.py:28
RK3_CONS_MULTIPLIER = ((0.25, 0.75), (2.0 / 3.0, 1.0 / 3.0))  # RK3.py:29


def rk3_stage_update(U, U_n, rhs, dt, stage, dom: Domain):
    """perform_stage_integration (time_integrator.py:108-227) for RK3:
    stage>0: U = c0*U + c1*U_n on the whole buffer (RK3.py:49-50);
    then U[interior] += (dt*m_stage) * rhs (RK3.py:60-61, time_integrator.py:57)."""
    if stage > 0:
        c0, c1 = RK3_CONS_MULTIPLIER[stage - 1]
        U = c0 * U + c1 * U_n
    else:
        U = U.copy()
    timestep = dt * RK3_TIMESTEP_MULTIPLIER[stage]
    U[dom.interior] = U[dom.interior] + timestep * rhs
    return U


def compute_dt(W, dom: Domain, s: Setup, eps=EPS64):
    """compute_time_step_size, convective part (time_step_size.py:40-43,103-109,154-155)."""
    if s.fixed_dt:
        return s.fixed_dt
    Wi = W[dom.interior]
    c = speed_of_sound(Wi[4], Wi[0], s.gamma)
    abs_velocity = 0.0
    for i in dom.active:
        abs_velocity = abs_velocity + (np.abs(Wi[i + 1]) + c)
    dt = dom.dx_min / (np.max(abs_velocity) + eps)
    tr = s.transport
    if tr is not None:  # time_step_size.py:112-135
        T = Wi[4] / (Wi[0] * tr.R)
        if tr.viscous:
            dt = min(dt, 3.0 / 14.0 * dom.dx_min ** 2 / (np.max(tr.mu(T) / Wi[0]) + eps))
        if tr.heat:
            dt = min(dt, 0.1 * dom.dx_min ** 2 / (np.max(tr.kappa(T, s.gamma) / (Wi[0] * tr.cp(s.gamma))) + eps))
    return dt * s.cfl


def rk3_step(W, U, dt, dom: Domain, s: Setup, rhs_fn: Callable = None):
    """One TVD-RK3 step in the order of SimulationManager.do_runge_kutta_stages
    (simulation_manager.py:770-1047): rhs -> stage integration -> W = prim(U) on the
    whole buffer -> face halo update with U halo = cons(W halo)."""
    rhs_fn = rhs_fn or (lambda W_, U_: compute_rhs(W_, U_, dom, s, dt))
    U_n = U
    for stage in range(3):
        rhs = rhs_fn(W, U)
        U = rk3_stage_update(U, U_n, rhs, dt, stage, dom)
        with np.errstate(all="ignore"):
            W = prim_from_cons(U, s.gamma)
        W, U = halo_update(W, U, dom, s)
    return W, U


def create_buffers(prim_interior, dom: Domain, s: Setup, eps=EPS64):
    """from_primitive_initial_condition (material_fields_initializer.py:590-690):
    eps-filled buffer (helper_functions.py:49-60), interior set, U = cons(W) on the
    whole buffer, face halo update at t=0."""
    W = np.ones((5,) + dom.shape) * eps
    W[dom.interior] = prim_interior
    U = cons_from_prim(W, s.gamma)
    return halo_update(W, U, dom, s)


def simulate(W, U, dom: Domain, s: Setup, end_time=None, end_step=None):
    """SimulationManager.advance loop (simulation_manager.py:325-328): first dt from the
    initial state (time_control_initializer.py:55-66); step with the current dt; new dt
    from the new state (:612-628); stop at the first t >= end_time (no clipping)."""
    t, step = 0.0, 0
    end_time = np.inf if end_time is None else end_time
    end_step = np.iinfo(np.int64).max if end_step is None else end_step
    dt = compute_dt(W, dom, s)
    while t < end_time and step < end_step:
        W, U = rk3_step(W, U, dt, dom, s)
        t += dt
        step += 1
        dt = compute_dt(W, dom, s)
    return W, U, t, step


# --------------------------------------------------------------------------
# Initial conditions of the workloads
# --------------------------------------------------------------------------
def sod_initial(n=1000):
    """1-D Sod on [0,1]: (rho,u,p) = (1,0,1) for x<=0.5 else (0.125,0,0.1)
    (examples/examples_1D/02_sod_shock_tube/sod.json)."""
    dom = Domain((n, 1, 1), ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), nh=5)
    x = dom.cell_centers()[0].reshape(-1, 1, 1)
    rho = 1.0 * (x <= 0.5) + 0.125 * (x > 0.5)
    p = 1.0 * (x <= 0.5) + 0.1 * (x > 0.5)
    z = np.zeros_like(x)
    return dom, np.stack([rho, z, z, z, p])


def tgv_initial(n=256, nh=5, gamma=1.4, mach=0.1):
    """Inviscid Taylor-Green vortex on [0,2pi)^3: rho=1, u=sin x cos y cos z,
    v=-cos x sin y cos z, w=0, p = p0 + (1/16)(cos2x+cos2y)(cos2z+2), p0 = 1/(gamma Ma^2)
    (examples/examples_3D/01_tgv/tgv.json)."""
    L = 2.0 * np.pi
    dom = Domain((n, n, n), ((0.0, L), (0.0, L), (0.0, L)), nh=nh)
    xs = dom.cell_centers()
    X, Y, Z = np.meshgrid(xs[0], xs[1], xs[2], indexing="ij")
    rho = np.ones_like(X)
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(X)
    p = 1.0 / gamma / mach ** 2 + 1.0 / 16.0 * ((np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2))
    return dom, np.stack([rho, u, v, w, p])
