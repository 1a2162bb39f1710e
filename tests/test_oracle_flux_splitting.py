"""Pins of the oracle's flux-splitting scheme (flux_splitting_flux_xi, FluxSplittingScheme of
solvers/convective_fluxes/flux_splitting_scheme.py:65-111) by what the mathematics fixes:

- consistency: on a uniform state every splitting returns the physical flux (R L = I and the two
  halves of the split add up);
- with WENO1 (first order) the LLF splitting is the Rusanov flux of the two cells, and the ROE
  splitting at the Roe state is the textbook Roe flux F = (F_L + F_R)/2 - sum |lambda_i| a_i K_i / 2,
  written here independently in wave strengths a_i and eigenvectors K_i (Toro, Riemann Solvers and
  Numerical Methods for Fluid Dynamics, 3rd ed., sec. 11.2);
- where the velocity and p/rho are constant the CLLF eigenvalues equal the ROE-splitting ones at
  the arithmetic state;
- on a contact wave (velocity and pressure constant, density varying) moving with u > 0 only the
  positive part varies across the window, so the flux is the physical flux of the left-biased
  reconstruction -- the HLLC flux of PRIMITIVE reconstruction with the same stencil.  This fixes
  which side each half is reconstructed from.
"""
import numpy as np
import pytest

from oracle import jxf_oracle as o

RNG = np.random.default_rng(2024)
STENCILS = ["WENO5-JS", "WENO5-Z", "WENO3-Z", "TENO5", "WENO6-CU"]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _dom(n=(12, 9, 7), nh=3):
    return o.Domain(n, ((0, 1.0), (0, 0.9), (0, 1.1)), nh=nh)


def _buffers(W0, dom, frozen="ARITHMETIC", fs="ROE", stencil="WENO5-JS"):
    s = o.Setup(gamma=1.4, stencil=stencil, riemann="FLUX-SPLITTING-" + fs, frozen_state=frozen)
    W, U = o.create_buffers(W0, dom, s)
    return s, W, U


@pytest.mark.parametrize("fs", ["ROE", "CLLF", "LLF"])
@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_uniform_state_gives_physical_flux(fs, stencil, axis):
    dom = _dom()
    W0 = np.empty((5,) + dom.cells)
    for i, val in enumerate((1.3, 0.4, -0.7, 0.2, 2.1)):
        W0[i] = val
    s, W, U = _buffers(W0, dom, fs=fs, stencil=stencil)
    assert s.convective_solver == "FLUX-SPLITTING" and s.flux_splitting == fs
    F = o.compute_flux_xi(W, U, dom, axis, s)
    Fp = o.physical_flux_xi(W, U, axis)[(slice(None),) + (dom.nh,) * 3]
    assert rel(F, np.broadcast_to(Fp.reshape(5, 1, 1, 1), F.shape)) < 1e-13


def _smooth(dom, seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = np.meshgrid(*dom.cell_centers(), indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 5)
    return np.stack([1.0 + 0.3 * np.sin(2 * np.pi * X + ph[0]) * np.cos(2 * np.pi * Y),
                     0.6 * np.cos(2 * np.pi * (X + Z) + ph[1]),
                     -0.4 * np.sin(2 * np.pi * Y + ph[2]),
                     0.3 * np.cos(2 * np.pi * Z + ph[3]),
                     1.0 + 0.25 * np.sin(2 * np.pi * (X - Y) + ph[4])])


def _cells(W, U, dom, axis):
    WL, WR = o.face_LR_cells(W, dom, axis)
    UL, UR = o.face_LR_cells(U, dom, axis)
    return WL, WR, UL, UR


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_weno1_llf_is_rusanov(axis):
    dom = _dom()
    s, W, U = _buffers(_smooth(dom, 1), dom, fs="LLF", stencil="WENO1")
    F = o.compute_flux_xi(W, U, dom, axis, s)
    ref = o.rusanov(*_cells(W, U, dom, axis), axis, 1.4)
    assert rel(F, ref) < 1e-13


def _toro_roe_flux(WL, WR, UL, UR, axis, gamma):
    """Roe's flux in wave strengths (Toro sec. 11.2.2, rotated to the normal `axis`)."""
    v = axis + 1
    t0, t1 = [i for i in (1, 2, 3) if i != v]
    sL, sR = np.sqrt(WL[0]), np.sqrt(WR[0])
    HL = (UL[4] + WL[4]) / WL[0]
    HR = (UR[4] + WR[4]) / WR[0]
    avg = lambda a, b: (sL * a + sR * b) / (sL + sR)  # noqa: E731
    u, vt0, vt1, H = avg(WL[v], WR[v]), avg(WL[t0], WR[t0]), avg(WL[t1], WR[t1]), avg(HL, HR)
    q2 = u * u + vt0 * vt0 + vt1 * vt1
    c = np.sqrt((gamma - 1.0) * (H - 0.5 * q2))
    rho = sL * sR
    dr, dp = WR[0] - WL[0], WR[4] - WL[4]
    du, dt0, dt1 = WR[v] - WL[v], WR[t0] - WL[t0], WR[t1] - WL[t1]
    a1 = (dp - rho * c * du) / (2 * c * c)
    a2 = dr - dp / (c * c)
    a5 = (dp + rho * c * du) / (2 * c * c)

    def K(n, t0v, t1v, e):
        k = np.zeros((5,) + np.shape(u))
        k[0], k[v], k[t0], k[t1], k[4] = n, t0v[0], t0v[1], t0v[2], e
        return k

    one = np.ones_like(u)
    zero = np.zeros_like(u)
    K1 = K(one, (u - c, vt0, vt1), None, H - u * c)
    K2 = K(one, (u, vt0, vt1), None, 0.5 * q2)
    K3 = K(zero, (zero, one, zero), None, vt0)
    K4 = K(zero, (zero, zero, one), None, vt1)
    K5 = K(one, (u + c, vt0, vt1), None, H + u * c)
    diss = (np.abs(u - c) * a1 * K1 + np.abs(u) * a2 * K2 + np.abs(u) * rho * dt0 * K3
            + np.abs(u) * rho * dt1 * K4 + np.abs(u + c) * a5 * K5)
    FL = o.physical_flux_xi(WL, UL, axis)
    FR = o.physical_flux_xi(WR, UR, axis)
    return 0.5 * (FL + FR) - 0.5 * diss


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_weno1_roe_is_textbook_roe_flux(axis):
    dom = _dom()
    s, W, U = _buffers(_smooth(dom, 2), dom, frozen="ROE", fs="ROE", stencil="WENO1")
    F = o.compute_flux_xi(W, U, dom, axis, s)
    ref = _toro_roe_flux(*_cells(W, U, dom, axis), axis, 1.4)
    assert rel(F, ref) < 1e-12


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_cllf_equals_roe_eigenvalues_where_u_and_c_are_constant(stencil, axis):
    dom = _dom()
    W0 = _smooth(dom, 3)
    X, Y, Z = np.meshgrid(*dom.cell_centers(), indexing="ij")
    W0[0] = 1.0 + 0.3 * np.sin(2 * np.pi * (X + Y - Z)) + 0.2 * (Z > 0.5)
    W0[1], W0[2], W0[3] = 0.45, -0.2, 0.1
    W0[4] = 0.8 * W0[0]  # p / rho constant: one sound speed
    s, W, U = _buffers(W0, dom, fs="CLLF", stencil=stencil)
    F = o.compute_flux_xi(W, U, dom, axis, s)
    s_roe, _, _ = _buffers(W0, dom, fs="ROE", stencil=stencil)
    assert rel(F, o.compute_flux_xi(W, U, dom, axis, s_roe)) < 1e-12
    # and LLF differs there (its one eigenvalue |u| + c is larger than |u - c| and |u|)
    s_llf, _, _ = _buffers(W0, dom, fs="LLF", stencil=stencil)
    if stencil != "WENO1":
        assert rel(F, o.compute_flux_xi(W, U, dom, axis, s_llf)) > 1e-9


@pytest.mark.parametrize("fs", ["ROE", "CLLF"])
@pytest.mark.parametrize("stencil", ["WENO5-JS", "WENO5-Z", "TENO5", "WENO3-JS"])
@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_contact_wave_is_upwind_reconstruction(fs, stencil, sign, axis):
    dom = _dom()
    X, Y, Z = np.meshgrid(*dom.cell_centers(), indexing="ij")
    W0 = np.empty((5,) + dom.cells)
    W0[0] = 1.0 + 0.4 * np.sin(2 * np.pi * (X + 2 * Y - Z)) + 0.2 * (X > 0.5)
    W0[1], W0[2], W0[3] = 0.3, -0.25, 0.35
    W0[axis + 1] = sign * 0.3  # subsonic (c ~ 1.2), moving in the sign's direction
    W0[4] = 1.0
    s, W, U = _buffers(W0, dom, fs=fs, stencil=stencil)
    F = o.compute_flux_xi(W, U, dom, axis, s)
    g = o.Setup(gamma=1.4, stencil=stencil, variable="PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT")
    ref = o.compute_flux_xi(W, U, dom, axis, g)
    assert rel(F, ref) < 1e-12
