"""The stage VJP (jfa_rk3_stage_vjp, host build of the library sources) against reverse-mode
automatic differentiation of an independent restatement: torch.autograd of oracle/torch_stage.py,
the stand-in for jax.grad of the reference.  Unlike the finite-difference checks
(tests/test_emu_adjoint.py), this compares every entry of J^T lam to rounding level.

Pins of the restatement itself: its forward equals the NumPy oracle's stage map, and the
derivative conventions at non-smooth points are jax's (d|x|/dx = 0 and d sign/dx = 0 at 0,
maximum / minimum split the cotangent evenly at a tie: lax.max's balanced JVP).
"""
import numpy as np
import pytest
import torch

from oracle import jxf_oracle as o
from oracle import torch_stage as ts
from test_emu_adjoint import CASES, FACES, VISCOUS_CASES, emu, run_vjp, smooth_state, stage_map  # noqa: F401

# Both sides compute in FP64 with different association (hand-derived reverse formulas against
# autograd's chain of elementwise derivatives); the derivatives of the WENO weights amplify
# rounding (terms in tau / beta^2), and the largest differences measured over these cases are
# 1.2e-12 of the largest entry.  A wrong or missing term of the reverse sweep is orders larger.
TOL = 1e-11


def _setup(case, variable="PRIMITIVE"):
    cells, bcs, (stencil, riemann, speed) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable=variable, riemann=riemann, signal_speed=speed,
                frozen_state="ARITHMETIC", bcs=dict(zip(FACES, bcs)))
    return dom, s


def _viscous_setup(case):
    cells, (stencil, riemann, speed), tr = VISCOUS_CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable="PRIMITIVE", riemann=riemann, signal_speed=speed,
                bcs={f: "PERIODIC" for f in FACES}, transport=tr)
    return dom, s


def _torch_map(dom, s, stage, dt, Un_int):
    return lambda U: ts.stage_map(U, torch.as_tensor(Un_int), dom, s.stencil, s.riemann, s.signal_speed, s.gamma,
                                  s.bcs, stage, dt, o.EPS64, s.transport, s.variable, s.frozen_state)


def _compare(emu, dom, s, stage, sweep, seed_state, seed):  # noqa: F811
    rng = np.random.default_rng(seed)
    W0 = smooth_state(dom, seed_state)
    Uint = o.cons_from_prim(W0, s.gamma)
    Un_int = Uint * (1.0 + 0.01 * rng.standard_normal(Uint.shape))
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    _, W_in = stage_map(dom, s, stage, dt, Un_int)(Uint)
    lam_int = rng.standard_normal(Uint.shape)
    lam_out = np.zeros((5,) + dom.shape)
    lam_out[dom.interior] = lam_int
    lam_in, lam_n = run_vjp(emu, dom, s, W_in, lam_out, stage, dt, sweep)
    U = torch.as_tensor(Uint).clone().requires_grad_(True)
    out = _torch_map(dom, s, stage, dt, Un_int)(U)
    (ref,) = torch.autograd.grad(out, U, torch.as_tensor(lam_int))
    ref = ref.numpy()
    err = np.max(np.abs(lam_in[dom.interior] - ref)) / np.max(np.abs(ref))
    assert err <= TOL, err
    if stage:
        c1 = o.RK3_CONS_MULTIPLIER[stage - 1][1]
        assert np.array_equal(lam_n[dom.interior], c1 * lam_int)


def test_tie_conventions_are_jax_conventions():
    """torch's derivatives at the kinks of abs, sign, maximum and minimum, which the
    restatement relies on to stand in for jax.grad."""
    x = torch.tensor([0.0, 2.0], dtype=torch.float64, requires_grad=True)
    torch.abs(x).sum().backward()
    assert x.grad.tolist() == [0.0, 1.0]
    x = torch.tensor([0.0, -3.0], dtype=torch.float64, requires_grad=True)
    torch.sign(x).sum().backward()
    assert x.grad.tolist() == [0.0, 0.0]
    for fn in (torch.maximum, torch.minimum):
        a = torch.tensor([1.5, 1.0], dtype=torch.float64, requires_grad=True)
        b = torch.tensor([1.5, 2.0], dtype=torch.float64, requires_grad=True)
        fn(a, b).sum().backward()
        assert a.grad[0].item() == 0.5 and b.grad[0].item() == 0.5, fn
        assert a.grad[1].item() + b.grad[1].item() == 1.0


@pytest.mark.parametrize("variable", ["PRIMITIVE", "CONSERVATIVE", "CHAR-PRIMITIVE"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_restatement_forward_matches_oracle(case, variable):
    dom, s = _setup(case, variable)
    W0 = smooth_state(dom, 50 + case)
    Uint = o.cons_from_prim(W0, s.gamma)
    Un_int = 1.01 * Uint
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    for stage in range(3):
        ref, _ = stage_map(dom, s, stage, dt, Un_int)(Uint)
        got = _torch_map(dom, s, stage, dt, Un_int)(torch.as_tensor(Uint)).numpy()
        assert np.max(np.abs(got - ref)) <= 1e-15 * np.max(np.abs(ref)), stage


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("stage", [0, 2])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_stage_vjp_matches_autograd(emu, case, stage, sweep):  # noqa: F811
    """lam_in = J^T lam_out over the interior, and lam_n = c1 lam_out, entry by entry."""
    dom, s = _setup(case)
    _compare(emu, dom, s, stage, sweep, 50 + case, 3000 + 10 * case + stage)


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("stage", [0, 2])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_char_primitive_stage_vjp_matches_autograd(emu, case, stage, sweep):  # noqa: F811
    """CHAR-PRIMITIVE reconstruction (ARITHMETIC frozen state): the reverse of the projection, of
    the frozen state's dependence on the face's two cells and of the back projection."""
    dom, s = _setup(case, "CHAR-PRIMITIVE")
    _compare(emu, dom, s, stage, sweep, 50 + case, 7000 + 10 * case + stage)


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("stage", [0, 2])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_conservative_stage_vjp_matches_autograd(emu, case, stage, sweep):  # noqa: F811
    """CONSERVATIVE reconstruction: WENO on the conservatives of the window cells (their
    dependence on the primitives, U = cons(W)), then the face's primitives prim(U_L), prim(U_R)."""
    dom, s = _setup(case, "CONSERVATIVE")
    _compare(emu, dom, s, stage, sweep, 50 + case, 8000 + 10 * case + stage)


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("variable", ["PRIMITIVE", "CONSERVATIVE", "CHAR-PRIMITIVE"])
@pytest.mark.parametrize("case", [0, 1, 3, 6])
def test_toro_speeds_stage_vjp_matches_autograd(emu, case, variable, sweep):  # noqa: F811
    """HLLC with Toro's pressure-based wave speeds (the PVRS estimate, its max with 0 and the
    shock / rarefaction factor q_K) on the boundary cases' grids."""
    dom, s = _setup(case, variable)
    s.riemann, s.signal_speed = "HLLC", "TORO"
    _compare(emu, dom, s, 2, sweep, 50 + case, 9000 + 10 * case)


@pytest.mark.parametrize("variable, frozen", [("CHAR-PRIMITIVE", "ROE"), ("CHAR-CONSERVATIVE", "ARITHMETIC"),
                                              ("CHAR-CONSERVATIVE", "ROE")])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_frozen_state_forward_matches_oracle(case, variable, frozen):
    dom, s = _setup(case, variable)
    s.frozen_state = frozen
    W0 = smooth_state(dom, 50 + case)
    Uint = o.cons_from_prim(W0, s.gamma)
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    for stage in range(3):
        ref, _ = stage_map(dom, s, stage, dt, 1.01 * Uint)(Uint)
        got = _torch_map(dom, s, stage, dt, 1.01 * Uint)(torch.as_tensor(Uint)).numpy()
        assert np.max(np.abs(got - ref)) <= 1e-15 * np.max(np.abs(ref)), stage


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("stage", [0, 2])
@pytest.mark.parametrize("case", [0, 1, 3, 6])
def test_roe_char_primitive_stage_vjp_matches_autograd(emu, case, stage, sweep):  # noqa: F811
    """CHAR-PRIMITIVE with the ROE frozen state: the projection's dependence on the Roe density
    sqrt(rho_L rho_R) and on c^2 = psi + g (psi + rho~ (sqrt rho_L + sqrt rho_R)^-2 |du|^2 / 2),
    which reaches every primitive of the face's two cells."""
    dom, s = _setup(case, "CHAR-PRIMITIVE")
    s.frozen_state = "ROE"
    _compare(emu, dom, s, stage, sweep, 50 + case, 9500 + 10 * case + stage)


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("frozen", ["ARITHMETIC", "ROE"])
@pytest.mark.parametrize("stage", [0, 2])
@pytest.mark.parametrize("case", [0, 1, 3, 6])
def test_char_conservative_stage_vjp_matches_autograd(emu, case, stage, frozen, sweep):  # noqa: F811
    """CHAR-CONSERVATIVE: the conservative window projected with the left eigenvectors, WENO5
    per characteristic, back with the right eigenvectors; the eigenvectors' dependence on the
    frozen state (rho~, u~, H~, c) of the face's two cells."""
    dom, s = _setup(case, "CHAR-CONSERVATIVE")
    s.frozen_state = frozen
    _compare(emu, dom, s, stage, sweep, 50 + case, 9600 + 10 * case + stage)


EXT_STENCILS = ["WENO1", "WENO3-JS", "WENO3-Z", "TENO5", "TENO5-A"]


@pytest.mark.parametrize("variable", ["PRIMITIVE", "CHAR-PRIMITIVE", "CHAR-CONSERVATIVE"])
@pytest.mark.parametrize("stencil", EXT_STENCILS)
@pytest.mark.parametrize("case", [0, 1, 6])
def test_stencil_restatement_forward_matches_oracle(case, stencil, variable):
    dom, s = _setup(case, variable)
    s.stencil = stencil
    W0 = smooth_state(dom, 50 + case)
    Uint = o.cons_from_prim(W0, s.gamma)
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    ref, _ = stage_map(dom, s, 2, dt, 1.01 * Uint)(Uint)
    got = _torch_map(dom, s, 2, dt, 1.01 * Uint)(torch.as_tensor(Uint)).numpy()
    assert np.max(np.abs(got - ref)) <= 1e-15 * np.max(np.abs(ref))


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("case, stencil, variable",
                         [(1, st, var) for st in EXT_STENCILS for var in ("PRIMITIVE", "CHAR-PRIMITIVE")]
                         + [(6, "WENO3-Z", "CONSERVATIVE"), (6, "TENO5", "PRIMITIVE"), (0, "TENO5-A", "CHAR-CONSERVATIVE")])
def test_stencil_stage_vjp_matches_autograd(emu, case, stencil, variable, sweep):  # noqa: F811
    """WENO1, WENO3-JS / -Z (their weights' derivatives) and TENO5 / TENO5-A (weights constant
    between cutoff switches, as jax differentiates the comparison)."""
    dom, s = _setup(case, variable)
    s.stencil = stencil
    _compare(emu, dom, s, 2, sweep, 50 + case, 9100 + 10 * case)


@pytest.mark.parametrize("riemann", ["HLL", "HLLC-LM", "AUSMP"])
@pytest.mark.parametrize("variable", ["PRIMITIVE", "CHAR-PRIMITIVE"])
@pytest.mark.parametrize("speed", ["EINFELDT", "DAVIS", "TORO"])
@pytest.mark.parametrize("case", [0, 1, 6])
def test_hll_restatement_forward_matches_oracle(case, speed, variable, riemann):
    dom, s = _setup(case, variable)
    s.riemann, s.signal_speed = riemann, speed
    W0 = smooth_state(dom, 50 + case)
    Uint = o.cons_from_prim(W0, s.gamma)
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    ref, _ = stage_map(dom, s, 1, dt, 1.01 * Uint)(Uint)
    got = _torch_map(dom, s, 1, dt, 1.01 * Uint)(torch.as_tensor(Uint)).numpy()
    assert np.max(np.abs(got - ref)) <= 1e-15 * np.max(np.abs(ref))


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("case, speed, variable",
                         [(1, sp, var) for sp in ("EINFELDT", "ARITHMETIC", "RUSANOV", "DAVIS", "TORO")
                          for var in ("PRIMITIVE", "CONSERVATIVE", "CHAR-PRIMITIVE")]
                         + [(0, "EINFELDT", "PRIMITIVE"), (6, "EINFELDT", "CHAR-PRIMITIVE"), (6, "TORO", "CONSERVATIVE")])
def test_hll_stage_vjp_matches_autograd(emu, case, speed, variable, sweep):  # noqa: F811
    """HLL with every signal speed: the two-wave flux, its clipped speeds min(S_L, 0),
    max(S_R, 0) and the precision epsilon in the denominator."""
    dom, s = _setup(case, variable)
    s.riemann, s.signal_speed = "HLL", speed
    _compare(emu, dom, s, 2, sweep, 50 + case, 9300 + 10 * case)


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("case, speed, variable",
                         [(1, sp, "PRIMITIVE") for sp in ("EINFELDT", "ARITHMETIC", "RUSANOV", "DAVIS", "TORO")]
                         + [(1, "EINFELDT", "CHAR-PRIMITIVE"), (6, "EINFELDT", "CONSERVATIVE"),
                            (0, "DAVIS", "CHAR-CONSERVATIVE")])
def test_hllc_lm_stage_vjp_matches_autograd(emu, case, speed, variable, sweep):  # noqa: F811
    """HLLC-LM: the HLLC star states, the low-Mach factor phi = sin(min(1, Ma / 0.1) pi / 2),
    |S*| and the (piecewise constant) sign blends."""
    dom, s = _setup(case, variable)
    s.riemann, s.signal_speed = "HLLC-LM", speed
    _compare(emu, dom, s, 2, sweep, 50 + case, 9400 + 10 * case)


@pytest.mark.parametrize("kind", ["ROE", "CLLF", "LLF"])
@pytest.mark.parametrize("frozen", ["ARITHMETIC", "ROE"])
@pytest.mark.parametrize("case", [0, 1, 6])
def test_flux_splitting_restatement_forward_matches_oracle(case, frozen, kind):
    dom, s = _setup(case)
    s.riemann, s.signal_speed, s.frozen_state = "FLUX-SPLITTING-" + kind, None, frozen
    s.convective_solver, s.flux_splitting = "FLUX-SPLITTING", kind
    W0 = smooth_state(dom, 50 + case)
    Uint = o.cons_from_prim(W0, s.gamma)
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    ref, _ = stage_map(dom, s, 2, dt, 1.01 * Uint)(Uint)
    got = _torch_map(dom, s, 2, dt, 1.01 * Uint)(torch.as_tensor(Uint)).numpy()
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("case, kind, frozen, stencil",
                         [(1, k, "ARITHMETIC", "WENO5-JS") for k in ("ROE", "CLLF", "LLF")]
                         + [(6, "ROE", "ROE", "WENO5-Z"), (0, "CLLF", "ROE", "WENO3-Z")])
def test_flux_splitting_stage_vjp_matches_autograd(emu, case, kind, frozen, stencil, sweep):  # noqa: F811
    """The flux-splitting scheme: the split characteristic fluxes of the window, their WENO
    reconstructions, the eigenvalue magnitudes and both eigenvector matrices' dependence on the
    frozen state."""
    dom, s = _setup(case)
    s.riemann, s.signal_speed, s.frozen_state, s.stencil = "FLUX-SPLITTING-" + kind, None, frozen, stencil
    s.convective_solver, s.flux_splitting = "FLUX-SPLITTING", kind
    _compare(emu, dom, s, 2, sweep, 50 + case, 9700 + 10 * case)


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("case, variable", [(0, "PRIMITIVE"), (1, "PRIMITIVE"), (1, "CHAR-PRIMITIVE"),
                                            (6, "CONSERVATIVE"), (3, "PRIMITIVE")])
def test_ausmp_stage_vjp_matches_autograd(emu, case, variable, sweep):  # noqa: F811
    """AUSM+: the split Mach numbers and pressures (subsonic polynomials at these states), the
    interface speed of sound and M+- = (M +- |M|)/2."""
    dom, s = _setup(case, variable)
    s.riemann, s.signal_speed = "AUSMP", None
    _compare(emu, dom, s, 2, sweep, 50 + case, 9450 + 10 * case)


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("stage", [0, 2])
@pytest.mark.parametrize("case", range(len(VISCOUS_CASES)))
def test_viscous_stage_vjp_matches_autograd(emu, case, stage, sweep):  # noqa: F811
    """The same with viscous and / or heat fluxes (periodic domains: power law, Sutherland,
    Prandtl, own conductivity laws, bulk viscosity, heat only)."""
    dom, s = _viscous_setup(case)
    _compare(emu, dom, s, stage, sweep, 80 + case, 4000 + 10 * case + stage)


def test_viscous_forward_matches_oracle():
    for case in range(len(VISCOUS_CASES)):
        dom, s = _viscous_setup(case)
        W0 = smooth_state(dom, 80 + case)
        Uint = o.cons_from_prim(W0, s.gamma)
        for stage in range(3):
            ref, _ = stage_map(dom, s, stage, 1e-3, 1.01 * Uint)(Uint)
            got = _torch_map(dom, s, stage, 1e-3, 1.01 * Uint)(torch.as_tensor(Uint)).numpy()
            assert np.max(np.abs(got - ref)) <= 1e-15 * np.max(np.abs(ref)), (case, stage)

