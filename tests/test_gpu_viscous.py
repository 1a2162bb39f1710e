"""GPU parity of the dissipative (viscous + heat) fluxes against the oracle (FP64, relative L-inf
<= 1e-12): RHS, fused RK3 steps with the viscous/thermal time-step limits, and the standalone
time step, on periodic domains (the stencils wrap) and on domains with ZEROGRADIENT and SYMMETRY
faces (the stencils read the face and edge halos)."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import jxf_oracle as o
from test_gpu_parity import dev, host, native_grid, random_state, rel, say

pytestmark = pytest.mark.gpu

TOL = 1e-12
FACES = ("east", "west", "north", "south", "top", "bottom")

TRANSPORTS = {
    "powerlaw-heat": o.Transport(R=1.0, mu_ref=2e-3, T_ref=1.0, mu_exp=0.76, bulk_viscosity=0.0, prandtl=0.71),
    "const-bulk": o.Transport(R=0.8, mu_ref=5e-3, bulk_viscosity=1e-3),
    "heat-only": o.Transport(R=1.2, mu_ref=3e-3, prandtl=0.9, viscous=False),
    "sutherland-prandtl": o.Transport(R=1.0, mu_ref=1.8e-3, T_ref=0.9, mu_sutherland=0.4, prandtl=0.72),
    "const-kappa": o.Transport(R=0.9, mu_ref=1e-3, kappa_ref=4e-3),
    "sutherland-kappa": o.Transport(R=1.1, mu_ref=2e-3, mu_exp=0.7, kappa_ref=3e-3, kappa_T_ref=0.8,
                                    kappa_sutherland=0.3),
}


def native_transport(tr):
    from jaxfluids_amd import _native

    return _native.make_transport(**dataclasses.asdict(tr))


def native_scheme_tr(s, tr):
    from jaxfluids_amd import _native

    return _native.make_scheme(s.variable, s.stencil, s.riemann, s.signal_speed, s.frozen_state, 0, s.gamma,
                               s.eps_stencil, native_transport(tr))


def _setup(tr, stencil="WENO5-Z", variable="CHAR-PRIMITIVE"):
    return o.Setup(gamma=1.4, stencil=stencil, variable=variable, riemann="HLLC", signal_speed="EINFELDT",
                   bcs={f: "PERIODIC" for f in FACES}, transport=tr)


@pytest.mark.parametrize("name", list(TRANSPORTS))
@pytest.mark.parametrize("cells", [(11, 9, 8), (13, 10, 1)], ids=["3d", "2d"])
def test_dissipative_rhs(name, cells):
    from jaxfluids_amd import _native

    say(f"[visc rhs] {name} {cells}")
    tr = TRANSPORTS[name]
    s = _setup(tr)
    dom, W0 = random_state(cells, 3, seed=41, jump=False)
    if cells[2] == 1:
        W0 = W0[..., :1]
    W, U = o.create_buffers(W0, dom, s)
    ref = o.compute_rhs(W, U, dom, s)
    lib = _native.load()
    grid = native_grid(dom)
    scheme = native_scheme_tr(s, tr)
    ws = torch.empty(lib.jfa_rhs_workspace_bytes(grid, 0), dtype=torch.uint8, device="cuda")
    out = torch.empty((5,) + dom.cells, dtype=torch.float64, device="cuda")
    Wd, Ud = dev(W), dev(U)
    rc = lib.jfa_compute_rhs(_native.ptr(Wd), _native.ptr(Ud), grid, scheme, 0.0, _native.ptr(out), _native.ptr(ws),
                             ws.numel(), _native.stream_ptr())
    _native.check(rc, "jfa_compute_rhs")
    got = host(out)
    assert rel(got, ref) < TOL, rel(got, ref)
    # the dissipative part is not negligible here
    s0 = _setup(None)
    conv = o.compute_rhs(W, U, dom, s0)
    assert np.max(np.abs(ref - conv)) > 1e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("name", ["powerlaw-heat", "const-bulk", "sutherland-prandtl", "sutherland-kappa"])
@pytest.mark.parametrize("cells", [(10, 8, 9), (12, 11, 1)], ids=["3d", "2d"])
def test_dissipative_rk3_step_and_dt(name, cells):
    """One fused RK3 step (3 stages, halos, next dt with the viscous / thermal limits)."""
    from jaxfluids_amd import _native

    say(f"[visc step] {name} {cells}")
    tr = TRANSPORTS[name]
    s = _setup(tr, stencil="WENO5-JS", variable="PRIMITIVE")
    dom, W0 = random_state(cells, 4, seed=43, jump=False)
    if cells[2] == 1:
        W0 = W0[..., :1]
    W, U = o.create_buffers(W0, dom, s)
    dt = o.compute_dt(W, dom, s)
    Wr, Ur = o.rk3_step(W, U, dt, dom, s)
    lib = _native.load()
    grid, scheme = native_grid(dom), native_scheme_tr(s, tr)
    bcs = _native.bc_array({f: "PERIODIC" for f in FACES})
    ws = torch.empty(lib.jfa_stage_workspace_bytes(grid, scheme), dtype=torch.uint8, device="cuda")
    Wd, Ud = dev(W), dev(U)
    Wb, Ub, Wc, Uc = (torch.empty_like(Wd) for _ in range(4))
    dtd = torch.full((1,), dt, dtype=torch.float64, device="cuda")
    P = _native.ptr
    chain = [(Wd, Ud, Wb, Ub), (Wb, Ub, Wc, Uc), (Wc, Uc, Wd, Ud)]
    for k, (wi, ui, wo, uo) in enumerate(chain):
        rc = lib.jfa_rk3_stage_fused(P(wi), P(ui), P(Ud if k else None), P(wo), P(uo), grid, scheme, bcs, k, P(dtd),
                                     0.0, s.cfl, o.EPS64, P(dtd) if k == 2 else P(None), P(ws), ws.numel(),
                                     _native.stream_ptr())
        _native.check(rc, "jfa_rk3_stage_fused")
    I = dom.interior
    assert rel(host(Wd)[I], Wr[I]) < TOL and rel(host(Ud)[I], Ur[I]) < TOL
    assert abs(host(dtd)[0] - o.compute_dt(Wr, dom, s)) <= 1e-14 * dt


def test_dissipative_dt_limits():
    """jfa_compute_dt with a transport: convective, viscous and thermal limits (a viscosity large
    enough that the viscous limit binds, then a conductivity large enough that the thermal one does)."""
    from jaxfluids_amd import _native

    say("[visc dt]")
    dom, W0 = random_state((12, 10, 8), 3, seed=45, jump=False)
    for tr in (o.Transport(mu_ref=0.5), o.Transport(mu_ref=0.05, prandtl=0.01), o.Transport(mu_ref=1e-6, prandtl=0.7),
               o.Transport(mu_ref=0.4, mu_sutherland=0.3), o.Transport(mu_ref=1e-6, kappa_ref=2.0, kappa_exp=0.5),
               o.Transport(mu_ref=0.01, viscous=False, kappa_ref=1.5, kappa_sutherland=0.2)):
        s = _setup(tr)
        W, _ = o.create_buffers(W0, dom, s)
        ref = o.compute_dt(W, dom, s)
        dt = torch.empty(1, dtype=torch.float64, device="cuda")
        ws = torch.empty(32, dtype=torch.uint8, device="cuda")
        lib = _native.load()
        Wd = dev(W)
        rc = lib.jfa_compute_dt(_native.ptr(Wd), native_grid(dom), 1.4, s.cfl, o.EPS64, 0,
                                native_transport(tr), _native.ptr(dt), _native.ptr(ws), _native.stream_ptr())
        _native.check(rc, "jfa_compute_dt")
        assert abs(host(dt)[0] - ref) <= 1e-14 * ref, (tr, host(dt)[0], ref)


# non-periodic boundary sets: every edge type (ANY_ANY, PERIODIC, SYMMETRY on either face)
NONPERIODIC_BCS = {
    "zerogradient": {f: "ZEROGRADIENT" for f in FACES},
    "symmetry": {f: "SYMMETRY" for f in FACES},
    "mixed": {"east": "PERIODIC", "west": "PERIODIC", "north": "SYMMETRY", "south": "ZEROGRADIENT",
              "top": "ZEROGRADIENT", "bottom": "SYMMETRY"},
    "upper-symmetry": {"east": "ZEROGRADIENT", "west": "ZEROGRADIENT", "north": "SYMMETRY", "south": "SYMMETRY",
                       "top": "SYMMETRY", "bottom": "ZEROGRADIENT"},
}


def _regions(dom):
    """Interior, face-halo slabs and edge-halo blocks of the buffer."""
    parts = (slice(0, dom.nh), slice(-dom.nh, None))
    out = [dom.interior]
    for i, a in enumerate(dom.active):
        for p in parts:
            sl = [dom.dom(0), dom.dom(1), dom.dom(2)]
            sl[a] = p
            out.append((Ellipsis,) + tuple(sl))
        for b in dom.active[i + 1:]:
            for pa in parts:
                for pb in parts:
                    sl = [dom.dom(0), dom.dom(1), dom.dom(2)]
                    sl[a], sl[b] = pa, pb
                    out.append((Ellipsis,) + tuple(sl))
    return out


@pytest.mark.parametrize("bcs", list(NONPERIODIC_BCS))
@pytest.mark.parametrize("cells", [(10, 8, 9), (12, 11, 1)], ids=["3d", "2d"])
def test_dissipative_rk3_step_non_periodic(bcs, cells):
    """One fused RK3 step with viscous and heat fluxes on a domain with walls and outflow faces:
    the new state, its face and edge halos, and the next dt."""
    from jaxfluids_amd import _native

    say(f"[visc step non-periodic] {bcs} {cells}")
    tr = TRANSPORTS["powerlaw-heat"]
    s = o.Setup(gamma=1.4, stencil="WENO5-Z", variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                bcs=NONPERIODIC_BCS[bcs], transport=tr)
    dom, W0 = random_state(cells, 3, seed=49, jump=False)
    if cells[2] == 1:
        W0 = W0[..., :1]
    W, U = o.create_buffers(W0, dom, s)
    dt = o.compute_dt(W, dom, s)
    Wr, Ur = o.rk3_step(W, U, dt, dom, s)
    lib = _native.load()
    grid, scheme = native_grid(dom), native_scheme_tr(s, tr)
    bc = _native.bc_array({f: (s.bcs[f] if o.FACE_AXIS[f] in dom.active else "PERIODIC") for f in FACES})
    ws = torch.empty(lib.jfa_stage_workspace_bytes(grid, scheme), dtype=torch.uint8, device="cuda")
    Wd, Ud = dev(W), dev(U)
    Wb, Ub, Wc, Uc = (torch.full_like(Wd, float("nan")) for _ in range(4))
    dtd = torch.full((1,), dt, dtype=torch.float64, device="cuda")
    P = _native.ptr
    chain = [(Wd, Ud, Wb, Ub), (Wb, Ub, Wc, Uc), (Wc, Uc, Wd, Ud)]
    for k, (wi, ui, wo, uo) in enumerate(chain):
        rc = lib.jfa_rk3_stage_fused(P(wi), P(ui), P(Ud if k else None), P(wo), P(uo), grid, scheme, bc, k, P(dtd),
                                     0.0, s.cfl, o.EPS64, P(dtd) if k == 2 else P(None), P(ws), ws.numel(),
                                     _native.stream_ptr())
        _native.check(rc, "jfa_rk3_stage_fused")
    Wg, Ug = host(Wd), host(Ud)
    for r in _regions(dom):
        assert rel(Wg[r], Wr[r]) < TOL and rel(Ug[r], Ur[r]) < TOL, (r, rel(Wg[r], Wr[r]))
    assert abs(host(dtd)[0] - o.compute_dt(Wr, dom, s)) <= 1e-14 * dt
    # the boundaries matter: the periodic step differs
    sp = dataclasses.replace(s, bcs={f: "PERIODIC" for f in FACES})
    Wp, _ = o.rk3_step(*o.create_buffers(W0, dom, sp), dt, dom, sp)
    assert rel(Wp[dom.interior], Wr[dom.interior]) > 1e-6


@pytest.mark.parametrize("bcs", ["zerogradient", "mixed"])
def test_dissipative_rhs_non_periodic(bcs):
    """jfa_compute_rhs with the transport's halos flag reads the buffer's face and edge halos."""
    from jaxfluids_amd import _native

    say(f"[visc rhs non-periodic] {bcs}")
    tr = TRANSPORTS["sutherland-kappa"]
    s = o.Setup(gamma=1.4, stencil="WENO5-Z", variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                bcs=NONPERIODIC_BCS[bcs], transport=tr)
    dom, W0 = random_state((11, 9, 8), 3, seed=51, jump=False)
    W, U = o.create_buffers(W0, dom, s)
    ref = o.compute_rhs(W, U, dom, s)
    lib = _native.load()
    grid = native_grid(dom)
    scheme = _native.make_scheme(s.variable, s.stencil, s.riemann, s.signal_speed, s.frozen_state, 0, s.gamma,
                                 s.eps_stencil, _native.make_transport(**dataclasses.asdict(tr), halos=True))
    ws = torch.empty(lib.jfa_rhs_workspace_bytes(grid, 0), dtype=torch.uint8, device="cuda")
    out = torch.empty((5,) + dom.cells, dtype=torch.float64, device="cuda")
    Wd, Ud = dev(W), dev(U)
    rc = lib.jfa_compute_rhs(_native.ptr(Wd), _native.ptr(Ud), grid, scheme, 0.0, _native.ptr(out), _native.ptr(ws),
                             ws.numel(), _native.stream_ptr())
    _native.check(rc, "jfa_compute_rhs")
    assert rel(host(out), ref) < TOL, rel(host(out), ref)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("bcs", list(NONPERIODIC_BCS))
def test_edge_halo_update(bcs, dtype):
    """jfa_edge_halo_update on random face halos: W as the oracle's edge fill (bit-exact in FP64),
    U = cons(W); the rest of the buffer untouched."""
    from jaxfluids_amd import _native

    say(f"[edge halos] {bcs} {dtype}")
    dom, _ = random_state((9, 7, 6), 3, seed=53, jump=False)
    s = o.Setup(gamma=1.4, bcs=NONPERIODIC_BCS[bcs])
    rng = np.random.default_rng(7)
    W = 1.0 + rng.random((5,) + dom.shape)
    U = rng.random((5,) + dom.shape)
    if dtype == torch.float32:
        W, U = W.astype(np.float32).astype(np.float64), U.astype(np.float32).astype(np.float64)
    Wr, Ur = o.edge_halo_update(W, U, dom, s)
    Wd, Ud = torch.tensor(W, dtype=dtype, device="cuda"), torch.tensor(U, dtype=dtype, device="cuda")
    bc = _native.bc_array(s.bcs)
    rc = _native.load().jfa_edge_halo_update(_native.ptr(Wd), _native.ptr(Ud), native_grid(dom), bc, 1.4,
                                             _native.dtype_code(Wd), _native.stream_ptr())
    _native.check(rc, "jfa_edge_halo_update")
    Wg, Ug = host(Wd).astype(np.float64), host(Ud).astype(np.float64)
    if dtype == torch.float64:
        assert np.array_equal(Wg, Wr)
        assert rel(Ug, Ur) <= 4e-16
    else:
        assert rel(Wg, Wr) <= 6e-8 and rel(Ug, Ur) <= 3e-7
    changed = np.any(Wr != W, axis=0)
    assert np.array_equal(np.any(Wg != W, axis=0), changed)


@pytest.mark.parametrize("fused", [True, False])
def test_viscous_tgv_through_simulation_manager(fused):
    """Viscous Taylor-Green vortex 16^3 from the reference's setup dicts (power-law CUSTOM viscosity
    lambda, PRANDTL conductivity, CENTRAL2 dissipative stencils): 4 CFL steps through
    InputManager -> InitializationManager -> SimulationManager against the oracle."""
    import copy

    from jaxfluids_amd import InitializationManager, InputManager, SimulationManager
    from test_setup_readers import NUM, _case

    say(f"[visc tgv] 16^3 x 4 steps fused={fused}")
    dom, W0 = o.tgv_initial(16, nh=3)
    case = _case()
    for a in "xyz":
        case["domain"][a]["cells"] = 16
    case["general"]["end_step"] = 4
    num = copy.deepcopy(NUM)
    im = InputManager(case, num)
    tr = o.Transport(R=1.0, mu_ref=0.0028, T_ref=1.0, mu_exp=0.76, prandtl=0.71)
    s = o.Setup(gamma=1.4, stencil="WENO5-Z", variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                transport=tr)
    W, U = o.create_buffers(W0, dom, s)
    Wr, Ur, t_ref, n_ref = o.simulate(W, U, dom, s, end_step=4)
    bufs = InitializationManager(im).initialization(W0)
    out = SimulationManager(im, fused=fused).simulate(bufs)
    assert out.time_control_variables.simulation_step == n_ref
    assert abs(out.time_control_variables.physical_simulation_time - t_ref) <= 1e-13 * t_ref
    I = dom.interior
    assert rel(host(out.material_fields.primitives)[I], Wr[I]) < TOL


@pytest.mark.parametrize("fused", [True, False])
def test_viscous_walls_through_simulation_manager(fused):
    """The viscous Taylor-Green setup with SYMMETRY walls on x, ZEROGRADIENT faces on y and
    PERIODIC z (16^3, 4 CFL steps through the setup dicts): the initialization fills the edge
    halos, every stage refills them, against the oracle."""
    import copy

    from jaxfluids_amd import InitializationManager, InputManager, SimulationManager
    from test_setup_readers import NUM, _case

    say(f"[visc walls] 16^3 x 4 steps fused={fused}")
    bcs = {"east": "SYMMETRY", "west": "SYMMETRY", "north": "ZEROGRADIENT", "south": "ZEROGRADIENT",
           "top": "PERIODIC", "bottom": "PERIODIC"}
    dom, W0 = o.tgv_initial(16, nh=3)
    case = _case()
    for a in "xyz":
        case["domain"][a]["cells"] = 16
    case["general"]["end_step"] = 4
    case["boundary_conditions"] = {f: {"type": t} for f, t in bcs.items()}
    im = InputManager(case, copy.deepcopy(NUM))
    tr = o.Transport(R=1.0, mu_ref=0.0028, T_ref=1.0, mu_exp=0.76, prandtl=0.71)
    s = o.Setup(gamma=1.4, stencil="WENO5-Z", variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                transport=tr, bcs=bcs)
    W, U = o.create_buffers(W0, dom, s)
    Wr, Ur, t_ref, n_ref = o.simulate(W, U, dom, s, end_step=4)
    bufs = InitializationManager(im).initialization(W0)
    out = SimulationManager(im, fused=fused).simulate(bufs)
    assert out.time_control_variables.simulation_step == n_ref
    assert abs(out.time_control_variables.physical_simulation_time - t_ref) <= 1e-13 * t_ref
    Wg = host(out.material_fields.primitives)
    for r in _regions(dom):
        assert rel(Wg[r], Wr[r]) < TOL, r
