"""GPU parity of the HIP path against the oracle (FP64, relative L-inf <= 1e-12).

Each test prints the case it runs before touching the GPU.  Inputs are seeded;
sizes are small enough for the NumPy oracle to finish in seconds.
"""
import math

import numpy as np
import pytest
import torch

from oracle import jxf_oracle as o

pytestmark = pytest.mark.gpu

TOL = 1e-12
FACES = ("east", "west", "north", "south", "top", "bottom")


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def dev(x, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def say(msg):
    print(msg, flush=True)


def random_state(cells, nh, seed, jump=True, ranges=None):
    """Smooth periodic field plus (optionally) a density/pressure jump."""
    rng = np.random.default_rng(seed)
    ranges = ranges or tuple((0.0, 1.0) for _ in range(3))
    dom = o.Domain(cells, ranges, nh=nh)
    xs = dom.cell_centers()
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 8)
    k = 2 * np.pi
    rho = 1.0 + 0.3 * np.sin(k * X + ph[0]) * np.cos(k * Y + ph[1]) + 0.1 * np.cos(k * Z + ph[2])
    u = 0.5 * np.sin(k * Y + ph[2]) + 0.2 * np.cos(k * Z + ph[3])
    v = 0.4 * np.cos(k * X + ph[4]) - 0.1 * np.sin(k * Z)
    w = 0.3 * np.sin(k * Z + ph[5]) * np.sin(k * X + ph[6])
    p = 1.0 + 0.2 * np.cos(k * X + k * Y + ph[7])
    if jump:
        m = (X + 0.3 * Y + 0.2 * Z) > 0.55
        rho = np.where(m, 0.25 * rho, rho)
        p = np.where(m, 0.2 * p, p)
    W = np.stack([rho, u, v, w, p])
    # Seeded noise breaks exact plateau ties [a, b, b, b, a]: there beta1 = 0 and
    # beta0 == beta2 in exact arithmetic, so WENO-Z's tau5 is pure rounding and
    # tau5 / (beta1 + 1e-30) decides the weights -- no two evaluation orders agree there.
    W = W * (1.0 + 1e-3 * rng.standard_normal(W.shape))
    return dom, W


def native_grid(dom):
    from jaxfluids_amd import _native

    return _native.make_grid(dom.cells, dom.nh, dom.dx)


def native_scheme(s, dtype=0):
    from jaxfluids_amd import _native

    return _native.make_scheme(s.variable, s.stencil, s.riemann, s.signal_speed, s.frozen_state, dtype, s.gamma,
                               s.eps_stencil, None, s.interpolation_limiter, s.limit_velocity, s.limiter_eps,
                               s.flux_limiter, s.flux_partition, s.flux_limiter_eps)


# ------------------------------------------------------------------ reconstruct_xi
@pytest.mark.parametrize("stencil", ["WENO5-JS", "WENO5-Z"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_weno5_reconstruct_parity(stencil, axis):
    from jaxfluids_amd import WENO5JS, WENO5Z

    say(f"[weno5_reconstruct] {stencil} axis={axis}")
    dom, W0 = random_state((9, 8, 7), 4, seed=11 + axis)
    s = o.Setup()
    W, U = o.create_buffers(W0, dom, s)
    cls = {"WENO5-JS": WENO5JS, "WENO5-Z": WENO5Z}[stencil]
    st = cls(nh=4, inactive_axes=[])
    for j in (0, 1):
        got = host(st.reconstruct_xi(dev(W), axis, j))
        ref = o.weno5_reconstruct_xi(W, dom, axis, j, stencil)
        assert rel(got, ref) < TOL, (j, rel(got, ref))


def test_weno5_reconstruct_1d_inactive_axes():
    from jaxfluids_amd import WENO5Z

    say("[weno5_reconstruct] 1-D buffer with inactive y, z")
    dom, W0 = o.sod_initial(64)
    s = o.Setup(bcs={"east": "ZEROGRADIENT", "west": "ZEROGRADIENT"})
    W, U = o.create_buffers(W0, dom, s)
    st = WENO5Z(nh=dom.nh, inactive_axes=["y", "z"])
    got = host(st.reconstruct_xi(dev(W), 0, 1))
    assert rel(got, o.weno5_reconstruct_xi(W, dom, 0, 1, "WENO5-Z")) < TOL


# ------------------------------------------------------------------ Riemann
@pytest.mark.parametrize("solver,speed", [("HLLC", s) for s in ("ARITHMETIC", "RUSANOV", "DAVIS", "EINFELDT", "TORO")]
                         + [("RUSANOV", "EINFELDT")])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_riemann_parity(solver, speed, axis):
    from jaxfluids_amd import DICT_RIEMANN_SOLVER, EquationManager, IdealGas, MaterialManager

    say(f"[riemann] {solver}/{speed} axis={axis}")
    rng = np.random.default_rng(5 + axis)
    n = 4096
    g = 1.4

    def prims():
        return np.stack([rng.uniform(0.1, 2, n), rng.uniform(-2, 2, n), rng.uniform(-1, 1, n),
                         rng.uniform(-1, 1, n), rng.uniform(0.05, 3, n)])

    WL, WR = prims(), prims()
    WR[:, :64] = WL[:, :64]            # equal states (S* == u, consistency)
    WR[axis + 1, 64:128] = WL[axis + 1, 64:128]
    WR[4, 64:128] = WL[4, 64:128]      # contacts
    UL, UR = o.cons_from_prim(WL, g), o.cons_from_prim(WR, g)
    ref = o.RIEMANN[solver](WL, WR, UL, UR, axis, g, speed)
    mm = MaterialManager(IdealGas(g))
    rs = DICT_RIEMANN_SOLVER[solver](mm, EquationManager(mm), speed)
    got, _, _ = rs.solve_riemann_problem_xi(dev(WL), dev(WR), dev(UL), dev(UR), axis)
    assert rel(host(got), ref) < TOL


# ------------------------------------------------------------------ compute_flux_xi / compute_rhs
FLUX_CASES = [
    ("WENO5-JS", "PRIMITIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
    ("WENO5-Z", "PRIMITIVE", "RUSANOV", "EINFELDT", "ARITHMETIC"),
    ("WENO5-Z", "CHAR-PRIMITIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
    ("WENO5-JS", "CHAR-PRIMITIVE", "HLLC", "TORO", "ROE"),
    ("WENO5-Z", "CONSERVATIVE", "HLLC", "DAVIS", "ARITHMETIC"),
    ("WENO5-JS", "CHAR-CONSERVATIVE", "HLLC", "ARITHMETIC", "ARITHMETIC"),
    ("WENO5-Z", "CHAR-CONSERVATIVE", "RUSANOV", "EINFELDT", "ROE"),
]


def _setup(case, bcs=None):
    st, var, rs, ss, fz = case
    return o.Setup(gamma=1.4, stencil=st, variable=var, riemann=rs, signal_speed=ss, frozen_state=fz,
                   bcs=bcs or {f: "PERIODIC" for f in FACES})


@pytest.mark.parametrize("case", FLUX_CASES, ids=lambda c: "-".join(c))
def test_compute_flux_xi_parity(case):
    from jaxfluids_amd import _native

    say(f"[compute_flux_xi] {case}")
    s = _setup(case)
    dom, W0 = random_state((10, 9, 8), 3, seed=21)
    W, U = o.create_buffers(W0, dom, s)
    lib = _native.load()
    Wd, Ud = dev(W), dev(U)
    for axis in range(3):
        ref = o.compute_flux_xi(W, U, dom, axis, s)
        out = torch.empty(ref.shape, dtype=torch.float64, device="cuda")
        rc = lib.jfa_compute_flux_xi(_native.ptr(Wd), _native.ptr(Ud), native_grid(dom), native_scheme(s), axis,
                                     _native.ptr(out), None, _native.stream_ptr())
        _native.check(rc, "flux")
        assert rel(host(out), ref) < TOL, (axis, rel(host(out), ref))


@pytest.mark.parametrize("case", FLUX_CASES[:3] + FLUX_CASES[5:6], ids=lambda c: "-".join(c))
def test_compute_rhs_parity(case):
    from jaxfluids_amd import _native

    say(f"[compute_rhs] {case}")
    s = _setup(case)
    dom, W0 = random_state((12, 10, 6), 5, seed=3)
    W, U = o.create_buffers(W0, dom, s)
    ref = o.compute_rhs(W, U, dom, s)
    lib = _native.load()
    grid = native_grid(dom)
    ws = torch.empty(lib.jfa_rhs_workspace_bytes(grid, 0), dtype=torch.uint8, device="cuda")
    out = torch.empty(ref.shape, dtype=torch.float64, device="cuda")
    Wd, Ud = dev(W), dev(U)
    rc = lib.jfa_compute_rhs(_native.ptr(Wd), _native.ptr(Ud), grid, native_scheme(s), 0.0, _native.ptr(out),
                             _native.ptr(ws), ws.numel(), _native.stream_ptr())
    _native.check(rc, "rhs")
    assert rel(host(out), ref) < TOL


# ------------------------------------------------------------------ halos, dt
@pytest.mark.parametrize("bc", ["PERIODIC", "ZEROGRADIENT", "SYMMETRY"])
def test_halo_and_dt_parity(bc):
    from jaxfluids_amd import _native

    say(f"[halo+dt] {bc}")
    s = o.Setup(bcs={f: bc for f in FACES}, cfl=0.7)
    dom, W0 = random_state((7, 6, 9), 4, seed=8, ranges=((0, 1.0), (0, 0.7), (0, 1.3)))
    W, U = o.create_buffers(W0, dom, s)
    # scramble the halos, then refill on the GPU
    Wd, Ud = dev(W), dev(U)
    I = dom.interior
    Wd_s = torch.full_like(Wd, 7.0)
    Ud_s = torch.full_like(Ud, 7.0)
    Wd_s[I] = Wd[I]
    Ud_s[I] = Ud[I]
    lib = _native.load()
    rc = lib.jfa_halo_update(_native.ptr(Wd_s), _native.ptr(Ud_s), native_grid(dom), _native.bc_array(s.bcs), 1.4,
                             0, _native.stream_ptr())
    _native.check(rc, "halo")
    Wh, Uh = host(Wd_s), host(Ud_s)
    for a in range(3):
        for side in (0, 1):
            sl = [dom.dom(0), dom.dom(1), dom.dom(2)]
            sl[a] = slice(0, dom.nh) if side == 0 else slice(-dom.nh, None)
            sl = (Ellipsis,) + tuple(sl)
            assert np.array_equal(Wh[sl], W[sl]), (a, side)
            assert rel(Uh[sl], U[sl]) < 1e-15, (a, side)
    dt = torch.empty(1, dtype=torch.float64, device="cuda")
    ws = torch.empty(32, dtype=torch.uint8, device="cuda")
    rc = lib.jfa_compute_dt(_native.ptr(Wd), native_grid(dom), 1.4, 0.7, o.EPS64, 0, None, _native.ptr(dt),
                            _native.ptr(ws), _native.stream_ptr())
    _native.check(rc, "dt")
    assert abs(host(dt)[0] - o.compute_dt(W, dom, s)) <= 1e-15 * o.compute_dt(W, dom, s)


# ------------------------------------------------------------------ whole steps
STEP_CASES = [
    (FLUX_CASES[0], "ZEROGRADIENT"),
    (FLUX_CASES[2], "PERIODIC"),
    (FLUX_CASES[3], "SYMMETRY"),
    (FLUX_CASES[5], "PERIODIC"),
]


@pytest.mark.parametrize("case,bc", STEP_CASES, ids=lambda c: "-".join(c) if isinstance(c, tuple) else c)
@pytest.mark.parametrize("fused", [True, False])
def test_rk3_step_parity(case, bc, fused):
    """One full RK3 step (3 stages, halos, next dt) through SimulationManager."""
    say(f"[rk3 step] fused={fused} {case} {bc}")
    from jaxfluids_amd import SimulationManager

    s = _setup(case, {f: bc for f in FACES})
    dom, W0 = random_state((11, 9, 8), 4, seed=17)
    im = _input_manager(dom, s)
    sim = SimulationManager(im, fused=fused)
    W, U = o.create_buffers(W0, dom, s)
    dt = o.compute_dt(W, dom, s)
    Wr, Ur = o.rk3_step(W, U, dt, dom, s)
    Wd, Ud = dev(W), dev(U)
    dtd = torch.full((1,), dt, dtype=torch.float64, device="cuda")
    Wd, Ud = sim.do_runge_kutta_stages(Wd, Ud, dtd)
    Wg, Ug = host(Wd), host(Ud)
    I = dom.interior
    assert rel(Wg[I], Wr[I]) < TOL and rel(Ug[I], Ur[I]) < TOL
    # face halos (edge halos are not part of the inviscid state)
    for a in range(3):
        sl = [dom.dom(0), dom.dom(1), dom.dom(2)]
        for part in (slice(0, dom.nh), slice(-dom.nh, None)):
            sl[a] = part
            t = (Ellipsis,) + tuple(sl)
            assert rel(Wg[t], Wr[t]) < TOL and rel(Ug[t], Ur[t]) < TOL
    dt2 = sim.compute_time_step(Wd)
    assert abs(host(dt2)[0] - o.compute_dt(Wr, dom, s)) < 1e-14 * dt


def _convective_fluxes_dict(s):
    if s.convective_solver == "FLUX-SPLITTING":
        return {"convective_solver": "FLUX-SPLITTING", "flux_splitting": {
            "flux_splitting": s.flux_splitting, "reconstruction_stencil": s.stencil, "frozen_state": s.frozen_state}}
    return {"convective_solver": "GODUNOV", "godunov": {
        "riemann_solver": s.riemann, "signal_speed": s.signal_speed or "EINFELDT",
        "reconstruction_stencil": s.stencil, "reconstruction_variable": s.variable, "frozen_state": s.frozen_state}}


def _input_manager(dom, s, end_time=None, end_step=None, positivity=None):
    from jaxfluids_amd import InputManager

    case = {
        "general": {"case_name": "t", "end_time": end_time if end_time is not None else 1e30,
                    "end_step": end_step if end_step is not None else 2 ** 40},
        "domain": {a: {"cells": dom.cells[i], "range": list(dom.ranges[i])} for i, a in enumerate("xyz")},
        "boundary_conditions": {f: {"type": s.bcs.get(f, "INACTIVE") if o.FACE_AXIS[f] in dom.active else "INACTIVE"}
                                for f in FACES},
        "initial_condition": {k: 1.0 for k in ("rho", "u", "v", "w", "p")},
        "material_properties": {"equation_of_state": {"model": "IdealGas", "specific_heat_ratio": s.gamma}},
    }
    num = {"conservatives": {"halo_cells": dom.nh,
                             "time_integration": {"integrator": "RK3", "CFL": s.cfl},
                             "convective_fluxes": _convective_fluxes_dict(s),
                             "positivity": positivity or {}},
           "active_physics": {"is_convective_flux": True},
           "precision": {"is_double_precision_compute": True}}
    return InputManager(case, num)


@pytest.mark.parametrize("variable", ["PRIMITIVE", "CHAR-PRIMITIVE"])
@pytest.mark.parametrize("use_graph", [False, True])
def test_sod_simulation_parity(variable, use_graph):
    """Config-1 Sod (1000 cells, WENO5-JS + HLLC + RK3, CFL 0.5, t_end 0.2) through the
    whole step loop: same number of steps, same final time, final state within 1e-11."""
    from jaxfluids_amd import InitializationManager, SimulationManager

    say(f"[sod] {variable} graph={use_graph}")
    dom, W0 = o.sod_initial(1000)
    s = o.Setup(stencil="WENO5-JS", variable=variable, riemann="HLLC", signal_speed="EINFELDT",
                bcs={"east": "ZEROGRADIENT", "west": "ZEROGRADIENT"})
    W, U = o.create_buffers(W0, dom, s)
    Wr, Ur, tr, nr = o.simulate(W, U, dom, s, end_time=0.2)
    im = _input_manager(dom, s, end_time=0.2)
    bufs = InitializationManager(im).initialization(W0)
    sim = SimulationManager(im, use_graph=use_graph)
    out = sim.simulate(bufs)
    tcv = out.time_control_variables
    assert tcv.simulation_step == nr
    assert abs(tcv.physical_simulation_time - tr) < 1e-12
    Wg = host(out.material_fields.primitives)
    I = dom.interior
    assert rel(Wg[I], Wr[I]) < 1e-11, rel(Wg[I], Wr[I])


def test_tgv_steps_parity():
    """Taylor-Green vortex 24^3 (WENO5-Z, CHAR-PRIMITIVE, HLLC), 5 steps with CFL dt."""
    from jaxfluids_amd import InitializationManager, SimulationManager

    say("[tgv] 24^3 x 5 steps")
    dom, W0 = o.tgv_initial(24, nh=5)
    s = o.Setup(stencil="WENO5-Z", variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT")
    W, U = o.create_buffers(W0, dom, s)
    Wr, Ur, tr, nr = o.simulate(W, U, dom, s, end_step=5)
    im = _input_manager(dom, s, end_step=5)
    bufs = InitializationManager(im).initialization(W0)
    out = SimulationManager(im).simulate(bufs)
    I = dom.interior
    assert rel(host(out.material_fields.primitives)[I], Wr[I]) < TOL


@pytest.mark.parametrize("case,bc", [(FLUX_CASES[2], "PERIODIC"), (FLUX_CASES[0], "ZEROGRADIENT"),
                                     (FLUX_CASES[6], "SYMMETRY")],
                         ids=["z-charprim-periodic", "js-prim-zerograd", "z-charcons-rusanov-sym"])
def test_rk3_step_parity_2d(case, bc):
    """2-D domain (z inactive): the innermost active axis is y, x is marched."""
    from jaxfluids_amd import SimulationManager

    say(f"[rk3 step 2d] {case} {bc}")
    s = _setup(case, {f: bc for f in FACES})
    dom, W0 = random_state((37, 70, 1), 3, seed=29, ranges=((0, 1.0), (0, 1.3), (0, 1.0)))
    W0 = W0[:, :, :, :1]
    dom = o.Domain((37, 70, 1), ((0, 1.0), (0, 1.3), (0, 1.0)), nh=3)
    s.bcs = {f: (bc if o.FACE_AXIS[f] in dom.active else "PERIODIC") for f in FACES}
    im = _input_manager(dom, s)
    sim = SimulationManager(im)
    W, U = o.create_buffers(W0, dom, s)
    dt = o.compute_dt(W, dom, s)
    Wr, Ur = o.rk3_step(W, U, dt, dom, s)
    Wd, Ud = dev(W), dev(U)
    dtd = torch.full((1,), dt, dtype=torch.float64, device="cuda")
    Wd, Ud = sim.do_runge_kutta_stages(Wd, Ud, dtd, dt_next=dtd)
    I = dom.interior
    assert rel(host(Wd)[I], Wr[I]) < TOL and rel(host(Ud)[I], Ur[I]) < TOL
    assert abs(host(dtd)[0] - o.compute_dt(Wr, dom, s)) < 1e-14 * dt


def test_fp32_storage_error_stated():
    """FP32 storage (FP64 arithmetic) against the FP64 oracle on one TGV step: the error is
    the float32 rounding of the state, stated here as < 1e-6 relative L-inf."""
    from jaxfluids_amd import SimulationManager

    say("[fp32 storage] tgv 24^3 one step")
    dom, W0 = o.tgv_initial(24, nh=5)
    s = o.Setup(stencil="WENO5-Z", variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT")
    W, U = o.create_buffers(W0, dom, s)
    dt = o.compute_dt(W, dom, s)
    Wr, Ur = o.rk3_step(W, U, dt, dom, s)
    im = _input_manager(dom, s)
    sim = SimulationManager(im)
    Wd, Ud = dev(W, torch.float32), dev(U, torch.float32)
    dtd = torch.full((1,), dt, dtype=torch.float64, device="cuda")
    Wd, Ud = sim.do_runge_kutta_stages(Wd, Ud, dtd)
    I = dom.interior
    err = rel(host(Wd)[I], Wr[I])
    say(f"    fp32-storage rel Linf vs FP64 oracle: {err:.3e}")
    assert err < 1e-6


@pytest.mark.parametrize("fused", [True, False])
def test_cfl_time_step_propagates_nan(fused):
    """A NaN cell makes dt NaN, as the reference's jnp.max does (time_step_size.py:109): the
    standalone reduction (jfa_compute_dt) and the fused stage's in-kernel partials; the
    reference's time loop then ends (NaN < end_time is false), and so does advance()."""
    from jaxfluids_amd import SimulationManager
    from jaxfluids_amd.data_types import JaxFluidsBuffers, MaterialFieldBuffers, TimeControlVariables

    say(f"[nan dt] fused={fused}")
    s = _setup(FLUX_CASES[0], {f: "PERIODIC" for f in FACES})
    dom, W0 = random_state((11, 9, 8), 4, seed=23)
    W, U = o.create_buffers(W0, dom, s)
    dt = o.compute_dt(W, dom, s)
    c = (0, dom.nh + 4, dom.nh + 3, dom.nh + 2)
    W[c] = np.nan
    U[c] = np.nan
    assert np.isnan(o.compute_dt(W, dom, s))
    sim = SimulationManager(_input_manager(dom, s, end_step=5), fused=fused)
    Wd, Ud = dev(W), dev(U)
    assert math.isnan(float(host(sim.compute_time_step(Wd))[0]))
    dtd = torch.full((1,), dt, dtype=torch.float64, device="cuda")
    tcv = TimeControlVariables(physical_simulation_time=0.0, simulation_step=0, physical_timestep_size=dtd,
                               end_time=1e30, end_step=5)
    out = sim.advance(JaxFluidsBuffers(MaterialFieldBuffers(Ud, Wd, None), tcv))
    t = out.time_control_variables
    assert t.simulation_step == 2 and math.isnan(t.physical_simulation_time), t  # step 1 ran with a finite dt
    assert math.isnan(float(host(t.physical_timestep_size)[0]))
