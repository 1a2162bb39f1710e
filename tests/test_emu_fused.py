"""CPU checks of the fused RK3 stage (jfa_fused.h: the MARCH and INNER kernels) through the
host build of the library sources (tests/emu): one RK3 step of three jfa_rk3_stage_fused calls on
host buffers, with the next CFL time step, against the oracle's step (FP64, relative L-inf
<= 1e-12).

Every state, workspace and time-step load and store of the kernels is checked against the
registered buffers, so an out-of-bounds index shows up as a violation here instead of a fault on
the GPU.  The grids cover every kernel combination: 3-D (MARCH x and y, INNER z), the three
2-D orientations (one MARCH, INNER) and 1-D (INNER alone), with odd extents so the 63-face
wavefront tiling of the INNER kernel crosses rows.
"""
import ctypes
import dataclasses
import os
import sys

import numpy as np
import pytest

from oracle import jxf_oracle as o
from test_gpu_parity import random_state, rel
from test_oracle_pins import _pocket_state

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "emu"))

TOL = 1e-12
FACES = ("east", "west", "north", "south", "top", "bottom")
SCHEMES = {
    "z-charprim-hllc": dict(stencil="WENO5-Z", variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                            frozen_state="ROE"),
    "js-prim-rusanov": dict(stencil="WENO5-JS", variable="PRIMITIVE", riemann="RUSANOV", signal_speed="EINFELDT",
                            frozen_state="ARITHMETIC"),
    "z-charcons-hllc": dict(stencil="WENO5-Z", variable="CHAR-CONSERVATIVE", riemann="HLLC", signal_speed="DAVIS",
                            frozen_state="ARITHMETIC"),
    "js-cons-hllc": dict(stencil="WENO5-JS", variable="CONSERVATIVE", riemann="HLLC", signal_speed="TORO",
                         frozen_state="ROE"),
}


@pytest.fixture(scope="module")
def emu():
    import build as emu_build
    from jaxfluids_amd import _native

    lib = ctypes.CDLL(emu_build.build())
    for name in ("jfa_stage_workspace_bytes", "jfa_rk3_stage_fused", "jfa_edge_halo_update"):
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = _native.SIGNATURES[name]
    lib.emu_add_range.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    lib.emu_violations.restype = ctypes.c_longlong
    lib.emu_accesses.restype = ctypes.c_longlong
    lib.emu_first_violation.restype = ctypes.c_ulonglong
    return lib


def fused_step(emu, dom, s, W, U, dt, transport=None, dtype=np.float64):
    """Three fused stages as SimulationManager.do_runge_kutta_stages runs them (the last one
    writes into the step's input buffers and the next dt); returns (W, U, dt_next).  dtype:
    the storage of W and U (float32: the FP32-storage instantiation, FP64 arithmetic)."""
    from jaxfluids_amd import _native

    grid = _native.make_grid(dom.cells, dom.nh, dom.dx)
    tr = None if transport is None else _native.make_transport(**dataclasses.asdict(transport))
    code = _native.F64 if dtype == np.float64 else _native.F32
    scheme = _native.make_scheme(s.variable, s.stencil, s.riemann, s.signal_speed, s.frozen_state, code, s.gamma,
                                 s.eps_stencil, tr, s.interpolation_limiter, s.limit_velocity, s.limiter_eps,
                                 s.flux_limiter, s.flux_partition, s.flux_limiter_eps)
    bcs = _native.bc_array({f: (s.bcs[f] if o.FACE_AXIS[f] in dom.active else "PERIODIC") for f in FACES})
    nbytes = emu.jfa_stage_workspace_bytes(grid, scheme)
    assert nbytes > 0
    Wa, Ua = np.ascontiguousarray(W.astype(dtype)), np.ascontiguousarray(U.astype(dtype))
    Wb, Ub, Wc, Uc = (np.full_like(Wa, np.nan) for _ in range(4))
    ws = np.zeros(nbytes, dtype=np.uint8)
    dta = np.array([dt])
    dtn = np.array([np.nan])
    emu.emu_clear_ranges()
    for a in (Wa, Ua, Wb, Ub, Wc, Uc, ws, dta, dtn):
        emu.emu_add_range(a.ctypes.data, a.nbytes)
    P = lambda a: None if a is None else a.ctypes.data  # noqa: E731
    for k, (wi, ui, wo, uo) in enumerate([(Wa, Ua, Wb, Ub), (Wb, Ub, Wc, Uc), (Wc, Uc, Wa, Ua)]):
        rc = emu.jfa_rk3_stage_fused(P(wi), P(ui), P(Ua if k else None), P(wo), P(uo), grid, scheme, bcs, k, P(dta),
                                     0.0, s.cfl, float(np.finfo(dtype).eps), P(dtn if k == 2 else None), P(ws),
                                     nbytes, None)
        assert rc == 0, rc
    assert emu.emu_violations() == 0, hex(emu.emu_first_violation())
    assert emu.emu_accesses() > 0
    return Wa, Ua, float(dtn[0])


def check_step(emu, dom, s, W0, transport=None, dt_scale=1.0):
    W, U = o.create_buffers(W0, dom, s)
    dt = dt_scale * o.compute_dt(W, dom, s)
    Wr, Ur = o.rk3_step(W, U, dt, dom, s)
    I = dom.interior
    assert np.all(np.isfinite(Wr[I])) and np.all(Wr[4][I] > 0)
    Wg, Ug, dtn = fused_step(emu, dom, s, W, U, dt, transport)
    assert rel(Wg[I], Wr[I]) < TOL and rel(Ug[I], Ur[I]) < TOL, (rel(Wg[I], Wr[I]), rel(Ug[I], Ur[I]))
    # face halos written by the last stage; with dissipative fluxes on a non-periodic domain the
    # edge halos too (periodic ones wrap onto the interior and leave the edge halos unwritten)
    periodic = all(s.bcs[f] == "PERIODIC" for f in FACES if o.FACE_AXIS[f] in dom.active)
    for t in halo_regions(dom, edges=transport is not None and not periodic):
        assert rel(Wg[t], Wr[t]) < TOL and rel(Ug[t], Ur[t]) < TOL, t
    assert abs(dtn - o.compute_dt(Wr, dom, s)) <= 1e-14 * dt


def halo_regions(dom, edges=False):
    """Index tuples of the face-halo slabs (interior along the other axes) and, with edges, of the
    edge-halo blocks of every pair of active axes (interior along the third)."""
    parts = (slice(0, dom.nh), slice(-dom.nh, None))
    out = []
    for a in dom.active:
        for part in parts:
            sl = [dom.dom(0), dom.dom(1), dom.dom(2)]
            sl[a] = part
            out.append((Ellipsis,) + tuple(sl))
    if edges:
        for i, a in enumerate(dom.active):
            for b in dom.active[i + 1:]:
                for pa in parts:
                    for pb in parts:
                        sl = [dom.dom(0), dom.dom(1), dom.dom(2)]
                        sl[a], sl[b] = pa, pb
                        out.append((Ellipsis,) + tuple(sl))
    return out


GRIDS = {
    "3d": ((9, 8, 7), ((0, 1.0), (0, 0.8), (0, 1.1))),
    "3d-wide": ((20, 6, 30), ((0, 1.0), (0, 0.5), (0, 1.4))),  # 60 wavefronts of z faces, 15 blocks
    "2d-xy": ((13, 11, 1), ((0, 1.0), (0, 1.3), (0, 1.0))),
    "2d-xz": ((10, 1, 12), ((0, 1.0), (0, 1.0), (0, 0.9))),
    "2d-yz": ((1, 9, 11), ((0, 1.0), (0, 1.2), (0, 1.0))),
    "1d": ((17, 1, 1), ((0, 1.0), (0, 1.0), (0, 1.0))),
}


def grid_state(name, seed, nh=3):
    cells, ranges = GRIDS[name]
    full = tuple(max(c, 2) for c in cells)  # random_state needs a mesh on every axis
    _, W0 = random_state(full, nh, seed=seed, ranges=ranges)
    W0 = np.ascontiguousarray(W0[:, :cells[0], :cells[1], :cells[2]])
    return o.Domain(cells, ranges, nh=nh), W0


@pytest.mark.parametrize("grid", list(GRIDS))
@pytest.mark.parametrize("bc", ["PERIODIC", "ZEROGRADIENT", "SYMMETRY"])
def test_fused_step_grids_and_boundaries(emu, grid, bc):
    dom, W0 = grid_state(grid, seed=3)
    s = o.Setup(gamma=1.4, bcs={f: bc for f in FACES}, cfl=0.6, **SCHEMES["z-charprim-hllc"])
    check_step(emu, dom, s, W0)


@pytest.mark.parametrize("scheme", list(SCHEMES)[1:])
@pytest.mark.parametrize("grid", ["3d", "2d-xy"])
def test_fused_step_schemes(emu, scheme, grid):
    dom, W0 = grid_state(grid, seed=7)
    s = o.Setup(gamma=1.4, bcs={f: "ZEROGRADIENT" for f in FACES}, cfl=0.5, **SCHEMES[scheme])
    check_step(emu, dom, s, W0)


@pytest.mark.parametrize("grid", ["3d", "2d-yz"])
def test_fused_step_viscous(emu, grid):
    """Dissipative fluxes through the MARCH and INNER kernels (periodic)."""
    dom, W0 = grid_state(grid, seed=5)
    tr = o.Transport(R=1.0, mu_ref=2e-2, T_ref=1.0, mu_exp=0.76, bulk_viscosity=1e-2, prandtl=0.71)
    s = o.Setup(gamma=1.4, bcs={f: "PERIODIC" for f in FACES}, cfl=0.5, transport=tr,
                **SCHEMES["z-charprim-hllc"])
    check_step(emu, dom, s, W0, transport=tr)


VISCOUS_BCS = {
    "zerogradient": {f: "ZEROGRADIENT" for f in FACES},
    "symmetry": {f: "SYMMETRY" for f in FACES},
    # periodic x with walls and outflow: PERIODIC decides the x edges, SYMMETRY the y-z edges
    "mixed": {"east": "PERIODIC", "west": "PERIODIC", "north": "SYMMETRY", "south": "ZEROGRADIENT",
              "top": "ZEROGRADIENT", "bottom": "SYMMETRY"},
    # SYMMETRY on the upper axis only: it decides the edges where the lower axis' face is ANY
    "upper-symmetry": {"east": "ZEROGRADIENT", "west": "ZEROGRADIENT", "north": "SYMMETRY", "south": "SYMMETRY",
                       "top": "SYMMETRY", "bottom": "ZEROGRADIENT"},
}


@pytest.mark.parametrize("bcs", list(VISCOUS_BCS))
@pytest.mark.parametrize("grid", ["3d", "2d-xy", "2d-xz", "1d"])
def test_fused_step_viscous_boundaries(emu, grid, bcs):
    """Dissipative fluxes on non-periodic domains: the stencils read the face halos and the edge
    halos, which every stage refills (the intermediate buffers start as NaN, so a halo read
    before it is written shows)."""
    dom, W0 = grid_state(grid, seed=9)
    tr = o.Transport(R=1.0, mu_ref=2e-2, T_ref=1.0, mu_exp=0.76, bulk_viscosity=1e-2, prandtl=0.71)
    s = o.Setup(gamma=1.4, bcs=VISCOUS_BCS[bcs], cfl=0.5, transport=tr, **SCHEMES["z-charprim-hllc"])
    check_step(emu, dom, s, W0, transport=tr)


@pytest.mark.parametrize("grid", ["3d", "2d-xz"])
def test_fused_step_viscous_boundaries_fp32_storage(emu, grid):
    """The FP32-storage instantiation with dissipative fluxes on a mixed non-periodic domain (the
    float edge-halo fill included) against the FP64 oracle: the float32 rounding of the state."""
    dom, W0 = grid_state(grid, seed=21)
    tr = o.Transport(R=1.0, mu_ref=2e-2, T_ref=1.0, mu_exp=0.76, bulk_viscosity=1e-2, prandtl=0.71)
    s = o.Setup(gamma=1.4, bcs=VISCOUS_BCS["mixed"], cfl=0.5, transport=tr, **SCHEMES["z-charprim-hllc"])
    W, U = o.create_buffers(W0, dom, s)
    dt = o.compute_dt(W, dom, s)
    Wr, _ = o.rk3_step(W, U, dt, dom, s)
    Wg, _, _ = fused_step(emu, dom, s, W, U, dt, transport=tr, dtype=np.float32)
    for t in [dom.interior] + halo_regions(dom, edges=True):
        err = rel(Wg[t].astype(np.float64), Wr[t])
        assert err < 1e-6, (t, err)


@pytest.mark.parametrize("bcs", list(VISCOUS_BCS))
@pytest.mark.parametrize("grid", ["3d", "2d-yz"])
def test_edge_halo_update(emu, grid, bcs):
    """jfa_edge_halo_update on a buffer with random face halos against the oracle's edge fill:
    W bit-exact (copies, sign flips and one mean), U = cons(W) to an ulp of the energy; vertex
    halos stay as they were."""
    dom, W0 = grid_state(grid, seed=13)
    s = o.Setup(gamma=1.4, bcs=VISCOUS_BCS[bcs])
    rng = np.random.default_rng(3)
    W = np.ascontiguousarray(1.0 + rng.random((5,) + dom.shape))
    U = np.ascontiguousarray(rng.random((5,) + dom.shape))
    Wr, Ur = o.edge_halo_update(W, U, dom, s)
    from jaxfluids_amd import _native

    grid_ = _native.make_grid(dom.cells, dom.nh, dom.dx)
    bcs_ = _native.bc_array({f: (s.bcs[f] if o.FACE_AXIS[f] in dom.active else "PERIODIC") for f in FACES})
    emu.emu_clear_ranges()
    for a in (W, U):
        emu.emu_add_range(a.ctypes.data, a.nbytes)
    assert emu.jfa_edge_halo_update(W.ctypes.data, U.ctypes.data, grid_, bcs_, 1.4, 0, None) == 0
    assert emu.emu_violations() == 0, hex(emu.emu_first_violation())
    assert np.array_equal(W, Wr)
    assert np.array_equal(U[:4], Ur[:4]) and rel(U, Ur) <= 4e-16, rel(U, Ur)


# HAS's blend weight theta = (eps - q_LF) / (q - q_LF) is ill-conditioned where the first-order
# flux's pseudo-updated state is itself inadmissible (on the pocket at CFL 0.9 |theta| reaches 1e14,
# so the result hangs on the last bit of every operand): HAS is compared at CFL 0.6, where every
# theta lies in [0, 1].
HAS_CFL = {"HAS": 0.6}


@pytest.mark.parametrize("fl,partition", [("SIMPLE", "CELLSIZE"), ("NASA", "CELLSIZE"), ("HAS", "CELLSIZE"),
                                          ("HAS", "WAVESPEED"), ("NASA", "WAVESPEED")])
def test_fused_step_with_limiters(emu, fl, partition):
    """Both positivity limiters on the near-vacuum pocket (2-D: MARCH x, INNER y)."""
    dom, W0 = _pocket_state()
    zg = {f: "ZEROGRADIENT" for f in ("east", "west", "north", "south")}
    s = o.Setup(stencil="WENO5-JS", variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT", bcs=zg,
                interpolation_limiter=True, flux_limiter=fl, flux_partition=partition)
    W, U = o.create_buffers(W0, dom, s)
    dt = HAS_CFL.get(fl, 0.9) / 0.5 * o.compute_dt(W, dom, s)
    Wr, Ur = o.rk3_step(W, U, dt, dom, s)
    Wg, Ug, _ = fused_step(emu, dom, s, W, U, dt)
    I = dom.interior
    assert rel(Wg[I], Wr[I]) < TOL and rel(Ug[I], Ur[I]) < TOL


NEW_STENCILS = ["WENO1", "WENO3-JS", "WENO3-Z", "TENO5", "TENO5-A", "WENO6-CU", "TENO6", "WENO6-CUM1", "WENO6-CUM2", "TENO6-A"]


@pytest.mark.parametrize("stencil", NEW_STENCILS)
@pytest.mark.parametrize("variable,grid", [("PRIMITIVE", "3d"), ("CHAR-PRIMITIVE", "2d-xy"),
                                           ("CHAR-CONSERVATIVE", "2d-yz"), ("CONSERVATIVE", "1d")])
def test_fused_step_further_stencils(emu, stencil, variable, grid):
    """WENO1, WENO3-JS/-Z, TENO5, TENO5-A, WENO6-CU and TENO6 (the run-time stencil switch of the
    kernels)."""
    dom, W0 = grid_state(grid, seed=11)
    s = o.Setup(gamma=1.4, bcs={f: "ZEROGRADIENT" for f in FACES}, cfl=0.5, stencil=stencil, variable=variable,
                riemann="HLLC", signal_speed="EINFELDT", frozen_state="ROE")
    check_step(emu, dom, s, W0)


@pytest.mark.parametrize("variable,grid,bc", [("PRIMITIVE", "3d", "ZEROGRADIENT"), ("CHAR-PRIMITIVE", "2d-xy", "PERIODIC"),
                                              ("CHAR-CONSERVATIVE", "2d-yz", "SYMMETRY"),
                                              ("CONSERVATIVE", "1d", "ZEROGRADIENT"),
                                              ("CHAR-PRIMITIVE", "3d-wide", "PERIODIC")])
@pytest.mark.parametrize("riemann", ["HLLC", "RUSANOV", "HLL"])
def test_fused_step_weno7(emu, variable, grid, bc, riemann):
    """WENO7-JS (its own 8-cell-window instantiation of the march and inner kernels), nh = 4."""
    dom, W0 = grid_state(grid, seed=17, nh=4)
    s = o.Setup(gamma=1.4, bcs={f: bc for f in FACES}, cfl=0.5, stencil="WENO7-JS", variable=variable,
                riemann=riemann, signal_speed="EINFELDT", frozen_state="ROE")
    check_step(emu, dom, s, W0)


def test_weno7_refusals(emu):
    """WENO7-JS needs 4 halo cells (JFA_EINVAL with 3) and is not built for flux splitting
    (JFA_EUNSUPPORTED)."""
    from jaxfluids_amd import _native

    for nh, riemann, want in ((3, "HLLC", _native.JFA_EINVAL), (4, "FLUX-SPLITTING-ROE", _native.JFA_EUNSUPPORTED)):
        dom, W0 = grid_state("2d-xy", seed=1, nh=nh)
        s = o.Setup(gamma=1.4, bcs={f: "PERIODIC" for f in FACES}, stencil="WENO7-JS", variable="PRIMITIVE",
                    riemann=riemann, signal_speed="EINFELDT")
        grid = _native.make_grid(dom.cells, dom.nh, dom.dx)
        scheme = _native.make_scheme(s.variable, s.stencil, s.riemann, s.signal_speed, s.frozen_state, 0, s.gamma,
                                     s.eps_stencil)
        W = np.zeros((5,) + dom.shape)
        ws = np.zeros(emu.jfa_stage_workspace_bytes(grid, scheme), dtype=np.uint8)
        bcs = _native.bc_array({f: "PERIODIC" for f in FACES})
        P = W.ctypes.data
        rc = emu.jfa_rk3_stage_fused(P, P, None, ws.ctypes.data, ws.ctypes.data + 8, grid, scheme, bcs, 0, None, 1e-3,
                                     0.5, o.EPS64, None, ws.ctypes.data, ws.nbytes, None)
        assert rc == want, (nh, riemann, rc)


@pytest.mark.parametrize("riemann", ["HLL", "HLLC-LM", "AUSMP"])
@pytest.mark.parametrize("speed", ["EINFELDT", "DAVIS", "TORO"])
@pytest.mark.parametrize("variable,grid", [("CHAR-PRIMITIVE", "3d"), ("PRIMITIVE", "2d-xz")])
def test_fused_step_further_riemann_solvers(emu, riemann, speed, variable, grid):
    """HLL (HLL.py:29-61), HLLC-LM (HLLCLM.py:59-136) and AUSM+ (AUSMP.py:30-83) in the fused
    sweeps; the state's Mach numbers straddle HLLC-LM's limit of 0.1."""
    dom, W0 = grid_state(grid, seed=13)
    s = o.Setup(gamma=1.4, bcs={f: "ZEROGRADIENT" for f in FACES}, cfl=0.5, stencil="WENO5-Z", variable=variable,
                riemann=riemann, signal_speed=speed, frozen_state="ROE")
    check_step(emu, dom, s, W0)


@pytest.mark.parametrize("fs", ["ROE", "CLLF", "LLF"])
@pytest.mark.parametrize("stencil,grid", [("WENO5-Z", "3d"), ("WENO5-JS", "2d-xy"), ("TENO6", "2d-yz"),
                                          ("WENO3-JS", "1d"), ("WENO6-CUM2", "3d-wide")])
def test_fused_step_flux_splitting(emu, fs, stencil, grid):
    """The flux-splitting scheme (flux_splitting_scheme.py:65-111) as the face flux of the fused
    sweeps (the kFluxSplit instantiation of the PRIMITIVE-window kernels)."""
    dom, W0 = grid_state(grid, seed=13)
    s = o.Setup(gamma=1.4, bcs={f: "ZEROGRADIENT" for f in FACES}, cfl=0.5, stencil=stencil,
                riemann="FLUX-SPLITTING-" + fs, frozen_state="ROE", variable="CHAR-CONSERVATIVE")
    check_step(emu, dom, s, W0)


@pytest.mark.parametrize("fl", ["SIMPLE", "NASA"])
def test_fused_step_flux_splitting_with_flux_limiter(emu, fl):
    dom, W0 = _pocket_state()
    zg = {f: "ZEROGRADIENT" for f in ("east", "west", "north", "south")}
    s = o.Setup(stencil="WENO5-Z", riemann="FLUX-SPLITTING-LLF", frozen_state="ARITHMETIC", bcs=zg,
                flux_limiter=fl)
    W, U = o.create_buffers(W0, dom, s)
    dt = 0.9 / 0.5 * o.compute_dt(W, dom, s)
    Wr, Ur = o.rk3_step(W, U, dt, dom, s)
    Wg, Ug, _ = fused_step(emu, dom, s, W, U, dt)
    I = dom.interior
    assert rel(Wg[I], Wr[I]) < TOL and rel(Ug[I], Ur[I]) < TOL


@pytest.mark.parametrize("fs", ["ROE", "LLF"])
@pytest.mark.parametrize("grid", ["3d", "2d-xz"])
def test_fused_step_flux_splitting_fp32_storage(emu, fs, grid):
    """The FP32-storage instantiation of the flux-splitting sweeps (float W and U, FP64
    arithmetic) against the FP64 oracle: the error is the float32 rounding of the state, stated
    here as < 1e-6 relative L-inf."""
    dom, W0 = grid_state(grid, seed=17)
    s = o.Setup(gamma=1.4, stencil="WENO5-Z", variable="PRIMITIVE", riemann="FLUX-SPLITTING-" + fs,
                signal_speed="EINFELDT", frozen_state="ROE", bcs={f: "ZEROGRADIENT" for f in FACES}, cfl=0.5)
    W, U = o.create_buffers(W0, dom, s)
    dt = o.compute_dt(W, dom, s)
    Wr, Ur = o.rk3_step(W, U, dt, dom, s)
    Wg, Ug, dtn = fused_step(emu, dom, s, W, U, dt, dtype=np.float32)
    I = dom.interior
    err = rel(Wg[I].astype(np.float64), Wr[I])
    assert 1e-9 < err < 1e-6, err  # float32 rounding, not FP64 agreement
    assert rel(Ug[I].astype(np.float64), Ur[I]) < 1e-6
    assert abs(dtn - o.compute_dt(Wr, dom, s)) <= 1e-6 * dt


@pytest.mark.parametrize("grid", ["3d", "2d-xy", "1d"])
def test_fused_cfl_time_step_propagates_nan(emu, grid):
    """A NaN cell makes the next dt NaN, as the reference's jnp.max (time_step_size.py:109) does:
    the wave-speed maxima (wave shuffles, block partials and the final reduction) keep a NaN
    instead of dropping it."""
    dom, W0 = grid_state(grid, seed=31)
    s = o.Setup(gamma=1.4, **SCHEMES["z-charprim-hllc"], bcs={f: "PERIODIC" for f in FACES})
    W, U = o.create_buffers(W0, dom, s)
    dt = o.compute_dt(W, dom, s)
    c = tuple(n // 2 + dom.nh if n > 1 else 0 for n in dom.cells)
    W[(0,) + c] = np.nan
    U[(0,) + c] = np.nan
    assert np.isnan(o.compute_dt(W, dom, s))  # the oracle's np.max propagates NaN
    _, _, dtn = fused_step(emu, dom, s, W, U, dt)
    assert np.isnan(dtn)
