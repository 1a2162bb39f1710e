"""GPU parity of the further reconstruction stencils -- WENO1, WENO3-JS, WENO3-Z, TENO5, TENO5-A,
WENO6-CU, TENO6, WENO6-CUM1, WENO6-CUM2, TENO6-A and WENO7-JS
(stencils/reconstruction/shock_capturing/weno/weno1_js.py, weno3_js.py, weno3_z.py, weno6_cu.py,
weno6_cum1.py, weno6_cum2.py, weno7_js.py, teno/teno5.py, teno5_a.py, teno6.py, teno6_a.py) -- against the oracle: the plugin's reconstruct_xi, compute_flux_xi for
every reconstruction variable, RK3 steps of the fused stage through the setup dicts (3-D and 2-D),
and the batched ensemble stage.  FP64, relative L-inf <= 1e-12.

The kernels run these stencils through one run-time switch (the kStencilExt instantiation) on the
same 6-cell face window as WENO5, so every case is also a check that the switch picks the right
stencil and sub-window; WENO7-JS has its own 8-cell-window instantiation (nh = 4).
"""
import numpy as np
import pytest
import torch

from oracle import jxf_oracle as o
from test_gpu_parity import _input_manager, dev, host, native_grid, native_scheme, random_state, rel, say

pytestmark = pytest.mark.gpu

TOL = 1e-12
FACES = ("east", "west", "north", "south", "top", "bottom")
STENCILS = ["WENO1", "WENO3-JS", "WENO3-Z", "TENO5", "TENO5-A", "WENO6-CU", "TENO6", "WENO6-CUM1", "WENO6-CUM2", "TENO6-A",
            "WENO7-JS"]
NH = {"WENO7-JS": 4}  # WENO7-JS reads the 8-cell window (weno7_js.py:29-31)


@pytest.mark.parametrize("stencil", STENCILS)
def test_reconstruct_xi_plugin(stencil):
    from jaxfluids_amd.stencils import DICT_SPATIAL_RECONSTRUCTION

    say(f"[reconstruct_xi] {stencil}")
    nh = NH.get(stencil, 3)
    dom, W0 = random_state((10, 9, 8), nh, seed=31)
    s = o.Setup(stencil="WENO5-JS")
    W, _ = o.create_buffers(W0, dom, s)
    plugin = DICT_SPATIAL_RECONSTRUCTION[stencil](nh=nh, inactive_axes=())
    Wd = dev(W)
    for axis in range(3):
        for j in (0, 1):
            ref = o.weno5_reconstruct_xi(W, dom, axis, j, stencil)
            out = plugin.reconstruct_xi(Wd, axis, j, dx=dom.dx[axis])
            assert rel(host(out), ref) < TOL, (axis, j, rel(host(out), ref))


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("variable", ["PRIMITIVE", "CONSERVATIVE", "CHAR-PRIMITIVE", "CHAR-CONSERVATIVE"])
def test_compute_flux_xi(stencil, variable):
    from jaxfluids_amd import _native

    say(f"[flux_xi] {stencil} {variable}")
    s = o.Setup(gamma=1.4, stencil=stencil, variable=variable, riemann="HLLC", signal_speed="EINFELDT",
                frozen_state="ROE", bcs={f: "PERIODIC" for f in FACES})
    dom, W0 = random_state((10, 9, 8), NH.get(stencil, 3), seed=23)
    W, U = o.create_buffers(W0, dom, s)
    lib = _native.load()
    Wd, Ud = dev(W), dev(U)
    for axis in range(3):
        with np.errstate(invalid="ignore"):
            ref = o.compute_flux_xi(W, U, dom, axis, s)
        out = torch.empty(ref.shape, dtype=torch.float64, device="cuda")
        rc = lib.jfa_compute_flux_xi(_native.ptr(Wd), _native.ptr(Ud), native_grid(dom), native_scheme(s), axis,
                                     _native.ptr(out), None, _native.stream_ptr())
        _native.check(rc, "jfa_compute_flux_xi")
        got = host(out)
        # The periodic wrap of the jump along y gives TENO5-A (the least dissipative here) a few
        # faces whose reconstructed pressure is negative: the reference's flux is NaN there too.
        bad = ~np.isfinite(ref)
        assert np.array_equal(bad, ~np.isfinite(got)), axis
        assert np.sum(bad) <= 10 * 5, np.sum(bad)
        assert rel(np.where(bad, 0.0, got), np.where(bad, 0.0, ref)) < TOL, (axis, rel(got, ref))


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("cells,bc", [((11, 9, 8), "ZEROGRADIENT"), ((37, 70, 1), "PERIODIC")], ids=["3d", "2d"])
def test_rk3_step_through_setup_dicts(stencil, cells, bc):
    """One fused RK3 step and the next CFL dt of SimulationManager built from setup dicts that
    name the stencil (CHAR-PRIMITIVE, HLLC-Einfeldt)."""
    from jaxfluids_amd import SimulationManager

    say(f"[rk3 step] {stencil} {cells} {bc}")
    ranges = ((0, 1.0), (0, 1.3), (0, 1.0))
    full = tuple(max(c, 2) for c in cells)
    _, W0 = random_state(full, 3, seed=19, ranges=ranges)
    W0 = np.ascontiguousarray(W0[:, :cells[0], :cells[1], :cells[2]])
    dom = o.Domain(cells, ranges, nh=NH.get(stencil, 3))
    s = o.Setup(gamma=1.4, stencil=stencil, variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                frozen_state="ARITHMETIC",
                bcs={f: (bc if o.FACE_AXIS[f] in dom.active else "PERIODIC") for f in FACES})
    sim = SimulationManager(_input_manager(dom, s))
    W, U = o.create_buffers(W0, dom, s)
    dt = o.compute_dt(W, dom, s)
    Wr, Ur = o.rk3_step(W, U, dt, dom, s)
    I = dom.interior
    assert np.all(np.isfinite(Wr[I]))
    Wd, Ud = dev(W), dev(U)
    dtd = torch.full((1,), dt, dtype=torch.float64, device="cuda")
    Wd, Ud = sim.do_runge_kutta_stages(Wd, Ud, dtd, dt_next=dtd)
    assert rel(host(Wd)[I], Wr[I]) < TOL and rel(host(Ud)[I], Ur[I]) < TOL
    assert abs(host(dtd)[0] - o.compute_dt(Wr, dom, s)) < 1e-14 * dt


@pytest.mark.parametrize("scheme", [("TENO5", "CHAR-PRIMITIVE", "HLLC", "EINFELDT", "ROE"),
                                    ("WENO3-Z", "PRIMITIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
                                    ("WENO1", "CONSERVATIVE", "RUSANOV", "EINFELDT", "ARITHMETIC")],
                         ids=lambda s: "-".join(s))
def test_ensemble_step_matches_oracle(scheme):
    """The batched stage of 5 uneven cases, one RK3 step each with its own CFL dt."""
    from test_gpu_ensemble import CASES, case_state, face_regions, make_ensemble, oracle_case

    say(f"[ensemble] {scheme}")
    prims = [case_state(c[0], c[1], seed=200 + i) for i, c in enumerate(CASES)]
    ens = make_ensemble(scheme, chunk=16)
    ens.initialize(prims)
    refs = [oracle_case(i, scheme, p) for i, p in enumerate(prims)]
    ens.step(1)
    got_dt = host(ens.dt)
    for i, (dom, s, W, U) in enumerate(refs):
        dt = o.compute_dt(W, dom, s)
        W, U = o.rk3_step(W, U, dt, dom, s)
        Wg, Ug = host(ens.state(i)[0]), host(ens.state(i)[1])
        for r in face_regions(dom):
            assert rel(Wg[r], W[r]) < TOL and rel(Ug[r], U[r]) < TOL, (i, rel(Wg[r], W[r]))
        assert abs(got_dt[i] - o.compute_dt(W, dom, s)) <= 1e-14 * dt, i


def test_weno7_is_refused_by_the_batched_stage():
    """The batched stage reads 6-cell windows: WENO7-JS returns JFA_EUNSUPPORTED there, before any
    launch."""
    from jaxfluids_amd import _native
    from jaxfluids_amd.ensemble import Ensemble2D

    say("[weno7 batch refusal]")
    zg = {f: "ZEROGRADIENT" for f in ("east", "west", "north", "south")}
    ens = Ensemble2D([{"cells": (16, 12), "ranges": ((0.0, 1.0), (0.0, 1.0)), "bcs": zg}], nh=4,
                     stencil="WENO7-JS", variable="PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT")
    _, W0 = random_state((16, 12, 2), 4, seed=5)
    ens.initialize([np.ascontiguousarray(W0[:, :, :, :1])])
    with pytest.raises(_native.JFAError, match="unsupported"):
        ens.step(1)
