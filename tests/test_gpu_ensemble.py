"""GPU parity of the batched 2-D ensemble (one persistent launch per RK3 stage) against the
oracle, case by case, and against the single-domain fused path.

Cases have uneven sizes (x chunks and 64-cell y tiles end ragged), different boundary
types and different CFL time steps.  FP64, relative L-inf <= 1e-12.
"""
import numpy as np
import pytest
import torch

from oracle import jxf_oracle as o

pytestmark = pytest.mark.gpu

TOL = 1e-12
FACES2 = ("east", "west", "north", "south")
FACES = FACES2 + ("top", "bottom")

# (cells, ranges, bcs by face) -- sizes straddle the 64-cell y tile and the x chunk
CASES = [
    ((37, 70), ((0.0, 1.0), (0.0, 1.3)), {"east": "PERIODIC", "west": "PERIODIC", "north": "PERIODIC", "south": "PERIODIC"}),
    ((64, 64), ((0.0, 2.0), (0.0, 2.0)), {f: "ZEROGRADIENT" for f in FACES2}),
    ((70, 130), ((0.0, 1.0), (0.0, 1.0)), {f: "SYMMETRY" for f in FACES2}),
    ((5, 9), ((0.0, 0.5), (0.0, 0.9)), {"east": "PERIODIC", "west": "PERIODIC", "north": "ZEROGRADIENT", "south": "SYMMETRY"}),
    ((129, 3), ((0.0, 1.0), (0.0, 0.1)), {"east": "ZEROGRADIENT", "west": "SYMMETRY", "north": "ZEROGRADIENT", "south": "ZEROGRADIENT"}),
]

SCHEMES = [
    ("WENO5-JS", "PRIMITIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
    ("WENO5-Z", "CHAR-PRIMITIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
    ("WENO5-Z", "CONSERVATIVE", "RUSANOV", "EINFELDT", "ARITHMETIC"),
    ("WENO5-JS", "CHAR-CONSERVATIVE", "HLLC", "TORO", "ROE"),
]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def say(msg):
    print(msg, flush=True)


def case_state(cells, ranges, seed):
    """Smooth field + an oblique density/pressure jump, with seeded 1e-3 noise (breaks exact
    WENO plateau ties, whose WENO-Z weights are rounding-determined)."""
    rng = np.random.default_rng(seed)
    nx, ny = cells
    x = ranges[0][0] + (np.arange(nx) + 0.5) * (ranges[0][1] - ranges[0][0]) / nx
    y = ranges[1][0] + (np.arange(ny) + 0.5) * (ranges[1][1] - ranges[1][0]) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 6)
    k = 2 * np.pi
    rho = 1.0 + 0.3 * np.sin(k * X + ph[0]) * np.cos(k * Y + ph[1])
    u = 0.3 * np.sin(k * Y + ph[2]) + rng.uniform(-0.3, 0.3)
    v = 0.3 * np.cos(k * X + ph[3]) + rng.uniform(-0.3, 0.3)
    p = 1.0 + 0.2 * np.cos(k * X + k * Y + ph[4])
    m = (X + 0.3 * Y) > 0.55 * (ranges[0][1] + 0.3 * ranges[1][1])
    rho = np.where(m, 0.5 * rho, rho)
    p = np.where(m, 0.4 * p, p)
    W = np.stack([rho, u, v, np.zeros_like(X), p])[..., None]
    return W * (1.0 + 1e-3 * rng.standard_normal(W.shape))


def oracle_case(c, scheme, prim):
    cells, ranges, bcs = CASES[c]
    dom = o.Domain((cells[0], cells[1], 1), (ranges[0], ranges[1], (0.0, 1.0)), nh=3)
    st, var, rs, ss, fz = scheme
    s = o.Setup(gamma=1.4, stencil=st, variable=var, riemann=rs, signal_speed=ss, frozen_state=fz,
                bcs={**bcs, "top": "PERIODIC", "bottom": "PERIODIC"})
    W, U = o.create_buffers(prim, dom, s)
    return dom, s, W, U


def make_ensemble(scheme, chunk=16, nh=3, dtype=torch.float64):
    from jaxfluids_amd.ensemble import Ensemble2D

    st, var, rs, ss, fz = scheme
    cases = [{"cells": c, "ranges": r, "bcs": b} for c, r, b in CASES]
    return Ensemble2D(cases, stencil=st, variable=var, riemann=rs, signal_speed=ss, frozen_state=fz, nh=nh,
                      chunk=chunk, dtype=dtype)


def face_regions(dom):
    """Interior plus the four face-halo slabs of a 2-D buffer."""
    out = [dom.interior]
    for a in range(2):
        sl = [dom.dom(0), dom.dom(1), slice(None)]
        for part in (slice(0, dom.nh), slice(-dom.nh, None)):
            sl[a] = part
            out.append((Ellipsis,) + tuple(sl))
    return out


@pytest.mark.parametrize("scheme", SCHEMES, ids=lambda s: "-".join(s))
@pytest.mark.parametrize("chunk", [16, 64])
def test_ensemble_two_steps_match_oracle(scheme, chunk):
    """Initial buffers, per-case dt and two RK3 steps of 5 uneven cases in one ensemble, each
    against the oracle run on that case alone."""
    say(f"[ensemble] {scheme} chunk={chunk}")
    prims = [case_state(c[0], c[1], seed=100 + i) for i, c in enumerate(CASES)]
    ens = make_ensemble(scheme, chunk=chunk)
    ens.initialize(prims)
    refs = [oracle_case(i, scheme, p) for i, p in enumerate(prims)]
    dts = [o.compute_dt(W, dom, s) for dom, s, W, U in refs]
    got_dt = host(ens.dt)
    for i, (dom, s, W, U) in enumerate(refs):
        Wg, Ug = host(ens.state(i)[0]), host(ens.state(i)[1])
        for r in face_regions(dom):
            assert rel(Wg[r], W[r]) < TOL and rel(Ug[r], U[r]) < TOL, (i, "initial")
        assert abs(got_dt[i] - dts[i]) <= 1e-14 * dts[i], (i, got_dt[i], dts[i])
    for step in range(2):
        ens.step(1)
        got_dt = host(ens.dt)
        for i, (dom, s, W, U) in enumerate(refs):
            W, U = o.rk3_step(W, U, dts[i], dom, s)
            refs[i] = (dom, s, W, U)
            Wg, Ug = host(ens.state(i)[0]), host(ens.state(i)[1])
            for r in face_regions(dom):
                assert rel(Wg[r], W[r]) < TOL, (step, i, rel(Wg[r], W[r]))
                assert rel(Ug[r], U[r]) < TOL, (step, i, rel(Ug[r], U[r]))
            dts[i] = o.compute_dt(W, dom, s)
            assert abs(got_dt[i] - dts[i]) <= 1e-14 * dts[i], (step, i, got_dt[i], dts[i])


def test_ensemble_graph_replays_match_eager():
    """A captured step replayed 3 times (with eager work and syncs between replays) gives the
    eager ensemble's state bit for bit after every replay."""
    say("[ensemble graph]")
    scheme = SCHEMES[1]
    prims = [case_state(c[0], c[1], seed=200 + i) for i, c in enumerate(CASES)]
    a = make_ensemble(scheme)
    b = make_ensemble(scheme)
    a.initialize(prims)
    b.initialize(prims)
    for r in range(3):
        a.step(1, use_graph=True)
        b.step(1)
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        assert torch.equal(a.A[0], b.A[0]) and torch.equal(a.A[1], b.A[1]), r
        assert torch.equal(a.dt, b.dt), r


def test_ensemble_matches_single_domain_fused_path():
    """Every ensemble case equals the same case advanced alone by the fused single-domain
    stage (jfa_rk3_stage_fused) with the same dt."""
    from jaxfluids_amd import _native
    import ctypes

    say("[ensemble vs fused]")
    scheme = SCHEMES[1]
    prims = [case_state(c[0], c[1], seed=300 + i) for i, c in enumerate(CASES)]
    ens = make_ensemble(scheme, chunk=32)
    ens.initialize(prims)
    singles = []
    for i in range(len(CASES)):
        W, U = ens.state(i)
        singles.append((W.clone(), U.clone()))
    dt0 = ens.dt.clone()
    ens.step(1)
    lib = _native.load()
    for i, (W, U) in enumerate(singles):
        grid, bcs = ens.grids[i], ens.bcs[i]
        nb = lib.jfa_stage_workspace_bytes(grid, ens.scheme)
        ws = torch.empty(max(int(nb), 16), dtype=torch.uint8, device="cuda")
        Wb, Ub, Wc, Uc = (torch.empty_like(W) for _ in range(4))
        dt = dt0[i:i + 1].clone()
        chain = [(W, U, Wb, Ub), (Wb, Ub, Wc, Uc), (Wc, Uc, W, U)]
        for stage, (wi, ui, wo, uo) in enumerate(chain):
            rc = lib.jfa_rk3_stage_fused(_native.ptr(wi), _native.ptr(ui), _native.ptr(U if stage else None),
                                         _native.ptr(wo), _native.ptr(uo), grid, ens.scheme, bcs, stage,
                                         _native.ptr(dt), 0.0, ens.cfl, ens.eps,
                                         _native.ptr(dt) if stage == 2 else _native.ptr(None),
                                         _native.ptr(ws), ws.numel(), _native.stream_ptr())
            _native.check(rc, "jfa_rk3_stage_fused")
        We, Ue = ens.state(i)
        nh = ens.nh
        I = (slice(None), slice(nh, -nh), slice(nh, -nh), slice(None))
        assert rel(host(We)[I], host(W)[I]) < TOL and rel(host(Ue)[I], host(U)[I]) < TOL, i
        assert abs(host(ens.dt)[i] - host(dt)[0]) <= 1e-14 * host(dt)[0], i


def test_ensemble_autotune_keeps_state():
    """The tuner times candidates on the live state and restores it; the tuned ensemble then
    steps to the same state as an untuned one."""
    say("[ensemble autotune]")
    from jaxfluids_amd import ensemble as E

    scheme = SCHEMES[0]
    prims = [case_state(c[0], c[1], seed=400 + i) for i, c in enumerate(CASES)]
    a = make_ensemble(scheme, chunk=8)
    b = make_ensemble(scheme, chunk=64)
    a.initialize(prims)
    b.initialize(prims)
    W0 = a.A[0].clone()
    E._TUNE_CACHE.clear()
    ch = a.autotune(chunks=(8, 32, 64), reps=1)
    assert ch in (8, 32, 64) and a.chunk == ch
    assert torch.equal(a.A[0], W0)
    a.step(2)
    b.step(2)
    assert rel(host(a.A[0]), host(b.A[0])) < TOL


def test_ensemble_fp32_storage_error_stated():
    """Float32 storage (FP64 arithmetic) against the FP64 oracle after one step: the error is
    the float32 rounding of the state, stated here as < 1e-6 relative L-inf."""
    say("[ensemble fp32]")
    scheme = SCHEMES[1]
    prims = [case_state(c[0], c[1], seed=500 + i) for i, c in enumerate(CASES)]
    ens = make_ensemble(scheme, dtype=torch.float32)
    ens.initialize(prims)
    ens.step(1)
    for i, p in enumerate(prims):
        dom, s, W, U = oracle_case(i, scheme, p)
        dt = float(host(ens.dt)[i])  # compare the step itself: use the ensemble's own first dt
        W1, U1 = o.rk3_step(W, U, o.compute_dt(W, dom, s), dom, s)
        assert rel(host(ens.state(i)[0])[dom.interior], W1[dom.interior]) < 1e-6, i
        assert dt > 0


@pytest.mark.parametrize("var", ["PRIMITIVE", "CHAR-PRIMITIVE"])
def test_ensemble_with_interpolation_limiter(var):
    """The batched stage with the positivity limiter of the face states: the near-vacuum pocket
    (not finite after one step without the limiter) beside a smooth case, one RK3 step each with
    its own CFL dt, against the oracle."""
    from jaxfluids_amd.ensemble import Ensemble2D
    from test_oracle_pins import _pocket_state

    say(f"[ensemble limiter] {var}")
    zg = {f: "ZEROGRADIENT" for f in FACES2}
    dom_p, Wp = _pocket_state()
    cells_s, ranges_s, bcs_s = CASES[1]
    prims = [Wp, case_state(cells_s, ranges_s, seed=7)]
    cases = [{"cells": dom_p.cells[:2], "ranges": dom_p.ranges[:2], "bcs": zg},
             {"cells": cells_s, "ranges": ranges_s, "bcs": bcs_s}]
    ens = Ensemble2D(cases, stencil="WENO5-JS", variable=var, riemann="HLLC", signal_speed="EINFELDT",
                     interpolation_limiter=True, chunk=16)
    ens.initialize(prims)
    ens.step(1)
    refs = []
    for (cells, ranges, bcs), prim in zip(((dom_p.cells[:2], dom_p.ranges[:2], zg), CASES[1]), prims):
        dom = o.Domain((cells[0], cells[1], 1), (ranges[0], ranges[1], (0.0, 1.0)), nh=3)
        s = o.Setup(stencil="WENO5-JS", variable=var, riemann="HLLC", signal_speed="EINFELDT",
                    bcs={**bcs, "top": "PERIODIC", "bottom": "PERIODIC"}, interpolation_limiter=True)
        W, U = o.create_buffers(prim, dom, s)
        refs.append((dom, o.rk3_step(W, U, o.compute_dt(W, dom, s), dom, s)[0]))
    for i, (dom, W) in enumerate(refs):
        Wg = host(ens.state(i)[0])
        assert np.all(np.isfinite(W[dom.interior]))
        assert rel(Wg[dom.interior], W[dom.interior]) < TOL, (i, rel(Wg[dom.interior], W[dom.interior]))


@pytest.mark.parametrize("fl,partition", [("SIMPLE", "UNIFORM"), ("NASA", "UNIFORM"), ("HAS", "UNIFORM"),
                                          ("HAS", "WAVESPEED"), ("SIMPLE", "WAVESPEED")])
def test_ensemble_with_flux_limiter(fl, partition):
    """The batched stage with both positivity limiters: the near-vacuum pocket beside a smooth
    case, one RK3 step each at the pocket's CFL 0.9 time step (fixed dt), against the oracle."""
    from jaxfluids_amd.ensemble import Ensemble2D
    from test_oracle_pins import _pocket_state

    say(f"[ensemble flux limiter] {fl} {partition}")
    zg = {f: "ZEROGRADIENT" for f in FACES2}
    dom_p, Wp = _pocket_state()
    cells_s, ranges_s, bcs_s = CASES[1]
    prims = [Wp, case_state(cells_s, ranges_s, seed=8)]
    specs = [(dom_p.cells[:2], dom_p.ranges[:2], zg), CASES[1]]
    ens = Ensemble2D([{"cells": c, "ranges": r, "bcs": b} for c, r, b in specs], stencil="WENO5-JS",
                     variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT", interpolation_limiter=True,
                     flux_limiter=fl, flux_partition=partition, chunk=16)
    ens.initialize(prims)
    dts = []
    refs = []
    for (cells, ranges, bcs), prim in zip(specs, prims):
        dom = o.Domain((cells[0], cells[1], 1), (ranges[0], ranges[1], (0.0, 1.0)), nh=3)
        s = o.Setup(stencil="WENO5-JS", variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                    bcs={**bcs, "top": "PERIODIC", "bottom": "PERIODIC"}, interpolation_limiter=True, flux_limiter=fl,
                    flux_partition=partition)
        W, U = o.create_buffers(prim, dom, s)
        # HAS at CFL 0.6: its blend weight is ill-conditioned where the first-order state is inadmissible
        dt = (0.6 if fl == "HAS" else 0.9) / 0.5 * o.compute_dt(W, dom, s)
        dts.append(dt)
        refs.append((dom, o.rk3_step(W, U, dt, dom, s)[0]))
    ens.dt.copy_(torch.tensor(dts, dtype=torch.float64, device="cuda"))
    ens.step(1)
    for i, (dom, W) in enumerate(refs):
        Wg = host(ens.state(i)[0])
        assert np.all(np.isfinite(W[dom.interior]))
        assert rel(Wg[dom.interior], W[dom.interior]) < TOL, (i, rel(Wg[dom.interior], W[dom.interior]))


@pytest.mark.parametrize("var", ["PRIMITIVE", "CHAR-PRIMITIVE"])
def test_four_quadrant_workload_matches_oracle(var):
    """The four-quadrant Riemann ensemble (seeded quadrant states rho ~ U(0.1, 1.5),
    p ~ U(0.03, 1.5), u, v ~ U(-1, 1); WENO5-JS + HLLC-Einfeldt, ZEROGRADIENT, nh = 5, per-case CFL
    dt) at reduced grid sizes drawn from {24, 32, 48}^2: three steps of 8 cases in one ensemble,
    each against the oracle run on that case alone, state and dt."""
    from jaxfluids_amd.ensemble import Ensemble2D
    from jaxfluids_amd.workloads import four_quadrant_cases, quadrant_primitives

    say(f"[four-quadrant] {var}")
    qc = four_quadrant_cases(8, seed=0, sizes=(24, 32, 48))
    zg = {f: "ZEROGRADIENT" for f in FACES2}
    ens = Ensemble2D([{"cells": (c.nx, c.ny), "ranges": ((0.0, 1.0), (0.0, 1.0)), "bcs": zg} for c in qc],
                     stencil="WENO5-JS", variable=var, riemann="HLLC", signal_speed="EINFELDT", nh=5, chunk=16)
    prims = [quadrant_primitives(c) for c in qc]
    ens.initialize(prims)
    refs = []
    for c, p in zip(qc, prims):
        dom = o.Domain((c.nx, c.ny, 1), ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), nh=5)
        s = o.Setup(gamma=1.4, stencil="WENO5-JS", variable=var, riemann="HLLC", signal_speed="EINFELDT",
                    bcs={**zg, "top": "PERIODIC", "bottom": "PERIODIC"})
        W, U = o.create_buffers(p, dom, s)
        refs.append([dom, s, W, U, o.compute_dt(W, dom, s)])
    for step in range(3):
        ens.step(1)
        got_dt = host(ens.dt)
        for i, r in enumerate(refs):
            dom, s, W, U, dt = r
            W, U = o.rk3_step(W, U, dt, dom, s)
            r[2:] = [W, U, o.compute_dt(W, dom, s)]
            I = dom.interior
            assert np.all(np.isfinite(W[I])), (step, i)
            Wg, Ug = host(ens.state(i)[0]), host(ens.state(i)[1])
            for reg in face_regions(dom):
                assert rel(Wg[reg], W[reg]) < TOL, (step, i, rel(Wg[reg], W[reg]))
                assert rel(Ug[reg], U[reg]) < TOL, (step, i, rel(Ug[reg], U[reg]))
            assert abs(got_dt[i] - r[4]) <= 1e-14 * r[4], (step, i, got_dt[i], r[4])


def test_ensemble_time_step_propagates_nan():
    """A NaN cell in one case makes that case's dt NaN (the reference's jnp.max propagates NaN,
    time_step_size.py:109); the other cases keep finite time steps."""
    say("[ensemble nan dt]")
    scheme = SCHEMES[1]
    prims = [case_state(c[0], c[1], seed=600 + i) for i, c in enumerate(CASES)]
    prims[2][0, 10, 20, 0] = np.nan
    ens = make_ensemble(scheme)
    ens.initialize(prims)
    dt0 = host(ens.dt)
    assert np.isnan(dt0[2]) and np.all(np.isfinite(np.delete(dt0, 2)))
    ens.step(1)
    dt1 = host(ens.dt)
    assert np.isnan(dt1[2]) and np.all(np.isfinite(np.delete(dt1, 2)))
