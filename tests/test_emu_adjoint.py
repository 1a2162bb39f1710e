"""CPU checks of the reverse-mode RK3 stage (jfa_adjoint.h) through the host build of the
library sources (tests/emu), against central finite differences of the oracle's own stage.

For the stage map  f(U_in) = c0 U_in + c1 U_n + m dt R(prim(U_in))  (face halos filled from the
interior, oracle functions only), the kernels' lam_in = J^T lam_out must satisfy the dot-product
identity  <lam_out, J d> = <lam_in, d>  for any direction d; J d comes from central differences
of the oracle.  Every state load and store of the kernels is bounds-checked.
"""
import ctypes
import dataclasses
import os
import sys

import numpy as np
import pytest

from oracle import jxf_oracle as o

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "emu"))

FACES = ("east", "west", "north", "south", "top", "bottom")


@pytest.fixture(scope="module")
def emu():
    import build as emu_build
    from jaxfluids_amd import _native

    lib = ctypes.CDLL(emu_build.build())
    for name in ("jfa_stage_vjp_workspace_bytes", "jfa_rk3_stage_vjp", "jfa_adjoint_set_march_min_waves"):
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = _native.SIGNATURES[name]
    lib.emu_add_range.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    lib.emu_violations.restype = ctypes.c_longlong
    lib.emu_accesses.restype = ctypes.c_longlong
    lib.emu_first_violation.restype = ctypes.c_ulonglong
    return lib


def smooth_state(dom, seed):
    rng = np.random.default_rng(seed)
    X = np.meshgrid(*dom.cell_centers(), indexing="ij")
    k = 2 * np.pi
    ph = rng.uniform(0, 2 * np.pi, 8)
    rho = 1.0 + 0.2 * np.sin(k * X[0] + ph[0]) * np.cos(k * X[1] + ph[1]) + 0.1 * np.cos(k * X[2] + ph[2])
    u = 0.3 * np.sin(k * X[1] + ph[3]) + 0.1
    v = 0.25 * np.cos(k * X[0] + ph[4]) - 0.05
    # w varies in 2-D too: WENO is not differentiable at an exactly constant field (all beta = 0,
    # weights that react to perturbations of order eps^(1/4)), so a constant w would make central
    # differences disagree with the derivative at the point for reasons unrelated to the kernels
    w = 0.2 * np.sin(k * X[2] + ph[5]) + 0.15 * np.cos(k * X[0] - k * X[1] + ph[7])
    p = 1.0 + 0.2 * np.cos(k * X[0] + k * X[1] + ph[6])
    m = (X[0] + 0.3 * X[1]) > 0.5
    rho = np.where(m, 0.6 * rho, rho)
    p = np.where(m, 0.5 * p, p)
    W = np.stack([rho, u, v, w, p])
    return W * (1.0 + 1e-3 * rng.standard_normal(W.shape))


def stage_map(dom, s, stage, dt, Un_int):
    """f(U_in interior) -> U_out interior with the oracle's prim, halo fill and rhs."""
    m = o.RK3_TIMESTEP_MULTIPLIER[stage]
    c0, c1 = (1.0, 0.0) if stage == 0 else o.RK3_CONS_MULTIPLIER[stage - 1]

    def f(Uint):
        W = np.full((5,) + dom.shape, o.EPS64)
        U = np.full((5,) + dom.shape, o.EPS64)
        W[dom.interior] = o.prim_from_cons(Uint, s.gamma)
        U[dom.interior] = Uint
        W, U = o.halo_update(W, U, dom, s)
        return c0 * Uint + c1 * Un_int + dt * m * o.compute_rhs(W, U, dom, s), W

    return f


CASES = [
    # cells, bcs (east, west, north, south, top, bottom), scheme (stencil, riemann, signal speed)
    ((13, 17, 1), ("PERIODIC", "PERIODIC", "PERIODIC", "PERIODIC", "PERIODIC", "PERIODIC"),
     ("WENO5-JS", "HLLC", "EINFELDT")),
    ((11, 9, 1), ("ZEROGRADIENT", "SYMMETRY", "SYMMETRY", "ZEROGRADIENT", "PERIODIC", "PERIODIC"),
     ("WENO5-Z", "HLLC", "EINFELDT")),
    ((10, 12, 1), ("PERIODIC", "PERIODIC", "ZEROGRADIENT", "ZEROGRADIENT", "PERIODIC", "PERIODIC"),
     ("WENO5-JS", "RUSANOV", None)),
    ((9, 10, 1), ("SYMMETRY", "PERIODIC", "PERIODIC", "SYMMETRY", "PERIODIC", "PERIODIC"),
     ("WENO5-Z", "HLLC", "DAVIS")),
    ((12, 11, 1), ("ZEROGRADIENT", "ZEROGRADIENT", "PERIODIC", "PERIODIC", "PERIODIC", "PERIODIC"),
     ("WENO5-JS", "HLLC", "ARITHMETIC")),
    ((10, 9, 1), ("PERIODIC", "PERIODIC", "PERIODIC", "PERIODIC", "PERIODIC", "PERIODIC"),
     ("WENO5-Z", "HLLC", "RUSANOV")),
    ((7, 6, 9), ("PERIODIC", "PERIODIC", "SYMMETRY", "ZEROGRADIENT", "PERIODIC", "PERIODIC"),
     ("WENO5-Z", "HLLC", "EINFELDT")),
]


def richardson(f, U, d, h):
    """Directional derivative by central differences with one Richardson step, (4 D(h/2) - D(h))/3:
    the WENO weights have large third derivatives where smoothness indicators are small, so the
    h^2 term of a plain central difference dominates at any step above rounding level."""
    D = lambda t: (f(U + t * d)[0] - f(U - t * d)[0]) / (2 * t)
    return (4.0 * D(0.5 * h) - D(h)) / 3.0


# marching transpose sweeps (threshold 0) or lane-parallel ones (threshold above any grid here)
SWEEPS = {"march": 0, "lanes": 1 << 40}


def run_vjp(emu, dom, s, W_in, lam_out, stage, dt, sweep="march"):
    from jaxfluids_amd import _native

    assert emu.jfa_adjoint_set_march_min_waves(SWEEPS[sweep]) == 0

    grid = _native.make_grid(dom.cells, dom.nh, dom.dx)
    transport = None if s.transport is None else _native.make_transport(**dataclasses.asdict(s.transport))
    scheme = _native.make_scheme(s.variable, s.stencil, s.riemann, s.signal_speed, s.frozen_state, _native.F64,
                                 s.gamma, s.eps_stencil, transport)
    bcs = _native.bc_array(s.bcs)
    nbytes = emu.jfa_stage_vjp_workspace_bytes(grid, scheme)
    ws = np.zeros(nbytes // 8)
    W = np.ascontiguousarray(W_in)
    lam = np.ascontiguousarray(lam_out)
    lam_in = np.zeros_like(lam)
    lam_n = np.zeros_like(lam)
    emu.emu_clear_ranges()
    for a in (W, lam, lam_in, lam_n, ws):
        emu.emu_add_range(a.ctypes.data, a.nbytes)
    rc = emu.jfa_rk3_stage_vjp(W.ctypes.data, lam.ctypes.data, lam_in.ctypes.data,
                               lam_n.ctypes.data if stage else None, grid, scheme, bcs, stage, dt, ws.ctypes.data,
                               nbytes, None)
    assert rc == 0
    assert emu.emu_violations() == 0, hex(emu.emu_first_violation())
    assert emu.emu_accesses() > 0
    return lam_in, lam_n


@pytest.mark.parametrize("sweep", ["march", "lanes"])
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("stage", [0, 2])
def test_stage_vjp_dot_product_against_oracle_fd(emu, case, stage, sweep):
    cells, bcs, (stencil, riemann, speed) = CASES[case]
    ranges = ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9))
    dom = o.Domain(cells, ranges, nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable="PRIMITIVE", riemann=riemann, signal_speed=speed,
                bcs=dict(zip(FACES, bcs)))
    rng = np.random.default_rng(1000 + 10 * case + stage)
    W0 = smooth_state(dom, 50 + case)
    Uint = o.cons_from_prim(W0, s.gamma)
    Un_int = Uint * (1.0 + 0.01 * rng.standard_normal(Uint.shape))
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    f = stage_map(dom, s, stage, dt, Un_int)
    _, W_in = f(Uint)
    lam_int = rng.standard_normal(Uint.shape)
    lam_out = np.zeros((5,) + dom.shape)
    lam_out[dom.interior] = lam_int
    lam_in, lam_n = run_vjp(emu, dom, s, W_in, lam_out, stage, dt, sweep)
    halo = np.ones(dom.shape, bool)
    halo[dom.interior[1:]] = False
    assert not np.any(lam_in[:, halo]) and not np.any(lam_n[:, halo])  # halos untouched
    g = lam_in[dom.interior]
    for trial in range(3):
        d = rng.standard_normal(Uint.shape) * np.abs(Uint).mean(axis=tuple(range(1, Uint.ndim)), keepdims=True)
        if trial == 2:  # boundary layers only: exercises the halo transpose
            mask = np.ones(Uint.shape[1:], bool)
            for a in dom.active:
                sl = [slice(None)] * 3
                sl[a] = slice(3, -3)
                mask[tuple(sl)] = False
            d = d * mask
        fd = richardson(f, Uint, d, 1e-7)
        lhs = float(np.sum(lam_int * fd))
        rhs = float(np.sum(g * d))
        assert abs(lhs - rhs) <= 1e-8 * np.sum(np.abs(lam_int * fd)), (trial, lhs, rhs)
    if stage:
        c1 = o.RK3_CONS_MULTIPLIER[stage - 1][1]
        assert np.array_equal(lam_n[dom.interior], c1 * lam_int)


@pytest.mark.parametrize("case", [
    ((6, 7, 1), ("PERIODIC", "PERIODIC", "SYMMETRY", "ZEROGRADIENT", "PERIODIC", "PERIODIC"),
     ("WENO5-Z", "HLLC", "EINFELDT")),
    ((7, 5, 1), ("ZEROGRADIENT", "SYMMETRY", "PERIODIC", "PERIODIC", "PERIODIC", "PERIODIC"),
     ("WENO5-JS", "RUSANOV", None)),
], ids=["z-hllc-sym-zg", "js-rusanov-zg-sym"])
def test_stage_vjp_every_entry_against_oracle_fd(emu, case):
    """Every entry of J^T lam on a tiny grid against its own central difference of the oracle's
    stage map (stage 1), so a wrong entry cannot hide in a sum."""
    cells, bcs, (stencil, riemann, speed) = case
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable="PRIMITIVE", riemann=riemann, signal_speed=speed,
                bcs=dict(zip(FACES, bcs)))
    rng = np.random.default_rng(77)
    W0 = smooth_state(dom, 91)
    Uint = o.cons_from_prim(W0, s.gamma)
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    f = stage_map(dom, s, 1, dt, Uint)
    _, W_in = f(Uint)
    lam_int = rng.standard_normal(Uint.shape)
    lam_out = np.zeros((5,) + dom.shape)
    lam_out[dom.interior] = lam_int
    g = run_vjp(emu, dom, s, W_in, lam_out, 1, dt, "lanes")[0][dom.interior]
    # per entry, the closest of four Richardson estimates: a kink of the scheme (S* = 0, a
    # min/max tie) within a step of the state spoils that step only; a wrong derivative
    # matches at none
    err = np.zeros_like(Uint)
    ref_max = 0.0
    for idx in np.ndindex(Uint.shape):
        e = np.zeros_like(Uint)
        scale = max(abs(Uint[idx]), 1e-3)
        e[idx] = scale
        refs = [np.sum(lam_int * richardson(f, Uint, e, h)) / scale for h in (1e-5, 1e-6, 1e-7, 3e-8)]
        err[idx] = min(abs(g[idx] - r) for r in refs)
        ref_max = max(ref_max, abs(refs[0]))
    assert np.max(err) <= 1e-7 * ref_max, np.max(err)


def _rollout_states(W0_int, dom, s, dt, nsteps):
    """Oracle rollout keeping, per step, the primitive buffers entering stages 0, 1, 2."""
    W, U = o.create_buffers(W0_int, dom, s)
    per_step = []
    for _ in range(nsteps):
        stage_W = []
        Un = U
        Wk, Uk = W, U
        for stage in range(3):
            stage_W.append(Wk)
            rhs = o.compute_rhs(Wk, Uk, dom, s)
            Uk = o.rk3_stage_update(Uk, Un, rhs, dt, stage, dom)
            Wk = o.prim_from_cons(Uk, s.gamma)
            Wk, Uk = o.halo_update(Wk, Uk, dom, s)
        per_step.append(stage_W)
        W, U = Wk, Uk
    return W, per_step


def test_rollout_gradient_emulated_vs_oracle_fd(emu):
    """d KE(t_end) / d W_0 of a 3-step periodic Kelvin-Helmholtz rollout (WENO5-JS + HLLC), with
    the step VJP sequence of jaxfluids_amd.adjoint.FixedDtStepper.step_vjp (stages 2, 1, 0; lam_n
    accumulated into the step-start cotangent) on the emulated stage VJPs and the oracle's stage
    states, against Richardson central differences of the oracle rollout."""
    from jaxfluids_amd.workloads import kelvin_helmholtz_primitives

    n, nsteps = 12, 3
    dom = o.Domain((n, n, 1), ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), nh=3)
    s = o.Setup(gamma=1.4, stencil="WENO5-JS", variable="PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                bcs={f: "PERIODIC" for f in FACES})
    W0 = kelvin_helmholtz_primitives(n, seed=3, thickness=0.08)
    W0[3] = 0.05 * np.sin(2 * np.pi * dom.cell_centers()[0])[:, None, None] + 0.02  # w varies (see smooth_state)
    dt = 0.5 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    dV = dom.dx[0] * dom.dx[1]

    def loss(Wi):
        WT, _ = _rollout_states(Wi, dom, s, dt, nsteps)
        Wt = WT[dom.interior]
        return 0.5 * dV * float(np.sum(Wt[0] * (Wt[1] ** 2 + Wt[2] ** 2 + Wt[3] ** 2)))

    WT, per_step = _rollout_states(W0, dom, s, dt, nsteps)
    Wt = WT[dom.interior]
    gW = np.zeros_like(Wt)
    gW[0] = 0.5 * dV * (Wt[1] ** 2 + Wt[2] ** 2 + Wt[3] ** 2)
    gW[1:4] = dV * Wt[0] * Wt[1:4]
    # (dW/dU)^T at the final state
    rho, u, v, w = Wt[0], Wt[1], Wt[2], Wt[3]
    gm1 = s.gamma - 1.0
    lam_int = np.stack([gW[0] - (u * gW[1] + v * gW[2] + w * gW[3]) / rho + gm1 * 0.5 * (u * u + v * v + w * w) * gW[4],
                        gW[1] / rho - gm1 * u * gW[4], gW[2] / rho - gm1 * v * gW[4], gW[3] / rho - gm1 * w * gW[4],
                        gm1 * gW[4]])
    lam = np.zeros((5,) + dom.shape)
    lam[dom.interior] = lam_int
    for stage_W in reversed(per_step):
        l2, l1 = run_vjp(emu, dom, s, stage_W[2], lam, 2, dt)  # l2 = lam_in of stage 2, l1 = its lam_n part
        acc = l1.copy()
        l1_in, acc1 = run_vjp(emu, dom, s, stage_W[1], l2, 1, dt)
        acc += acc1
        l0_in, _ = run_vjp(emu, dom, s, stage_W[0], l1_in, 0, dt)
        lam = acc + l0_in
    lamU = lam[dom.interior]
    W0i = W0
    q2 = W0i[1] ** 2 + W0i[2] ** 2 + W0i[3] ** 2
    g0 = np.stack([lamU[0] + lamU[1] * W0i[1] + lamU[2] * W0i[2] + lamU[3] * W0i[3] + lamU[4] * 0.5 * q2,
                   (lamU[1] + lamU[4] * W0i[1]) * W0i[0], (lamU[2] + lamU[4] * W0i[2]) * W0i[0],
                   (lamU[3] + lamU[4] * W0i[3]) * W0i[0], lamU[4] / gm1])
    rng = np.random.default_rng(5)
    for trial in range(2):
        # the initial pressure is exactly constant (p = 2.5): WENO is not differentiable at a
        # constant field, so directions that move p are left out; after the first stage p varies
        d = rng.standard_normal(W0.shape) * np.array([0.1, 0.05, 0.05, 0.05, 0.0])[:, None, None, None]
        if trial == 1:
            d[[0, 3]] = 0.0  # the velocity perturbation field only
        D = lambda h: (loss(W0 + h * d) - loss(W0 - h * d)) / (2 * h)
        fd = (4.0 * D(0.5e-6) - D(1e-6)) / 3.0
        an = float(np.sum(g0 * d))
        assert abs(fd - an) <= 1e-7 * abs(an), (trial, fd, an)


VISCOUS_CASES = [
    ((7, 6, 8), ("WENO5-Z", "HLLC", "EINFELDT"),
     o.Transport(R=1.0, mu_ref=0.02, T_ref=1.0, mu_exp=0.76, bulk_viscosity=0.01, prandtl=0.71)),
    ((10, 9, 1), ("WENO5-JS", "HLLC", "EINFELDT"),
     o.Transport(R=0.9, mu_ref=0.03, T_ref=0.8, mu_sutherland=0.3, kappa_ref=0.05, kappa_T_ref=1.1, kappa_exp=0.6)),
    ((9, 1, 10), ("WENO5-Z", "RUSANOV", None),
     o.Transport(R=1.2, mu_ref=0.02, prandtl=0.9, viscous=False)),
    ((8, 9, 1), ("WENO5-Z", "HLLC", "DAVIS"),
     o.Transport(R=1.0, mu_ref=0.025, kappa_ref=0.04, kappa_T_ref=0.9, kappa_sutherland=0.4)),
]


@pytest.mark.parametrize("case", range(len(VISCOUS_CASES)))
@pytest.mark.parametrize("stage", [0, 1])
def test_viscous_stage_vjp_against_oracle_fd(emu, case, stage):
    sweep = ("march", "lanes")[(case + stage) % 2]
    """Periodic stage with viscous and / or heat fluxes: the dot-product identity of its dissipative
    part alone (VJP with the transport minus VJP without, against central differences of the
    oracle's stage with minus without) to 1e-9, and of the whole stage."""
    cells, (stencil, riemann, speed), tr = VISCOUS_CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    kw = dict(gamma=1.4, stencil=stencil, variable="PRIMITIVE", riemann=riemann, signal_speed=speed,
              bcs={f: "PERIODIC" for f in FACES})
    sv, si = o.Setup(transport=tr, **kw), o.Setup(**kw)
    rng = np.random.default_rng(3000 + 10 * case + stage)
    W0 = smooth_state(dom, 80 + case)
    Uint = o.cons_from_prim(W0, 1.4)
    Un_int = Uint * (1.0 + 0.01 * rng.standard_normal(Uint.shape))
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, sv)[0], dom, sv)
    fv, fi = stage_map(dom, sv, stage, dt, Un_int), stage_map(dom, si, stage, dt, Un_int)
    _, W_in = fv(Uint)
    lam_int = rng.standard_normal(Uint.shape)
    lam_out = np.zeros((5,) + dom.shape)
    lam_out[dom.interior] = lam_int
    gv = run_vjp(emu, dom, sv, W_in, lam_out, stage, dt, sweep)[0][dom.interior]
    gi = run_vjp(emu, dom, si, W_in, lam_out, stage, dt, sweep)[0][dom.interior]
    for trial in range(2):
        d = rng.standard_normal(Uint.shape) * np.abs(Uint).mean(axis=tuple(range(1, Uint.ndim)), keepdims=True)
        # dissipative part: a smooth rational map, so a larger step leaves only rounding
        fdv, fdi = richardson(fv, Uint, d, 1e-5), richardson(fi, Uint, d, 1e-5)
        dl = float(np.sum(lam_int * (fdv - fdi)))
        dr = float(np.sum((gv - gi) * d))
        scale = np.sum(np.abs(lam_int * (fdv - fdi)))
        assert scale > 1e-3 * np.sum(np.abs(lam_int * fdv))  # the dissipative part is not negligible
        assert abs(dl - dr) <= 1e-9 * scale, (trial, dl, dr, scale)
        # whole stage: the WENO part's differences need the step chosen per state (see richardson)
        errs = []
        for h in (1e-5, 1e-6, 1e-7):
            fd = richardson(fv, Uint, d, h)
            errs.append(abs(float(np.sum(lam_int * fd)) - float(np.sum(gv * d))) / np.sum(np.abs(lam_int * fd)))
        assert min(errs) <= 1e-6, (trial, errs)
