"""GPU checks of the reverse mode (jfa_rk3_stage_vjp, FixedDtStepper, RK3Rollout).

* stage VJP on the device: dot-product identity against central differences of the oracle's
  stage map (as the CPU emulation test does for the same source), and every entry against
  torch.autograd of the PyTorch restatement (oracle/torch_stage.py);
* rollout gradient (torch.autograd through RK3Rollout) against Richardson central differences
  of the oracle rollout (small case) and of the device forward (Kelvin-Helmholtz 64^2, 40 steps,
  the 1e-6 bar of the adjoint workload);
* binomial checkpointing: the gradient with 3 snapshot slots equals the stored-every-step
  gradient bit for bit, and the forward steps taken equal the schedule's count.
"""
import numpy as np
import pytest
import torch

from oracle import jxf_oracle as o
from test_emu_adjoint import CASES, FACES, _rollout_states, richardson, smooth_state, stage_map

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def dev(x):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("case", [0, 1, 2, 6])
@pytest.mark.parametrize("stage", [0, 1, 2])
def test_stage_vjp_device_dot_product(case, stage):
    from jaxfluids_amd import _native

    say(f"[stage vjp] case {case} stage {stage}")
    cells, bcs, (stencil, riemann, speed) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable="PRIMITIVE", riemann=riemann, signal_speed=speed,
                bcs=dict(zip(FACES, bcs)))
    rng = np.random.default_rng(2000 + 10 * case + stage)
    W0 = smooth_state(dom, 60 + case)
    Uint = o.cons_from_prim(W0, s.gamma)
    Un_int = Uint * (1.0 + 0.01 * rng.standard_normal(Uint.shape))
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    f = stage_map(dom, s, stage, dt, Un_int)
    _, W_in = f(Uint)
    lam_int = rng.standard_normal(Uint.shape)
    lam_out = np.zeros((5,) + dom.shape)
    lam_out[dom.interior] = lam_int
    grid = _native.make_grid(dom.cells, dom.nh, dom.dx)
    scheme = _native.make_scheme("PRIMITIVE", stencil, riemann, speed, "ARITHMETIC", _native.F64, 1.4,
                                 s.eps_stencil)
    lib = _native.load()
    ws = torch.empty(lib.jfa_stage_vjp_workspace_bytes(grid, scheme), dtype=torch.uint8, device="cuda")
    Wd, ld = dev(W_in), dev(lam_out)
    li, ln = torch.zeros_like(ld), torch.zeros_like(ld)
    P = _native.ptr
    rc = lib.jfa_rk3_stage_vjp(P(Wd), P(ld), P(li), P(ln if stage else None), grid, scheme,
                               _native.bc_array(s.bcs), stage, dt, P(ws), ws.numel(), _native.stream_ptr())
    _native.check(rc, "jfa_rk3_stage_vjp")
    g = host(li)[dom.interior]
    for trial in range(2):
        d = rng.standard_normal(Uint.shape) * np.abs(Uint).mean(axis=tuple(range(1, Uint.ndim)), keepdims=True)
        rhs = float(np.sum(g * d))
        # the step that resolves the WENO weights' higher derivatives depends on the state
        errs = []
        for h in (1e-6, 1e-7, 3e-8):
            fd = richardson(f, Uint, d, h)
            errs.append(abs(float(np.sum(lam_int * fd)) - rhs) / np.sum(np.abs(lam_int * fd)))
        assert min(errs) <= 1e-8, (trial, errs)
    if stage:
        c1 = o.RK3_CONS_MULTIPLIER[stage - 1][1]
        assert np.allclose(host(ln)[dom.interior], c1 * lam_int, rtol=0, atol=0)


def _vjp_vs_autograd(dom, s, stage, seed_state, seed):
    import dataclasses

    from jaxfluids_amd import _native
    from oracle import torch_stage as ts

    rng = np.random.default_rng(seed)
    W0 = smooth_state(dom, seed_state)
    Uint = o.cons_from_prim(W0, s.gamma)
    Un_int = Uint * (1.0 + 0.01 * rng.standard_normal(Uint.shape))
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    _, W_in = stage_map(dom, s, stage, dt, Un_int)(Uint)
    lam_int = rng.standard_normal(Uint.shape)
    lam_out = np.zeros((5,) + dom.shape)
    lam_out[dom.interior] = lam_int
    grid = _native.make_grid(dom.cells, dom.nh, dom.dx)
    tr = None if s.transport is None else _native.make_transport(**dataclasses.asdict(s.transport))
    scheme = _native.make_scheme(s.variable, s.stencil, s.riemann, s.signal_speed, s.frozen_state, _native.F64, 1.4,
                                 s.eps_stencil, tr)
    lib = _native.load()
    ws = torch.empty(lib.jfa_stage_vjp_workspace_bytes(grid, scheme), dtype=torch.uint8, device="cuda")
    Wd, ld = dev(W_in), dev(lam_out)
    li, ln = torch.zeros_like(ld), torch.zeros_like(ld)
    P = _native.ptr
    rc = lib.jfa_rk3_stage_vjp(P(Wd), P(ld), P(li), P(ln if stage else None), grid, scheme,
                               _native.bc_array(s.bcs), stage, dt, P(ws), ws.numel(), _native.stream_ptr())
    _native.check(rc, "jfa_rk3_stage_vjp")
    U = torch.as_tensor(Uint).clone().requires_grad_(True)
    out = ts.stage_map(U, torch.as_tensor(Un_int), dom, s.stencil, s.riemann, s.signal_speed, s.gamma, s.bcs, stage,
                       dt, o.EPS64, s.transport, s.variable, s.frozen_state)
    (ref,) = torch.autograd.grad(out, U, torch.as_tensor(lam_int))
    ref = ref.numpy()
    got = host(li)[dom.interior]
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err <= 1e-11, err


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("stage", [0, 1, 2])
def test_stage_vjp_device_matches_autograd(case, stage):
    """Every entry of the device J^T lam against torch.autograd of the PyTorch restatement of the
    stage (oracle/torch_stage.py), the stand-in for jax.grad; 1e-11 of the largest entry (the
    rounding of the two reverse evaluations, as in the host-emulated test)."""
    say(f"[stage vjp vs autograd] case {case} stage {stage}")
    cells, bcs, (stencil, riemann, speed) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable="PRIMITIVE", riemann=riemann, signal_speed=speed,
                bcs=dict(zip(FACES, bcs)))
    _vjp_vs_autograd(dom, s, stage, 70 + case, 5000 + 10 * case + stage)


@pytest.mark.parametrize("case", [0, 1, 2, 3])
@pytest.mark.parametrize("stage", [0, 2])
def test_viscous_stage_vjp_device_matches_autograd(case, stage):
    """The same with viscous and / or heat fluxes (periodic)."""
    from test_emu_adjoint import VISCOUS_CASES

    say(f"[viscous stage vjp vs autograd] case {case} stage {stage}")
    cells, (stencil, riemann, speed), tr = VISCOUS_CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable="PRIMITIVE", riemann=riemann, signal_speed=speed,
                bcs={f: "PERIODIC" for f in FACES}, transport=tr)
    _vjp_vs_autograd(dom, s, stage, 85 + case, 6000 + 10 * case + stage)


def _kh_stepper(n, dt, variable="PRIMITIVE"):
    from jaxfluids_amd.adjoint import FixedDtStepper

    return FixedDtStepper((n, n, 1), ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), {f: "PERIODIC" for f in FACES}, dt,
                          nh=3, stencil="WENO5-JS", riemann="HLLC", signal_speed="EINFELDT", variable=variable)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("stage", [0, 2])
def test_char_primitive_stage_vjp_device_matches_autograd(case, stage):
    """CHAR-PRIMITIVE reconstruction (ARITHMETIC frozen state): every entry of the device J^T lam
    against torch.autograd of the restatement."""
    say(f"[char stage vjp vs autograd] case {case} stage {stage}")
    cells, bcs, (stencil, riemann, speed) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable="CHAR-PRIMITIVE", riemann=riemann, signal_speed=speed,
                frozen_state="ARITHMETIC", bcs=dict(zip(FACES, bcs)))
    _vjp_vs_autograd(dom, s, stage, 90 + case, 8000 + 10 * case + stage)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("stage", [0, 2])
def test_conservative_stage_vjp_device_matches_autograd(case, stage):
    """CONSERVATIVE reconstruction: every entry of the device J^T lam against torch.autograd of
    the restatement."""
    say(f"[conservative stage vjp vs autograd] case {case} stage {stage}")
    cells, bcs, (stencil, riemann, speed) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable="CONSERVATIVE", riemann=riemann, signal_speed=speed,
                frozen_state="ARITHMETIC", bcs=dict(zip(FACES, bcs)))
    _vjp_vs_autograd(dom, s, stage, 90 + case, 8500 + 10 * case + stage)


@pytest.mark.parametrize("variable", ["PRIMITIVE", "CONSERVATIVE", "CHAR-PRIMITIVE"])
@pytest.mark.parametrize("case", [0, 1, 6])
def test_toro_speeds_stage_vjp_device_matches_autograd(case, variable):
    """HLLC with Toro's wave speeds: every entry of the device J^T lam against autograd."""
    say(f"[toro stage vjp vs autograd] case {case} {variable}")
    cells, bcs, (stencil, _, _) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable=variable, riemann="HLLC", signal_speed="TORO",
                frozen_state="ARITHMETIC", bcs=dict(zip(FACES, bcs)))
    _vjp_vs_autograd(dom, s, 2, 95 + case, 9500 + 10 * case)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_roe_char_primitive_stage_vjp_device_matches_autograd(case):
    """CHAR-PRIMITIVE with the ROE frozen state: every entry of the device J^T lam against
    torch.autograd of the restatement."""
    say(f"[roe char stage vjp vs autograd] case {case}")
    cells, bcs, (stencil, riemann, speed) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable="CHAR-PRIMITIVE", riemann=riemann, signal_speed=speed,
                frozen_state="ROE", bcs=dict(zip(FACES, bcs)))
    _vjp_vs_autograd(dom, s, 2, 99 + case, 9900 + 10 * case)


@pytest.mark.parametrize("variable", ["PRIMITIVE", "CHAR-PRIMITIVE"])
@pytest.mark.parametrize("stencil", ["WENO1", "WENO3-JS", "WENO3-Z", "TENO5", "TENO5-A"])
@pytest.mark.parametrize("case", [0, 1, 6])
def test_stencil_stage_vjp_device_matches_autograd(case, stencil, variable):
    """WENO1, WENO3-JS/-Z, TENO5 and TENO5-A: every entry of the device J^T lam against
    torch.autograd of the restatement."""
    say(f"[stencil stage vjp vs autograd] case {case} {stencil} {variable}")
    cells, bcs, (_, riemann, speed) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable=variable, riemann=riemann, signal_speed=speed,
                frozen_state="ARITHMETIC", bcs=dict(zip(FACES, bcs)))
    _vjp_vs_autograd(dom, s, 2, 91 + case, 9150 + 10 * case)


@pytest.mark.parametrize("frozen", ["ARITHMETIC", "ROE"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_char_conservative_stage_vjp_device_matches_autograd(case, frozen):
    """CHAR-CONSERVATIVE (ARITHMETIC and ROE frozen states): every entry of the device J^T lam
    against torch.autograd of the restatement."""
    say(f"[char-conservative stage vjp vs autograd] case {case} {frozen}")
    cells, bcs, (stencil, riemann, speed) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable="CHAR-CONSERVATIVE", riemann=riemann, signal_speed=speed,
                frozen_state=frozen, bcs=dict(zip(FACES, bcs)))
    _vjp_vs_autograd(dom, s, 2, 93 + case, 9800 + 10 * case)


@pytest.mark.parametrize("frozen", ["ARITHMETIC", "ROE"])
@pytest.mark.parametrize("kind", ["ROE", "CLLF", "LLF"])
@pytest.mark.parametrize("case", [0, 1, 6])
def test_flux_splitting_stage_vjp_device_matches_autograd(case, kind, frozen):
    """The flux-splitting scheme: every entry of the device J^T lam against autograd."""
    say(f"[flux splitting stage vjp vs autograd] case {case} {kind} {frozen}")
    cells, bcs, (stencil, _, _) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable="PRIMITIVE", riemann="FLUX-SPLITTING-" + kind,
                signal_speed=None, frozen_state=frozen, bcs=dict(zip(FACES, bcs)))
    _vjp_vs_autograd(dom, s, 2, 92 + case, 9750 + 10 * case)


@pytest.mark.parametrize("variable", ["PRIMITIVE", "CHAR-PRIMITIVE"])
@pytest.mark.parametrize("case", [0, 1, 6])
def test_ausmp_stage_vjp_device_matches_autograd(case, variable):
    """AUSM+: every entry of the device J^T lam against autograd."""
    say(f"[ausmp stage vjp vs autograd] case {case} {variable}")
    cells, bcs, (stencil, _, _) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable=variable, riemann="AUSMP", signal_speed=None,
                frozen_state="ARITHMETIC", bcs=dict(zip(FACES, bcs)))
    _vjp_vs_autograd(dom, s, 2, 94 + case, 9450 + 10 * case)


@pytest.mark.parametrize("variable", ["PRIMITIVE", "CHAR-PRIMITIVE"])
@pytest.mark.parametrize("speed", ["EINFELDT", "ARITHMETIC", "TORO"])
@pytest.mark.parametrize("case", [0, 1, 6])
def test_hllc_lm_stage_vjp_device_matches_autograd(case, speed, variable):
    """HLLC-LM: every entry of the device J^T lam against autograd."""
    say(f"[hllc-lm stage vjp vs autograd] case {case} {speed} {variable}")
    cells, bcs, (stencil, _, _) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable=variable, riemann="HLLC-LM", signal_speed=speed,
                frozen_state="ARITHMETIC", bcs=dict(zip(FACES, bcs)))
    _vjp_vs_autograd(dom, s, 2, 96 + case, 9600 + 10 * case)


@pytest.mark.parametrize("variable", ["PRIMITIVE", "CHAR-PRIMITIVE"])
@pytest.mark.parametrize("speed", ["EINFELDT", "DAVIS", "TORO"])
@pytest.mark.parametrize("case", [0, 1, 6])
def test_hll_stage_vjp_device_matches_autograd(case, speed, variable):
    """HLL: every entry of the device J^T lam against autograd."""
    say(f"[hll stage vjp vs autograd] case {case} {speed} {variable}")
    cells, bcs, (stencil, _, _) = CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    s = o.Setup(gamma=1.4, stencil=stencil, variable=variable, riemann="HLL", signal_speed=speed,
                frozen_state="ARITHMETIC", bcs=dict(zip(FACES, bcs)))
    _vjp_vs_autograd(dom, s, 2, 97 + case, 9700 + 10 * case)


@pytest.mark.parametrize("variable", ["CHAR-PRIMITIVE", "CONSERVATIVE"])
def test_reconstruction_variable_rollout_gradient_against_device_fd(variable):
    """Kelvin-Helmholtz 48^2, 24 fixed-dt steps with CHAR-PRIMITIVE or CONSERVATIVE
    reconstruction through RK3Rollout (4 binomial slots): the gradient along a Fourier-mode
    direction against Richardson central differences of the device forward, to 1e-6."""
    from jaxfluids_amd.adjoint import kinetic_energy, rollout
    from jaxfluids_amd.workloads import fourier_velocity_direction, kelvin_helmholtz_primitives

    say(f"[{variable} kh gradient vs device fd] 48^2 x 24 steps")
    n, nsteps = 48, 24
    W0 = kelvin_helmholtz_primitives(n, seed=4, thickness=0.06)
    dom = o.Domain((n, n, 1), ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), nh=3)
    s = o.Setup(gamma=1.4, stencil="WENO5-JS", variable=variable, riemann="HLLC", signal_speed="EINFELDT",
                bcs={f: "PERIODIC" for f in FACES})
    dt = 0.5 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    dV = 1.0 / n ** 2
    st = _kh_stepper(n, dt, variable)
    W0d = dev(W0).requires_grad_(True)
    L = kinetic_energy(rollout(W0d, st, nsteps, snapshots=4), dV)
    L.backward()
    g = W0d.grad.detach()
    dd = dev(fourier_velocity_direction((n, n, 1), seed=11))

    def loss(h):
        with torch.no_grad():
            return float(kinetic_energy(rollout(dev(W0) + h * dd, st, nsteps, snapshots=0), dV))

    D = lambda h: (loss(h) - loss(-h)) / (2 * h)
    fd = (4.0 * D(0.5e-4) - D(1e-4)) / 3.0
    an = float(torch.sum(g * dd))
    say(f"  fd {fd:.15e} adjoint {an:.15e} rel {abs(fd - an) / abs(an):.2e}")
    assert abs(fd - an) <= 1e-6 * abs(an)


def test_rollout_gradient_against_oracle_fd():
    """The small rollout of the CPU chain test, now through RK3Rollout on the device."""
    from jaxfluids_amd.adjoint import kinetic_energy, rollout
    from jaxfluids_amd.workloads import kelvin_helmholtz_primitives

    say("[rollout vs oracle fd] 12^2 x 3 steps")
    n, nsteps = 12, 3
    dom = o.Domain((n, n, 1), ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), nh=3)
    s = o.Setup(gamma=1.4, stencil="WENO5-JS", variable="PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                bcs={f: "PERIODIC" for f in FACES})
    W0 = kelvin_helmholtz_primitives(n, seed=3, thickness=0.08)
    W0[3] = 0.05 * np.sin(2 * np.pi * dom.cell_centers()[0])[:, None, None] + 0.02
    dt = 0.5 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    dV = dom.dx[0] * dom.dx[1]
    st = _kh_stepper(n, dt)
    W0d = dev(W0).requires_grad_(True)
    L = kinetic_energy(rollout(W0d, st, nsteps), dV)
    L.backward()
    g = host(W0d.grad)

    def loss(Wi):
        WT, _ = _rollout_states(Wi, dom, s, dt, nsteps)
        Wt = WT[dom.interior]
        return 0.5 * dV * float(np.sum(Wt[0] * (Wt[1] ** 2 + Wt[2] ** 2 + Wt[3] ** 2)))

    assert abs(float(L) - loss(W0)) <= 1e-12 * abs(float(L))
    rng = np.random.default_rng(5)
    d = rng.standard_normal(W0.shape) * np.array([0.1, 0.05, 0.05, 0.05, 0.0])[:, None, None, None]
    D = lambda h: (loss(W0 + h * d) - loss(W0 - h * d)) / (2 * h)
    fd = (4.0 * D(0.5e-6) - D(1e-6)) / 3.0
    an = float(np.sum(g * d))
    assert abs(fd - an) <= 1e-7 * abs(an), (fd, an)


def test_kelvin_helmholtz_gradient_against_device_fd():
    """Kelvin-Helmholtz 64^2 (shear layers of thickness 0.05, resolved at this size), 40 fixed-dt
    steps: d KE / d (initial velocity) along a seeded Fourier-mode direction (k = 1..8) against
    Richardson central differences of the device forward, relative error <= 1e-6.  (With
    unresolved layers the loss is flat to 1e-11 between cells where the fields saturate, and the
    WENO weights bend it on scales far below any usable step.)"""
    from jaxfluids_amd.adjoint import kinetic_energy, rollout
    from jaxfluids_amd.workloads import fourier_velocity_direction, kelvin_helmholtz_primitives

    say("[kh gradient vs device fd] 64^2 x 40 steps")
    n, nsteps = 64, 40
    W0 = kelvin_helmholtz_primitives(n, seed=1, thickness=0.05)
    dom = o.Domain((n, n, 1), ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), nh=3)
    s = o.Setup(gamma=1.4, stencil="WENO5-JS", variable="PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                bcs={f: "PERIODIC" for f in FACES})
    dt = 0.5 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    dV = 1.0 / n ** 2
    st = _kh_stepper(n, dt)
    W0d = dev(W0).requires_grad_(True)
    L = kinetic_energy(rollout(W0d, st, nsteps, snapshots=8), dV)
    L.backward()
    g = W0d.grad.detach()
    dd = dev(fourier_velocity_direction((n, n, 1), seed=9))

    def loss(h):
        with torch.no_grad():
            return float(kinetic_energy(rollout(dev(W0) + h * dd, st, nsteps, snapshots=0), dV))

    D = lambda h: (loss(h) - loss(-h)) / (2 * h)
    fd = (4.0 * D(0.5e-4) - D(1e-4)) / 3.0
    an = float(torch.sum(g * dd))
    say(f"  fd {fd:.15e} adjoint {an:.15e} rel {abs(fd - an) / abs(an):.2e}")
    assert abs(fd - an) <= 1e-6 * abs(an)


def test_binomial_checkpointing_is_exact():
    """3 snapshot slots vs one per step: the same gradient bit for bit (the recomputed forward
    steps are the same kernels on the same inputs) and the schedule's forward-step count."""
    from jaxfluids_amd.adjoint import binomial_table, first_level_snapshots, kinetic_energy, rollout
    from jaxfluids_amd.workloads import kelvin_helmholtz_primitives

    say("[checkpointing] 32^2 x 30 steps, 3 slots vs all")
    n, nsteps = 32, 30
    W0 = kelvin_helmholtz_primitives(n, seed=2)
    grads = []
    for snaps in (nsteps, 3):
        st = _kh_stepper(n, 2e-3)
        W0d = dev(W0).requires_grad_(True)
        L = kinetic_energy(rollout(W0d, st, nsteps, snapshots=snaps), 1.0 / n ** 2)
        before = st.forward_steps
        L.backward()
        grads.append(host(W0d.grad))
        T, _ = binomial_table(nsteps, snaps)
        chain = first_level_snapshots(nsteps, snaps)
        # the forward sweep already made the first-level snapshots (the last at chain[-1])
        assert st.forward_steps - before == T[nsteps, snaps] - (chain[-1] if chain else 0)
    assert np.array_equal(grads[0], grads[1])


def test_segment_checkpointing_is_exact():
    """A checkpoint every 8 steps of 30 (the last segment 6 steps) vs one per step: the same
    gradient bit for bit, and the backward recomputes each step exactly once."""
    from jaxfluids_amd.adjoint import kinetic_energy, rollout
    from jaxfluids_amd.workloads import kelvin_helmholtz_primitives

    say("[checkpointing] 32^2 x 30 steps, segments of 8 vs all")
    n, nsteps = 32, 30
    W0 = kelvin_helmholtz_primitives(n, seed=2)
    grads = []
    for snaps, schedule in ((nsteps, "binomial"), (4, "segments")):
        st = _kh_stepper(n, 2e-3)
        W0d = dev(W0).requires_grad_(True)
        L = kinetic_energy(rollout(W0d, st, nsteps, snapshots=snaps, schedule=schedule), 1.0 / n ** 2)
        assert st.forward_steps == nsteps
        L.backward()
        grads.append(host(W0d.grad))
        if schedule == "segments":
            assert st.forward_steps == 2 * nsteps
    assert np.array_equal(grads[0], grads[1])


@pytest.mark.parametrize("case", [0, 1])
@pytest.mark.parametrize("stage", [0, 2])
def test_viscous_stage_vjp_device(case, stage):
    """The viscous / heat stage VJP on the device: dot-product identity of the dissipative part
    (VJP with minus without the transport) against oracle central differences, to 1e-9."""
    import dataclasses

    from jaxfluids_amd import _native
    from test_emu_adjoint import VISCOUS_CASES

    say(f"[viscous stage vjp] case {case} stage {stage}")
    cells, (stencil, riemann, speed), tr = VISCOUS_CASES[case]
    dom = o.Domain(cells, ((0.0, 1.0), (0.0, 1.1), (0.0, 0.9)), nh=3)
    kw = dict(gamma=1.4, stencil=stencil, variable="PRIMITIVE", riemann=riemann, signal_speed=speed,
              bcs={f: "PERIODIC" for f in FACES})
    sv, si = o.Setup(transport=tr, **kw), o.Setup(**kw)
    rng = np.random.default_rng(4000 + 10 * case + stage)
    W0 = smooth_state(dom, 90 + case)
    Uint = o.cons_from_prim(W0, 1.4)
    Un_int = Uint * (1.0 + 0.01 * rng.standard_normal(Uint.shape))
    dt = 0.2 * o.compute_dt(o.create_buffers(W0, dom, sv)[0], dom, sv)
    fv, fi = stage_map(dom, sv, stage, dt, Un_int), stage_map(dom, si, stage, dt, Un_int)
    _, W_in = fv(Uint)
    lam_int = rng.standard_normal(Uint.shape)
    lam_out = np.zeros((5,) + dom.shape)
    lam_out[dom.interior] = lam_int
    grid = _native.make_grid(dom.cells, dom.nh, dom.dx)
    lib = _native.load()
    P = _native.ptr
    g = {}
    for name, s in (("visc", sv), ("inv", si)):
        transport = None if s.transport is None else _native.make_transport(**dataclasses.asdict(s.transport))
        scheme = _native.make_scheme("PRIMITIVE", stencil, riemann, speed, "ARITHMETIC", _native.F64, 1.4,
                                     s.eps_stencil, transport)
        ws = torch.empty(lib.jfa_stage_vjp_workspace_bytes(grid, scheme), dtype=torch.uint8, device="cuda")
        Wd, ld = dev(W_in), dev(lam_out)
        li, ln = torch.zeros_like(ld), torch.zeros_like(ld)
        rc = lib.jfa_rk3_stage_vjp(P(Wd), P(ld), P(li), P(ln if stage else None), grid, scheme,
                                   _native.bc_array(s.bcs), stage, dt, P(ws), ws.numel(), _native.stream_ptr())
        _native.check(rc, "jfa_rk3_stage_vjp")
        g[name] = host(li)[dom.interior]
    d = rng.standard_normal(Uint.shape) * np.abs(Uint).mean(axis=tuple(range(1, Uint.ndim)), keepdims=True)
    fdv, fdi = richardson(fv, Uint, d, 1e-5), richardson(fi, Uint, d, 1e-5)
    dl, dr = float(np.sum(lam_int * (fdv - fdi))), float(np.sum((g["visc"] - g["inv"]) * d))
    scale = np.sum(np.abs(lam_int * (fdv - fdi)))
    assert abs(dl - dr) <= 1e-9 * scale, (dl, dr, scale)


def viscous_tgv(n, mach=0.1, reynolds=1600.0, prandtl=0.71):
    """Viscous Taylor-Green vortex on [0, 2pi)^3: the inviscid initial state at Mach 0.1 with a
    power-law viscosity mu = mu_ref (T / T0)^0.76, mu_ref = rho0 U0 L / Re (U0 = L = rho0 = 1,
    T0 = p0 / (rho0 R)), Prandtl conductivity."""
    from jaxfluids_amd import _native

    dom, W0 = o.tgv_initial(n, nh=3, mach=mach)
    T0 = (1.0 / (1.4 * mach ** 2)) / 1.0
    kw = dict(R=1.0, mu_ref=1.0 / reynolds, T_ref=T0, mu_exp=0.76, prandtl=prandtl)
    return dom, W0, o.Transport(**kw), _native.make_transport(**kw)


def test_viscous_tgv_gradient_against_device_fd():
    """Viscous Taylor-Green 16^3 (Re 1600, Pr 0.71, WENO5-Z + HLLC) with a small smooth density
    and w perturbation (no field exactly constant), 20 fixed-dt steps with 4 snapshot slots:
    d enstrophy / d (initial u, v) along a seeded Fourier-mode direction against central
    differences of the device forward, relative error <= 1e-6.  The differences converge at
    first order in h here (Riemann-solver branch switches crossed by the perturbation), so one
    small step is taken rather than a Richardson pair."""
    from jaxfluids_amd.adjoint import FixedDtStepper, enstrophy, rollout
    from jaxfluids_amd.workloads import fourier_velocity_direction

    say("[viscous tgv gradient vs device fd] 16^3 x 20 steps")
    n, nsteps = 16, 20
    dom, W0, tr, ntr = viscous_tgv(n)
    pert = fourier_velocity_direction((n, n, n), seed=21, modes=3, amplitude=1e-2)
    W0 = W0.copy()
    W0[0] += pert[1]
    W0[3] += pert[2]
    s = o.Setup(gamma=1.4, stencil="WENO5-Z", variable="PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                transport=tr)
    dt = 0.5 * o.compute_dt(o.create_buffers(W0, dom, s)[0], dom, s)
    st = FixedDtStepper(dom.cells, dom.ranges, {f: "PERIODIC" for f in FACES}, dt, nh=3, stencil="WENO5-Z",
                        riemann="HLLC", signal_speed="EINFELDT", transport=ntr)
    W0d = dev(W0).requires_grad_(True)
    L = enstrophy(rollout(W0d, st, nsteps, snapshots=4), dom.dx)
    L.backward()
    g = W0d.grad.detach()
    dd = dev(fourier_velocity_direction((n, n, n), seed=13))

    def loss(h):
        with torch.no_grad():
            return float(enstrophy(rollout(dev(W0) + h * dd, st, nsteps, snapshots=0), dom.dx))

    fd = (loss(1e-5) - loss(-1e-5)) / 2e-5
    an = float(torch.sum(g * dd))
    say(f"  fd {fd:.15e} adjoint {an:.15e} rel {abs(fd - an) / abs(an):.2e}")
    assert abs(fd - an) <= 1e-6 * abs(an)


def test_viscous_rollout_matches_oracle():
    """The stepper's viscous forward equals the oracle's fixed-dt steps (8^3, 3 steps)."""
    from jaxfluids_amd.adjoint import FixedDtStepper, rollout

    say("[viscous rollout vs oracle] 8^3 x 3 steps")
    dom, W0, tr, ntr = viscous_tgv(8)
    s = o.Setup(gamma=1.4, stencil="WENO5-Z", variable="PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT",
                transport=tr)
    W, U = o.create_buffers(W0, dom, s)
    dt = 0.5 * o.compute_dt(W, dom, s)
    for _ in range(3):
        W, U = o.rk3_step(W, U, dt, dom, s)
    st = FixedDtStepper(dom.cells, dom.ranges, {f: "PERIODIC" for f in FACES}, dt, nh=3, stencil="WENO5-Z",
                        riemann="HLLC", signal_speed="EINFELDT", transport=ntr)
    with torch.no_grad():
        WT = host(rollout(dev(W0), st, 3, snapshots=0))
    ref = W[dom.interior]
    assert np.max(np.abs(WT - ref)) <= 1e-12 * np.max(np.abs(ref))
