"""Benchmark: inviscid Taylor-Green vortex 256^3, FP64, WENO5-Z (characteristic/primitive
variables) + HLLC (Einfeldt speeds) + TVD-RK3, CFL 0.5, periodic, halo_cells 5 -- the
reference's own TGV numerical setup (examples/examples_3D/01_tgv/numerical_setup.json)
on the 256^3 grid.

One "step" = one RK3 step (3 stages + the CFL time step) over the whole grid, with the
state resident in HBM.  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps 100] [--warmup 10]

value = Mcell-updates/s per RK3 stage = N_cells * 3 * steps / wall_s / 1e6.
roofline: algorithmic bytes of the stage kernel (160 B/cell for stage 1, 200 B/cell for
stages 2-3: read W,U(,U_n), write W,U; FP64) divided by its average launch duration,
measured with HIP events on the launch stream over the timed region; peak 8 TB/s.
cpu_baseline: the C/OpenMP restatement (oracle/c) of the same step on the host cores,
timed on a bounded sample of the same workload.
ensemble: the batched-ensemble workload (64 uneven 2-D Riemann problems, one launch per RK3
stage for all cases) against plain per-case launches of the same stages.
adjoint: the Kelvin-Helmholtz 256^2 adjoint (500 fixed-dt steps): backward / forward wall time
and a finite-difference check of the gradient, with primitive and (char_primitive) characteristic
reconstruction.
adjoint_viscous: the viscous Taylor-Green 128^3 adjoint (Re 1600, 1000 fixed-dt steps, a checkpoint
every 50 steps, enstrophy loss): the same measurements, with binomial checkpointing in 20 slots
beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from jaxfluids_amd.workloads import tgv_primitives  # noqa: E402

METRIC = "Mcell-updates/s per RK3 stage at 256^3 FP64; %HBM peak; adjoint/fwd ratio"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell and stage (W and U resident, 5 fields each)
STAGE_BYTES_PER_CELL = {8: (160.0, 200.0, 200.0), 4: (80.0, 100.0, 100.0)}


def tgv_setups(n: int, nh: int, dtype: str, variable: str):
    L = 2 * np.pi
    case = {
        "general": {"case_name": f"tgv{n}", "end_time": 1e9},
        "domain": {a: {"cells": n, "range": [0.0, L]} for a in ("x", "y", "z")},
        "boundary_conditions": {f: {"type": "PERIODIC"} for f in ("east", "west", "north", "south", "top", "bottom")},
        "initial_condition": {"rho": 1.0, "u": 0.0, "v": 0.0, "w": 0.0, "p": 1.0},  # replaced below
        "material_properties": {"equation_of_state": {"model": "IdealGas", "specific_heat_ratio": 1.4,
                                                      "specific_gas_constant": 1.0}},
    }
    num = {
        "conservatives": {
            "halo_cells": nh,
            "time_integration": {"integrator": "RK3", "CFL": 0.5},
            "convective_fluxes": {"convective_solver": "GODUNOV", "godunov": {
                "riemann_solver": "HLLC", "signal_speed": "EINFELDT", "reconstruction_stencil": "WENO5-Z",
                "reconstruction_variable": variable}},
        },
        "active_physics": {"is_convective_flux": True},
        "precision": {"is_double_precision_compute": dtype == "f64"},
    }
    return case, num


def measured_stage_figure(n: int, dtype: str, variable: str, nh: int, key: str = "stage_hbm_bytes"):
    """A per-stage figure of this workload from the committed rocprofv3 summary
    (profiles/tgv256_f64_stage_kernels.md): ``stage_hbm_bytes`` (FETCH_SIZE + WRITE_SIZE, scaled
    by a plain copy) or ``stage_fp64_flop`` (SQ FP64 instruction counters); None for other
    workloads."""
    if (n, dtype, variable, nh) != (256, "f64", "CHAR-PRIMITIVE", 5):
        return None
    path = os.path.join(ROOT, "profiles", "tgv256_f64_stage_kernels.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak


def cpu_baseline(n: int, nh: int, variable: str, seconds: float):
    """Time the C/OpenMP restatement on a bounded number of steps of the same workload."""
    from oracle import c_port
    from oracle import jxf_oracle as o

    dom = o.Domain((n, n, n), ((0, 2 * np.pi),) * 3, nh=nh)
    s = o.Setup(gamma=1.4, stencil="WENO5-Z", variable=variable, riemann="HLLC", signal_speed="EINFELDT")
    prim = tgv_primitives(n)
    W = np.full((5,) + dom.shape, o.EPS64)
    W[dom.interior] = prim
    del prim
    U = o.cons_from_prim(W, s.gamma)
    st = c_port.CPUStepper(dom, s)
    dt = st.dt(W, o.EPS64)
    steps, t0 = 0, time.perf_counter()
    while True:
        st.step(W, U, dt)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    return {"value": round(n ** 3 * 3 * steps / el / 1e6, 3), "unit": "Mcell-updates/s",
            "cores": int(c_port.max_threads()), "kind": "port",
            "sample": f"{steps} RK3 step(s) of the same TGV {n}^3 workload, C/OpenMP restatement "
                      f"(oracle/c/jxf_cpu.c, gcc -O3 -march=x86-64-v3), {el:.1f} s"}


def bench_ensemble(steps: int, warmup: int, ncases: int = 64, seed: int = 0, nh: int = 5):
    """Batched ensemble of independent 2-D four-quadrant Riemann problems (sizes drawn from
    {128,192,256,384,512}^2, seeded states), WENO5-JS CHAR-PRIMITIVE + HLLC-Einfeldt + RK3,
    CFL 0.5, zero-gradient boundaries, each case with its own CFL dt.  Times the batched path
    (one launch per RK3 stage for all cases, eager and replayed from a HIP graph) against the
    same cases advanced by plain per-case launches of the fused single-domain stage."""
    import ctypes

    import torch

    from jaxfluids_amd import _native
    from jaxfluids_amd.ensemble import Ensemble2D
    from jaxfluids_amd.workloads import four_quadrant_cases, quadrant_primitives

    qc = four_quadrant_cases(ncases, seed=seed)
    zg = {f: "ZEROGRADIENT" for f in ("east", "west", "north", "south")}
    scheme = dict(stencil="WENO5-JS", variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT")
    ens = Ensemble2D([{"cells": (c.nx, c.ny), "ranges": ((0.0, 1.0), (0.0, 1.0)), "bcs": zg} for c in qc],
                     nh=nh, cfl=0.5, chunk=16, **scheme)
    ens.initialize([quadrant_primitives(c) for c in qc])
    chunk = ens.autotune(chunks=(8, 16, 32, 64), reps=2)
    ncell = ens.cells
    W0, U0, dt0 = ens.A[0].clone(), ens.A[1].clone(), ens.dt.clone()

    def restore():
        ens.A[0].copy_(W0)
        ens.A[1].copy_(U0)
        ens.dt.copy_(dt0)

    def timed(fn, n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn(n)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # batched, eager, with per-stage kernel times from HIP events on the launch stream
    stream = torch.cuda.current_stream()
    ev = []

    def eager_step_events(n):
        for _ in range(n):
            for k, (src, dst) in enumerate(((ens.A, ens.B), (ens.B, ens.C), (ens.C, ens.A))):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                ens._stage(src, dst, k, k == 2)
                e1.record(stream)
                ev.append((k, e0, e1))

    restore()
    ens.step(warmup)
    restore()
    t_eager = timed(eager_step_events, steps)
    W_batched = ens.A[0].clone()
    kern_ms = sum(e0.elapsed_time(e1) for _, e0, e1 in ev)
    byt = STAGE_BYTES_PER_CELL[8]
    kern_bytes = sum(byt[k] * ncell for k, _, _ in ev)
    # batched, graph replay
    restore()
    ens.step(1, use_graph=True)  # capture + one replay
    restore()
    t_graph = timed(lambda n: ens.step(n, use_graph=True), steps)

    # per-case: each case in its own buffers, three fused stages per step
    lib = _native.load()
    per = []
    for i in range(ncases):
        W, U = ens.view(W0, i).clone(), ens.view(U0, i).clone()  # the initial state
        grid, bcs = ens.grids[i], ens.bcs[i]
        nb = max(int(lib.jfa_stage_workspace_bytes(grid, ens.scheme)), 16)
        per.append(dict(W=W, U=U, B=(torch.empty_like(W), torch.empty_like(U)), C=(torch.empty_like(W), torch.empty_like(U)),
                        ws=torch.empty(nb, dtype=torch.uint8, device="cuda"), grid=grid, bcs=bcs,
                        dt=torch.empty(1, dtype=torch.float64, device="cuda")))
    W0s = [(q["W"].clone(), q["U"].clone()) for q in per]

    def per_restore():
        for i, q in enumerate(per):
            q["W"].copy_(W0s[i][0])
            q["U"].copy_(W0s[i][1])
            q["dt"].copy_(dt0[i:i + 1])

    def per_step(n):
        st = _native.stream_ptr()
        P = _native.ptr
        for _ in range(n):
            for q in per:
                W, U, (Wb, Ub), (Wc, Uc), dt = q["W"], q["U"], q["B"], q["C"], q["dt"]
                for k, (wi, ui, wo, uo) in enumerate(((W, U, Wb, Ub), (Wb, Ub, Wc, Uc), (Wc, Uc, W, U))):
                    rc = lib.jfa_rk3_stage_fused(P(wi), P(ui), P(U if k else None), P(wo), P(uo), q["grid"],
                                                 ens.scheme, q["bcs"], k, P(dt), 0.0, ens.cfl, ens.eps,
                                                 P(dt) if k == 2 else P(None), P(q["ws"]), q["ws"].numel(), st)
                    if rc:
                        _native.check(rc, "jfa_rk3_stage_fused")

    def case_diffs(Wpool):
        """relative max difference per case between the batched state and the per-case state"""
        out = []
        for i, q in enumerate(per):
            We = ens.view(Wpool, i)
            if not (torch.isfinite(We).all() and torch.isfinite(q["W"]).all()):
                out.append(float("nan"))
                continue
            out.append((q["W"] - We).abs().max().item() / We.abs().max().item())
        return out

    # agreement of the two paths over a short horizon from the same state
    restore()
    ens.step(10)
    per_restore()
    per_step(10)
    d10 = [d for d in case_diffs(ens.A[0]) if d == d]
    per_restore()
    per_step(warmup)
    per_restore()
    t_per = timed(per_step, steps)
    d200 = case_diffs(W_batched)
    # per-case, graph replay (the same launches captured once)
    per_restore()
    g = torch.cuda.CUDAGraph()
    s2 = torch.cuda.Stream()
    s2.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(g, stream=s2):
        per_step(1)
    torch.cuda.current_stream().wait_stream(s2)
    per_restore()
    t_per_graph = timed(lambda n: [g.replay() for _ in range(n)], steps)
    nonfinite = [i for i in range(ncases) if not bool(torch.isfinite(ens.view(W_batched, i)).all().item())]
    nonfinite_per = [i for i, q in enumerate(per) if not bool(torch.isfinite(q["W"]).all().item())]
    rate = lambda t: ncell * 3 * steps / t / 1e6

    # the same cases with the positivity limiter of the face states (positivity.is_interpolation_limiter)
    lim = Ensemble2D([{"cells": (c.nx, c.ny), "ranges": ((0.0, 1.0), (0.0, 1.0)), "bcs": zg} for c in qc],
                     nh=nh, cfl=0.5, chunk=chunk, interpolation_limiter=True, **scheme)
    lim.initialize([quadrant_primitives(c) for c in qc])
    lim.step(warmup)
    lim.initialize([quadrant_primitives(c) for c in qc])
    t_lim = timed(lambda n: lim.step(n), steps)
    limited = {"batched_eager": round(rate(t_lim), 2),
               "finite_cases": int(sum(bool(torch.isfinite(lim.state(i)[0]).all().item()) for i in range(ncases)))}
    del lim

    achieved = kern_bytes / (kern_ms * 1e-3) / 1e9
    return {
        "workload": f"{ncases} independent 2-D four-quadrant Riemann problems, sizes from "
                    "{128,192,256,384,512}^2 (seed 0), WENO5-JS CHAR-PRIMITIVE + HLLC-Einfeldt + RK3, CFL 0.5, "
                    f"zero-gradient, halo_cells {nh}, FP64, per-case CFL dt",
        "cells": ncell, "steps": steps, "chunk": chunk, "unit": "Mcell-updates/s",
        "batched_graph": round(rate(t_graph), 2), "batched_eager": round(rate(t_eager), 2),
        "per_case_eager": round(rate(t_per), 2), "per_case_graph": round(rate(t_per_graph), 2),
        "speedup_vs_per_case_eager": round(t_per / t_graph, 3),
        "speedup_vs_per_case_graph": round(t_per_graph / t_graph, 3),
        "batched_vs_per_case_max_rel_diff_10_steps": max(d10) if d10 else None,
        "cases_within_1e-10_after_all_steps": int(sum(1 for d in d200 if d == d and d <= 1e-10)),
        "finite_cases": ncases - len(nonfinite),
        "nonfinite_cases": nonfinite, "nonfinite_cases_per_case_path": nonfinite_per,
        "with_interpolation_limiter": limited,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "batch2d_stage_kernel", "avg_stage_ms": round(kern_ms / len(ev), 4)},
    }


def bench_adjoint(nsteps: int = 500, n: int = 256, slots: int = 20, fd_check: bool = True):
    """Adjoint of the 2-D Kelvin-Helmholtz workload: 256^2 periodic, WENO5-JS + HLLC-Einfeldt + RK3,
    primitive reconstruction, fixed dt (CFL 0.5 of the initial state), 500 steps, FP64; loss =
    domain-integrated kinetic energy at t_end; gradient w.r.t. the initial state (its velocity
    part is the gradient w.r.t. the seeded perturbation field).  Reports forward and backward
    wall times with every step stored and with binomial checkpointing in `slots` slots, and the
    gradient along a random velocity perturbation against Richardson central differences of the
    forward."""
    import torch

    from jaxfluids_amd.adjoint import FixedDtStepper, binomial_table, kinetic_energy, rollout
    from jaxfluids_amd.workloads import fourier_velocity_direction, kelvin_helmholtz_primitives
    from jaxfluids_amd import _native

    W0 = kelvin_helmholtz_primitives(n, seed=0)
    faces = {f: "PERIODIC" for f in _native.FACE_LOCATIONS}
    # fixed dt: CFL 0.5 of the initial state, sum over the two axes of (|u_i| + c) (time_step_size.py:103-109)
    c = np.sqrt(1.4 * W0[4] / W0[0])
    dt = 0.5 * (1.0 / n) / float(np.max(np.abs(W0[1]) + c + np.abs(W0[2]) + c))
    st = FixedDtStepper((n, n, 1), ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), faces, dt, nh=3, stencil="WENO5-JS",
                        riemann="HLLC", signal_speed="EINFELDT")
    dV = 1.0 / n ** 2
    W0d = torch.as_tensor(W0, dtype=torch.float64, device="cuda")

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out

    # warm up kernels and caches; the binomial schedule table is computed once per (steps, slots)
    binomial_table(nsteps, slots)
    with torch.no_grad():
        rollout(W0d, st, 5, 0)
    x = W0d.clone().requires_grad_(True)
    kinetic_energy(rollout(x, st, 5), dV).backward()

    with torch.no_grad():
        t_fwd, WT = timed(lambda: rollout(W0d, st, nsteps, 0))
    L0 = float(kinetic_energy(WT, dV))
    res = {}
    grads = {}
    for name, sl, schedule in (("store_every_step", nsteps, "binomial"), ("binomial", slots, "binomial"),
                               ("segments", slots, "segments")):
        x = W0d.clone().requires_grad_(True)
        t_f, out = timed(lambda: kinetic_energy(rollout(x, st, nsteps, sl, schedule), dV))
        f1 = st.forward_steps
        t_b, _ = timed(lambda: out.backward())
        res[name] = {"snapshot_slots": sl, "forward_s": round(t_f, 4), "backward_s": round(t_b, 4),
                     "backward_over_forward": round(t_b / t_fwd, 3),
                     "taped_forward_over_forward": round(t_f / t_fwd, 3),
                     "gradient_over_forward": round((t_f + t_b) / t_fwd, 3),
                     "recomputed_steps": int(st.forward_steps - f1)}
        if name == "binomial":
            T, _ = binomial_table(nsteps, sl)
            res[name]["schedule_steps"] = int(T[nsteps, sl])
        if name == "segments":
            res[name]["segment_steps"] = -(-nsteps // sl)
        grads[name] = x.grad.detach()
    same = bool(torch.equal(grads["store_every_step"], grads["binomial"])
                and torch.equal(grads["store_every_step"], grads["segments"]))
    out = {"workload": f"adjoint of 2-D Kelvin-Helmholtz {n}^2 periodic, WENO5-JS PRIMITIVE + HLLC-Einfeldt + RK3, "
                       f"fixed dt {dt:.6e}, {nsteps} steps, FP64, loss = kinetic energy at t_end",
           "forward_s": round(t_fwd, 4), "ms_per_forward_step": round(t_fwd / nsteps * 1e3, 4),
           "adjoint_over_forward": res["store_every_step"]["backward_over_forward"],
           "adjoint_schedule": "store every step",
           "gradient_over_forward": res["store_every_step"]["gradient_over_forward"],
           "modes": res, "checkpointed_gradient_identical": same, "loss": L0}
    if fd_check:
        out["fd_check"] = fd_check_along(lambda x: kinetic_energy(rollout(x, st, nsteps, 0), dV), W0d,
                                         grads["store_every_step"], fourier_velocity_direction((n, n, 1), seed=1))
    # the same rollout with CHAR-PRIMITIVE reconstruction (the TGV line's variable mode), every step stored
    sc = FixedDtStepper((n, n, 1), ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), faces, dt, nh=3, stencil="WENO5-JS",
                        riemann="HLLC", signal_speed="EINFELDT", variable="CHAR-PRIMITIVE")
    with torch.no_grad():
        rollout(W0d, sc, 5, 0)
        t_fc, _ = timed(lambda: rollout(W0d, sc, nsteps, 0))
    x = W0d.clone().requires_grad_(True)
    t_f, lo = timed(lambda: kinetic_energy(rollout(x, sc, nsteps), dV))
    t_b, _ = timed(lambda: lo.backward())
    out["char_primitive"] = {"forward_s": round(t_fc, 4), "backward_s": round(t_b, 4),
                             "backward_over_forward": round(t_b / t_fc, 3),
                             "taped_forward_over_forward": round(t_f / t_fc, 3)}
    if fd_check:
        out["char_primitive"]["fd_check"] = fd_check_along(
            lambda y: kinetic_energy(rollout(y, sc, nsteps, 0), dV), W0d, x.grad.detach(),
            fourier_velocity_direction((n, n, 1), seed=1))
    return out


def fd_check_along(loss_fn, W0d, grad, direction, steps=(1e-4, 5e-5, 2.5e-5, 1.25e-5)):
    """Central differences of loss_fn along `direction` at several steps, against the adjoint
    gradient's projection.  The rollouts' losses are only piecewise smooth (Riemann-solver branch
    switches, WENO weights near flat data), so the differences are reported at every step with
    their relative spread; rel_error is that of the smallest step."""
    import torch

    dd = torch.as_tensor(direction, dtype=torch.float64, device=W0d.device)

    def loss(h):
        with torch.no_grad():
            return float(loss_fn(W0d + h * dd))

    D = {h: (loss(h) - loss(-h)) / (2 * h) for h in steps}
    an = float(torch.sum(grad * dd))
    vals = list(D.values())
    return {"direction": "seeded (u, v) field on Fourier modes k = 1..8, amplitude 1e-2",
            "central_fd": {f"{h:g}": v for h, v in D.items()}, "adjoint": an,
            "rel_error": abs(D[min(steps)] - an) / abs(an),
            "fd_spread": (max(vals) - min(vals)) / abs(an)}


def bench_adjoint_viscous(nsteps: int = 1000, n: int = 128, slots: int = 20, store_steps: int = 100,
                          fd_check: bool = True):
    """Adjoint of the 3-D viscous Taylor-Green vortex: 128^3 periodic [0, 2pi)^3, Re 1600, Mach 0.1,
    Prandtl 0.71, power-law viscosity mu = mu_ref (T / T0)^0.76, WENO5-Z PRIMITIVE + HLLC-Einfeldt
    + RK3 with second-order central viscous and heat fluxes, fixed dt (CFL 0.5 of the initial
    state), 1000 steps, FP64; loss = enstrophy at t_end; gradient w.r.t. the initial velocity
    field.  Schedules: a checkpoint every 50 steps (20 checkpoints), each 50-step segment
    recomputed once with a slot per step (the headline adjoint_over_forward); binomial
    checkpointing in 20 snapshot slots; and every step stored, measured over the first
    `store_steps` steps (a full store of 1000 steps would hold 190 GB)."""
    import torch

    from jaxfluids_amd import _native
    from jaxfluids_amd.adjoint import FixedDtStepper, binomial_table, enstrophy, rollout
    from jaxfluids_amd.workloads import fourier_velocity_direction

    slots = min(slots, nsteps)
    store_steps = min(store_steps, nsteps)
    W0 = tgv_primitives(n)
    gamma, mach = 1.4, 0.1
    p0 = 1.0 / (gamma * mach ** 2)
    transport = _native.make_transport(R=1.0, mu_ref=1.0 / 1600.0, T_ref=p0, mu_exp=0.76, prandtl=0.71)
    h = 2 * np.pi / n
    c = np.sqrt(gamma * W0[4] / W0[0])
    conv = h / float(np.max(np.abs(W0[1]) + np.abs(W0[2]) + np.abs(W0[3]) + 3 * c))
    mu_max = float(np.max(1.0 / 1600.0 * (W0[4] / W0[0] / p0) ** 0.76 / W0[0]))
    dt = 0.5 * min(conv, 3.0 / 14.0 * h * h / mu_max)  # time_step_size.py:103-121 at t = 0
    faces = {f: "PERIODIC" for f in _native.FACE_LOCATIONS}
    st = FixedDtStepper((n, n, n), ((0.0, 2 * np.pi),) * 3, faces, dt, nh=3, stencil="WENO5-Z", riemann="HLLC",
                        signal_speed="EINFELDT", transport=transport)
    dx = (h, h, h)
    W0d = torch.as_tensor(W0, dtype=torch.float64, device="cuda")

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    binomial_table(nsteps, slots)  # computed once per (steps, slots)
    with torch.no_grad():
        rollout(W0d, st, 3, 0)
    x = W0d.clone().requires_grad_(True)
    enstrophy(rollout(x, st, 3), dx).backward()

    res = {}
    with torch.no_grad():
        t_fwd_s, _ = timed(lambda: rollout(W0d, st, store_steps, 0))
    x = W0d.clone().requires_grad_(True)
    t_f, lo = timed(lambda: enstrophy(rollout(x, st, store_steps, store_steps), dx))
    t_b, _ = timed(lambda: lo.backward())
    res["store_every_step"] = {"steps": store_steps, "forward_s": round(t_fwd_s, 4), "backward_s": round(t_b, 4),
                               "backward_over_forward": round(t_b / t_fwd_s, 3),
                               "taped_forward_over_forward": round(t_f / t_fwd_s, 3),
                               "gradient_over_forward": round((t_f + t_b) / t_fwd_s, 3)}
    del x, lo
    st.release_buffers()
    torch.cuda.empty_cache()
    with torch.no_grad():
        t_fwd, WT = timed(lambda: rollout(W0d, st, nsteps, 0))
    L0 = float(enstrophy(WT, dx))
    grads = {}
    for name, schedule in (("segments", "segments"), ("binomial", "binomial")):
        x = W0d.clone().requires_grad_(True)
        t_f, lo = timed(lambda: enstrophy(rollout(x, st, nsteps, slots, schedule), dx))
        f1 = st.forward_steps
        t_b, _ = timed(lambda: lo.backward())
        res[name] = {"steps": nsteps, "snapshot_slots": slots, "forward_s": round(t_f, 4),
                     "backward_s": round(t_b, 4), "backward_over_forward": round(t_b / t_fwd, 3),
                     "taped_forward_over_forward": round(t_f / t_fwd, 3),
                     "gradient_over_forward": round((t_f + t_b) / t_fwd, 3),
                     "recomputed_steps": int(st.forward_steps - f1)}
        if name == "segments":
            res[name]["segment_steps"] = -(-nsteps // slots)
        else:
            T, _ = binomial_table(nsteps, slots)
            res[name]["schedule_steps"] = int(T[nsteps, slots])
        grads[name] = x.grad.detach()
        del x, lo
        st.release_buffers()
        torch.cuda.empty_cache()
    g = grads["segments"]
    out = {"workload": f"adjoint of 3-D viscous Taylor-Green {n}^3 periodic, Re 1600, Mach 0.1, Pr 0.71, power-law "
                       f"viscosity (0.76), WENO5-Z PRIMITIVE + HLLC-Einfeldt + RK3 + CENTRAL2 viscous/heat fluxes, "
                       f"fixed dt {dt:.6e}, {nsteps} steps, FP64, loss = enstrophy at t_end",
           "forward_s": round(t_fwd, 4), "ms_per_forward_step": round(t_fwd / nsteps * 1e3, 4),
           "forward_mcell_updates_per_s": round(n ** 3 * 3 * nsteps / t_fwd / 1e6, 1),
           "adjoint_over_forward": res["segments"]["backward_over_forward"],
           "adjoint_schedule": f"checkpoint every {res['segments']['segment_steps']} steps ({slots} checkpoints), "
                               "each segment recomputed once with a slot per step",
           "gradient_over_forward": res["segments"]["gradient_over_forward"],
           "checkpointed_gradient_identical": bool(torch.equal(grads["segments"], grads["binomial"])),
           "modes": res, "loss": L0, "grad_finite": bool(torch.isfinite(g).all().item())}
    if fd_check:
        out["fd_check"] = fd_check_along(lambda y: enstrophy(rollout(y, st, nsteps, 0), dx), W0d, g,
                                         fourier_velocity_direction((n, n, n), seed=1),
                                         steps=(4e-5, 2e-5, 1e-5))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nh", type=int, default=5)
    ap.add_argument("--dtype", choices=("f64", "f32"), default="f64")
    ap.add_argument("--variable", default="CHAR-PRIMITIVE")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("tgv", "ensemble", "adjoint", "adjoint-viscous", "all"), default="all",
                    help="tgv: the headline line only; ensemble / adjoint: only that measurement; "
                         "all: headline line with the ensemble and adjoint measurements attached")
    ap.add_argument("--adjoint-steps", type=int, default=500)
    ap.add_argument("--viscous-adjoint-steps", type=int, default=1000)
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device-to-device copy of GIB GiB (5 launches): the reference "
                         "kernel whose FETCH_SIZE/WRITE_SIZE scale the traffic counters of the stage kernels")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU domain decomposition is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch

    from jaxfluids_amd import InitializationManager, InputManager, SimulationManager

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.calibrate_copy > 0:
        n = int(args.calibrate_copy * (1 << 30)) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n * 8, "launches": 5}), flush=True)
        return
    if args.workload == "ensemble":
        print(json.dumps(bench_ensemble(args.ensemble_steps, 5)), flush=True)
        return
    if args.workload == "adjoint":
        print(json.dumps(bench_adjoint(args.adjoint_steps)), flush=True)
        return
    if args.workload == "adjoint-viscous":
        print(json.dumps(bench_adjoint_viscous(args.viscous_adjoint_steps)), flush=True)
        return
    case, num = tgv_setups(args.n, args.nh, args.dtype, args.variable)
    im = InputManager(case, num)
    bufs = InitializationManager(im).initialization(tgv_primitives(args.n))
    sim = SimulationManager(im)
    W, U = bufs.material_fields.primitives, bufs.material_fields.conservatives
    dt = bufs.time_control_variables.physical_timestep_size
    sim._ensure_buffers(W)
    (Wb, Ub), (Wc, Uc) = sim._bufs["B"], sim._bufs["C"]
    stream = torch.cuda.current_stream()

    def step(events=None):
        # one RK3 step: three fused stages; the last also writes the next CFL dt in place
        seq = ((W, U, None, Wb, Ub), (Wb, Ub, U, Wc, Uc), (Wc, Uc, U, W, U))
        for k, (wi, ui, un, wo, uo) in enumerate(seq):
            if events is not None:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
            sim._fused_stage(wi, ui, un, wo, uo, k, dt, dt if k == 2 else None)
            if events is not None:
                e1.record(stream)
                events.append((k, e0, e1))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    events = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(events)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    ncell = args.n ** 3
    esize = 8 if args.dtype == "f64" else 4
    byt = STAGE_BYTES_PER_CELL[esize]
    tot_ms = sum(e0.elapsed_time(e1) for _, e0, e1 in events)
    tot_bytes = sum(byt[k] * ncell for k, _, _ in events)
    achieved = tot_bytes / (tot_ms * 1e-3) / 1e9
    value = ncell * 3 * args.steps / wall / 1e6
    W_h = W[(slice(None),) + im.domain_information.domain_slices_conservatives]
    finite = bool(torch.isfinite(W_h).all().item())

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": f"inviscid Taylor-Green vortex {args.n}^3 periodic, WENO5-Z {args.variable} + "
                               f"HLLC-Einfeldt + RK3, CFL 0.5, halo_cells {args.nh}",
                   "cells": ncell, "halo_cells": args.nh, "storage": args.dtype, "finite": finite},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_stage_figure(args.n, args.dtype, args.variable, args.nh),
                     "kernel": "RK3 stage (jfa_rk3_stage_fused: x/y marching sweeps + z sweep with update)",
                     "avg_stage_ms": round(tot_ms / len(events), 4),
                     "alg_bytes_per_cell_stage": round(sum(byt) / 3, 2)},
    }
    flop = measured_stage_figure(args.n, args.dtype, args.variable, args.nh, "stage_fp64_flop")
    if flop:
        # the stage's actual bound: FP64 arithmetic issue (flop count from the SQ counters of the
        # committed profile, time from this run's events)
        tflops = flop / (tot_ms / len(events) * 1e-3) / 1e12
        out["roofline"]["fp64"] = {"achieved": round(tflops, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                   "frac": round(tflops / FP64_PEAK_TFLOPS, 4), "flop_per_stage": flop}
    del sim, bufs, W, U, Wb, Ub, Wc, Uc, W_h
    torch.cuda.empty_cache()
    if args.workload == "all":
        out["ensemble"] = bench_ensemble(args.ensemble_steps, 5)
        torch.cuda.empty_cache()
        out["adjoint"] = bench_adjoint(args.adjoint_steps)
        torch.cuda.empty_cache()
        out["adjoint_viscous"] = bench_adjoint_viscous(args.viscous_adjoint_steps)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.n, args.nh, args.variable if args.variable != "CHAR-CONSERVATIVE"
                                           else "CHAR-PRIMITIVE", args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
