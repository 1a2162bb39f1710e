"""Reverse mode of fixed-dt RK3 rollouts on the GPU.

The reference differentiates a simulation with ``jax.grad`` through
``SimulationManager.do_runge_kutta_stages`` (simulation_manager.py:770-1047; feed-forward
training in simulation_manager.py:1180-1294) and relies on XLA to store or rematerialise the
forward.  Here the derivative of one RK3 stage is a hand-written HIP VJP
(``jfa_rk3_stage_vjp``), one step's VJP recomputes that step's stage states from the state at
its start, and a rollout of many steps is reversed with binomial checkpointing: with ``s``
snapshot slots the schedule that minimises recomputation is found by dynamic programming over
T(t, s) = min_m [m + T(t-m, s-1) + T(m, s)] (Griewank's binomial checkpointing), and the
reverse sweep follows it with an explicit stack.

``RK3Rollout`` wraps a rollout as a ``torch.autograd.Function``: its input is the interior
primitive state (5, Nx, Ny, Nz) at t = 0, its output the interior primitives after ``nsteps``
steps at fixed ``dt``; any torch loss of the output back-propagates into the initial state.
"""
from __future__ import annotations

import ctypes
import functools
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _native
from .config import precision


# ------------------------------------------------------------------ binomial checkpointing
@functools.lru_cache(maxsize=16)
def binomial_table(nsteps: int, snapshots: int) -> Tuple[np.ndarray, np.ndarray]:
    """T[t, s]: fewest forward steps to reverse t steps from a stored start state with s free
    snapshot slots; M[t, s]: the first advance of an optimal schedule."""
    S = max(0, min(snapshots, nsteps))
    T = np.zeros((nsteps + 1, S + 1), dtype=np.int64)
    M = np.ones((nsteps + 1, S + 1), dtype=np.int64)
    t = np.arange(nsteps + 1, dtype=np.int64)
    T[:, 0] = t * (t - 1) // 2
    T[0, :] = 0
    for s in range(1, S + 1):
        T[1:, s] = t[1:] - 1  # with a slot for every step: advance one step at a time
        for n in range(s + 2, nsteps + 1):
            m = np.arange(1, n)
            c = m + T[n - m, s - 1] + T[m, s]
            j = int(np.argmin(c))
            T[n, s] = c[j]
            M[n, s] = m[j]
    return T, M


def reverse_rollout(snaps: dict, nsteps: int, snapshots: int, lam, advance, step_vjp, store, release=None):
    """Reverse sweep of a rollout.  snaps[p] holds the state at step p (snaps[0] at least);
    advance(state, n) returns a working state n steps later, step_vjp(state, lam) the cotangent
    one step earlier, store(state) a copy to keep, release(state) (optional) takes back a stored
    state no longer needed.  At most `snapshots` states beyond snaps[0] are alive at once, and
    the forward steps taken are T[nsteps, snapshots] of binomial_table (fewer where snaps already
    holds first-level snapshots)."""
    if all(k in snaps for k in range(nsteps)):  # every step stored: no schedule to compute
        for k in range(nsteps - 1, -1, -1):
            lam = step_vjp(snaps[k], lam)
        return lam
    T, M = binomial_table(nsteps, snapshots)
    stack = [("seg", 0, nsteps, min(snapshots, nsteps))]
    while stack:
        item = stack.pop()
        if item[0] == "free":
            state = snaps.pop(item[1], None)
            if state is not None and release is not None:
                release(state)
            continue
        _, st, en, s = item
        if en - st == 1:
            lam = step_vjp(snaps[st], lam)
            continue
        if s == 0:
            for k in range(en - 1, st - 1, -1):
                lam = step_vjp(snaps[st] if k == st else advance(snaps[st], k - st), lam)
            continue
        m = int(M[en - st, s])
        if st + m not in snaps:
            snaps[st + m] = store(advance(snaps[st], m))
        stack.append(("seg", st, st + m, s))
        stack.append(("free", st + m))
        stack.append(("seg", st + m, en, s - 1))
    return lam


def first_level_snapshots(nsteps: int, snapshots: int) -> List[int]:
    """Positions the forward sweep stores so the reverse sweep starts without re-advancing."""
    T, M = binomial_table(nsteps, snapshots)
    out, p, s = [], 0, min(snapshots, nsteps)
    while nsteps - p > 1 and s > 0:
        p += int(M[nsteps - p, s])
        out.append(p)
        s -= 1
    return out


class StatePool:
    """Snapshot storage allocated once: `nslots` (W, U) buffer pairs in one device tensor, handed
    out by store() and taken back by release(), so a rollout's snapshots cost copies, not
    allocations."""

    def __init__(self, nslots: int, shape, device, fill: float) -> None:
        self.buf = torch.full((max(1, nslots), 2) + tuple(shape), fill, dtype=torch.float64, device=device)
        self._free = list(range(self.buf.shape[0] - 1, -1, -1))
        self._slot = {}

    @property
    def nslots(self) -> int:
        return self.buf.shape[0]

    def acquire(self) -> Tuple[torch.Tensor, torch.Tensor]:
        if not self._free:
            raise RuntimeError("snapshot pool exhausted")
        i = self._free.pop()
        W, U = self.buf[i, 0], self.buf[i, 1]
        self._slot[W.data_ptr()] = i
        return W, U

    def store(self, state) -> Tuple[torch.Tensor, torch.Tensor]:
        W, U = self.acquire()
        W.copy_(state[0])
        U.copy_(state[1])
        return W, U

    def release(self, state) -> None:
        i = self._slot.pop(state[0].data_ptr(), None)
        if i is not None:
            self._free.append(i)


# ------------------------------------------------------------------ torch pointwise VJPs
def prim_vjp(W: torch.Tensor, gW: torch.Tensor, gamma: float) -> torch.Tensor:
    """(dW/dU)^T gW for W = prim(U) (equation_manager.py:164-171)."""
    rho, u, v, w = W[0], W[1], W[2], W[3]
    g0, g1, g2, g3, g4 = gW[0], gW[1], gW[2], gW[3], gW[4]
    ir = 1.0 / rho
    gm1 = gamma - 1.0
    q2 = u * u + v * v + w * w
    return torch.stack([g0 - (u * g1 + v * g2 + w * g3) * ir + gm1 * 0.5 * q2 * g4,
                        g1 * ir - gm1 * u * g4, g2 * ir - gm1 * v * g4, g3 * ir - gm1 * w * g4, gm1 * g4])


def cons_vjp(W: torch.Tensor, gU: torch.Tensor, gamma: float) -> torch.Tensor:
    """(dU/dW)^T gU for U = cons(W) (equation_manager.py:93-101)."""
    rho, u, v, w = W[0], W[1], W[2], W[3]
    q2 = u * u + v * v + w * w
    return torch.stack([gU[0] + gU[1] * u + gU[2] * v + gU[3] * w + gU[4] * 0.5 * q2,
                        (gU[1] + gU[4] * u) * rho, (gU[2] + gU[4] * v) * rho, (gU[3] + gU[4] * w) * rho,
                        gU[4] / (gamma - 1.0)])


# ------------------------------------------------------------------ stepper
class FixedDtStepper:
    """Fixed-dt RK3 steps through the fused HIP stage and their VJPs, FP64, with PRIMITIVE,
    CONSERVATIVE, CHAR-PRIMITIVE or CHAR-CONSERVATIVE reconstruction (the characteristic ones
    with the ARITHMETIC or ROE frozen state, ``frozen_state``).
    ``cells``/``ranges``/``bcs`` as in the case setup; the scheme strings are the reference's
    registry names; ``transport`` (``_native.make_transport``) adds the viscous and heat fluxes
    on periodic domains."""

    def __init__(self, cells, ranges, bcs: Dict[str, str], dt: float, *, nh: int = 3, stencil: str = "WENO5-JS",
                 riemann: str = "HLLC", signal_speed: str = "EINFELDT", gamma: float = 1.4,
                 transport: Optional[_native.JfaTransport] = None, device=None,
                 variable: str = "PRIMITIVE", frozen_state: str = "ARITHMETIC") -> None:
        if variable not in ("PRIMITIVE", "CONSERVATIVE", "CHAR-PRIMITIVE", "CHAR-CONSERVATIVE"):
            raise NotImplementedError(f"reconstruction_variable {variable} has no reverse mode")
        cells = tuple(int(c) for c in cells)
        dx = [(r[1] - r[0]) / n for r, n in zip(ranges, cells)]
        scheme = _native.make_scheme(variable, stencil, riemann,
                                     signal_speed if riemann in ("HLLC", "HLL", "HLLC-LM") else None,
                                     frozen_state, _native.F64, gamma, precision.get_spatial_stencil_eps(), transport)
        full = {f: bcs.get(f, "PERIODIC") for f in _native.FACE_LOCATIONS}
        self._setup(cells, nh, gamma, _native.make_grid(cells, int(nh), dx), scheme, _native.bc_array(full), transport,
                    dt, device)

    @classmethod
    def from_simulation_manager(cls, sim, dt: float, device=None) -> "FixedDtStepper":
        """The stepper of a SimulationManager's own numerical setup (its scheme in FP64, grid and
        boundary conditions).  Raises NotImplementedError for a setup the stage VJP has no
        reverse mode for, before any step runs."""
        if precision.get_eps() > 1e-10:
            raise NotImplementedError("the reverse mode runs in double precision only")
        dom = sim.domain_information
        transport = sim.space_solver.transport
        scheme = sim.convective_flux_solver.scheme(_native.F64, transport)
        bc = sim.halo_manager.bc_array
        # the entry point checks the scheme before its buffers: NULL buffers tell the two apart
        rc = _native.load().jfa_rk3_stage_vjp(None, None, None, None, dom.grid, scheme, bc, 0, 1.0, None, 0, None)
        if rc == _native.JFA_EUNSUPPORTED:
            raise NotImplementedError("this numerical setup has no reverse mode (jfa_rk3_stage_vjp)")
        self = cls.__new__(cls)
        self._setup(dom.global_number_of_cells, dom.nh_conservatives, sim.material_manager.get_gamma(), dom.grid,
                    scheme, bc, transport, dt, device)
        return self

    def _setup(self, cells, nh, gamma, grid, scheme, bc, transport, dt, device) -> None:
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else device
        self.cells = tuple(int(c) for c in cells)
        self.nh = int(nh)
        self.gamma = float(gamma)
        self.grid = grid
        self.transport = transport  # viscous / heat fluxes (periodic domains); None: inviscid
        self.scheme = scheme
        self.bc = bc
        self.shape = (5,) + tuple(n + 2 * self.nh if n > 1 else 1 for n in self.cells)
        self.interior = (slice(None),) + tuple(slice(self.nh, -self.nh) if n > 1 else slice(None) for n in self.cells)
        lib = _native.load()
        e = lambda: torch.empty(self.shape, dtype=torch.float64, device=self.device)
        self._B = (e(), e())
        self._C = (e(), e())
        self._ws = torch.empty(max(int(lib.jfa_stage_workspace_bytes(self.grid, self.scheme)), 16), dtype=torch.uint8,
                               device=self.device)
        nb = int(lib.jfa_stage_vjp_workspace_bytes(self.grid, self.scheme))
        if nb == 0:
            raise _native.JFAError("invalid grid for the adjoint")
        self._aws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        self._lam = [e(), e(), e()]
        self._lam_out = [e(), e()]
        self._nvjp = 0
        self._work = (e(), e())
        self._pools: Dict[str, torch.Tensor] = {}
        self.tape = 0  # id of the taped rollout whose steps the rollout buffers hold
        self.dt_value = float(dt)
        self.dt = torch.full((1,), float(dt), dtype=torch.float64, device=self.device)
        self.forward_steps = 0  # forward steps executed (rollouts and recomputation)
        self.last_state = None  # (W, U) buffers the last rollout ended with

    # -- state
    def fill(self, W_int: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        """Buffers from interior primitives: eps fill, interior set, U = cons(W), face halos
        (material_fields_initializer.py:590-690)."""
        lib = _native.load()
        W = torch.full(self.shape, precision.get_eps(), dtype=torch.float64, device=self.device)
        W[self.interior] = W_int.to(device=self.device, dtype=torch.float64)
        U = torch.empty_like(W)
        _native.check(lib.jfa_cons_from_prim(_native.ptr(W), _native.ptr(U), W[0].numel(), self.gamma, _native.F64,
                                             _native.stream_ptr()), "jfa_cons_from_prim")
        _native.check(lib.jfa_halo_update(_native.ptr(W), _native.ptr(U), self.grid, self.bc, self.gamma, _native.F64,
                                          _native.stream_ptr()), "jfa_halo_update")
        return W, U

    def _stage(self, wi, ui, un, wo, uo, k) -> None:
        P = _native.ptr
        rc = _native.load().jfa_rk3_stage_fused(P(wi), P(ui), P(un), P(wo), P(uo), self.grid, self.scheme, self.bc, k,
                                                P(self.dt), 0.0, 1.0, 0.0, P(None), P(self._ws), self._ws.numel(),
                                                _native.stream_ptr())
        _native.check(rc, "jfa_rk3_stage_fused")

    def step(self, W: torch.Tensor, U: torch.Tensor, n: int = 1, keep_stages: bool = False):
        """n RK3 steps in place (simulation_manager.py:770-1047 at fixed dt).  With keep_stages
        (n = 1), returns copies of the primitives that enter stages 2 and 3, which let step_vjp
        skip recomputing them."""
        (Wb, Ub), (Wc, Uc) = self._B, self._C
        kept = None
        for _ in range(n):
            self._stage(W, U, None, Wb, Ub, 0)
            self._stage(Wb, Ub, U, Wc, Uc, 1)
            if keep_stages:
                kept = (Wb.clone(), Wc.clone())
            self._stage(Wc, Uc, U, W, U, 2)  # U_out aliases U_n (cell-local read-before-write)
        self.forward_steps += n
        return kept

    def step_into(self, W: torch.Tensor, U: torch.Tensor, W2: torch.Tensor, W3: torch.Tensor, W_out: torch.Tensor,
                  U_out: torch.Tensor) -> None:
        """One RK3 step from (W, U) into (W_out, U_out), writing the primitives that enter stages 2
        and 3 into W2 and W3 (taped forward: nothing is copied or allocated)."""
        (_, Ub), (_, Uc) = self._B, self._C
        self._stage(W, U, None, W2, Ub, 0)
        self._stage(W2, Ub, U, W3, Uc, 1)
        self._stage(W3, Uc, U, W_out, U_out, 2)
        self.forward_steps += 1

    def buffers(self, name: str, count: int) -> torch.Tensor:
        """`count` state-shaped FP64 buffers (first index), kept by the stepper and reused by later
        rollouts that need no more (grown when they do); edge halos hold the precision eps."""
        t = self._pools.get(name)
        if t is None or t.shape[0] < count:
            self._pools.pop(name, None)
            t = torch.full((max(1, count),) + self.shape, precision.get_eps(), dtype=torch.float64,
                           device=self.device)
            self._pools[name] = t
        return t[:count]

    def release_buffers(self) -> None:
        """Frees the rollout buffers kept by `buffers`."""
        self._pools.clear()

    def _vjp(self, W_in, lam_out, lam_in, lam_n, k) -> None:
        P = _native.ptr
        rc = _native.load().jfa_rk3_stage_vjp(P(W_in), P(lam_out), P(lam_in), P(lam_n), self.grid, self.scheme,
                                              self.bc, k, self.dt_value, P(self._aws), self._aws.numel(),
                                              _native.stream_ptr())
        _native.check(rc, "jfa_rk3_stage_vjp")

    def step_vjp(self, W: torch.Tensor, U: torch.Tensor, lam: torch.Tensor,
                 stages: Optional[Tuple[torch.Tensor, torch.Tensor]] = None) -> torch.Tensor:
        """dL/dU_n from lam = dL/dU_{n+1} (interior of buffer-layout tensors); (W, U) is the state at
        the start of the step and is not modified.  The primitives entering stages 2 and 3 are
        taken from `stages` (kept by step) or recomputed.  The result is one of two buffers the
        stepper reuses in turn: it stays valid until the second call after this one."""
        (Wb, Ub), (Wc, Uc) = self._B, self._C
        if stages is not None:
            Wb, Wc = stages
        else:
            self._stage(W, U, None, Wb, Ub, 0)
            self._stage(Wb, Ub, U, Wc, Uc, 1)
        l2, l1, l0 = self._lam
        for t in (l2, l1, l0):
            t.zero_()  # fill kernels (graph-safe), halos stay zero
        self._vjp(Wc, lam, l2, l0, 2)
        self._vjp(Wb, l2, l1, l0, 1)
        self._vjp(W, l1, l0, None, 0)
        # two output buffers in turn: the returned cotangent is the next call's input
        out = self._lam_out[self._nvjp % 2]
        if out is lam:
            out = self._lam_out[(self._nvjp + 1) % 2]
        self._nvjp += 1
        out.copy_(l0)
        return out

    # -- rollouts
    def reverse(self, snaps: Dict[int, Tuple[torch.Tensor, torch.Tensor]], nsteps: int, snapshots: int,
                lam: torch.Tensor, pool: Optional[StatePool] = None) -> torch.Tensor:
        """Reverse nsteps steps from the stored state snaps[0] (plus any first-level snapshots) with
        at most `snapshots` extra stored states, following the binomial schedule.  States the sweep
        stores come from `pool` (which must have a free slot for each), else from clones."""
        work = self._work

        def advance(src, n):
            work[0].copy_(src[0])
            work[1].copy_(src[1])
            self.step(work[0], work[1], n)
            return work

        store = pool.store if pool is not None else (lambda state: (state[0].clone(), state[1].clone()))
        # a stored state may carry its step's stage primitives (W, U, W2, W3)
        return reverse_rollout(snaps, nsteps, snapshots, lam, advance,
                               lambda state, l: self.step_vjp(state[0], state[1], l,
                                                              state[2:] if len(state) == 4 else None),
                               store, None if pool is None else pool.release)

    def reverse_segments(self, snaps: Dict[int, Tuple[torch.Tensor, torch.Tensor]], nsteps: int, segment: int,
                         lam: torch.Tensor, pool: Optional[StatePool] = None) -> torch.Tensor:
        """Reverse nsteps steps checkpointed every `segment` steps (snaps[0], snaps[segment], ...):
        each segment, last first, is recomputed once from its checkpoint with every step's state
        and stage primitives kept, then reversed with the three stage VJPs per step -- the
        schedule of a rematerialised segment (jax.checkpoint around `segment` steps)."""
        S = self.buffers("segment_states", 2 * (segment + 1))
        K = self.buffers("segment_stages", 2 * segment)
        for start in range(((nsteps - 1) // segment) * segment, -1, -segment):
            m = min(segment, nsteps - start)
            state = snaps.pop(start)
            S[0].copy_(state[0])
            S[1].copy_(state[1])
            if pool is not None:
                pool.release(state)
            for n in range(m):
                self.step_into(S[2 * n], S[2 * n + 1], K[2 * n], K[2 * n + 1], S[2 * n + 2], S[2 * n + 3])
            for n in range(m - 1, -1, -1):
                lam = self.step_vjp(S[2 * n], S[2 * n + 1], lam, (K[2 * n], K[2 * n + 1]))
        return lam


class RK3Rollout(torch.autograd.Function):
    """W_T = rollout(W_0): nsteps fixed-dt RK3 steps from interior primitives W_0.

    Forward stores snaps[0] and the first-level binomial snapshots (or, with segments, a
    checkpoint every ceil(nsteps / snapshots) steps); backward reverses the rollout with
    ``stepper.reverse`` (``reverse_segments``) and maps dL/dU_0 to dL/dW_0."""

    @staticmethod
    def forward(ctx, W0: torch.Tensor, stepper: FixedDtStepper, nsteps: int, snapshots: int,
                segments: bool = False, shared: bool = True, start=None) -> torch.Tensor:
        # start: the (W, U) buffers a previous rollout ended with (stepper.last_state), whose
        # interior primitives W0 are; the state continues from them instead of U = cons(W0)
        if start is not None:
            W, U = start[0].clone(), start[1].clone()
        else:
            W, U = stepper.fill(W0.detach())
        ctx.segment = 0
        # with a snapshot at every step (snapshots >= nsteps - 1) the stage primitives are kept too,
        # so the reverse sweep evaluates only the three stage VJPs per step; the steps write straight
        # into the stepper's rollout buffers (one pending rollout per stepper: shared=False keeps the
        # snapshots in a pool of this rollout's own instead)
        keep = shared and snapshots >= nsteps - 1
        pool = None
        if keep:
            S = stepper.buffers("states", 2 * (nsteps + 1))   # (W_n, U_n) of every step
            K = stepper.buffers("stages", 2 * max(nsteps, 1))  # primitives entering stages 2 and 3
            S[0].copy_(W)
            S[1].copy_(U)
            snaps = {}
            for n in range(nsteps):
                Wi, Ui, Wo, Uo = S[2 * n], S[2 * n + 1], S[2 * n + 2], S[2 * n + 3]
                stepper.step_into(Wi, Ui, K[2 * n], K[2 * n + 1], Wo, Uo)
                snaps[n] = (Wi, Ui, K[2 * n], K[2 * n + 1])
            W, U = S[2 * nsteps], S[2 * nsteps + 1]
            stepper.tape += 1
        elif segments:
            seg = -(-nsteps // max(1, snapshots))
            pool = StatePool(-(-nsteps // seg), stepper.shape, stepper.device, precision.get_eps())
            snaps = {}
            for n in range(nsteps):
                if n % seg == 0:
                    snaps[n] = pool.store((W, U))
                stepper.step(W, U)
            ctx.segment = seg
        else:
            pool = StatePool(min(snapshots, nsteps) + 1, stepper.shape, stepper.device, precision.get_eps())
            snaps = {0: pool.store((W, U))}
            chain = set(first_level_snapshots(nsteps, snapshots))
            for n in range(1, nsteps + 1):
                stepper.step(W, U)
                if n in chain:
                    snaps[n] = pool.store((W, U))
        ctx.stepper, ctx.nsteps, ctx.snapshots, ctx.snaps, ctx.pool = stepper, nsteps, snapshots, snaps, pool
        ctx.tape = stepper.tape if keep else None
        ctx.W0 = W0.detach()
        stepper.last_state = (W, U)  # the buffers the rollout ended with (shared ones are reused)
        WT = W[stepper.interior].clone()
        ctx.WT = WT
        return WT

    @staticmethod
    def backward(ctx, gWT: torch.Tensor):
        st: FixedDtStepper = ctx.stepper
        if ctx.tape is not None and ctx.tape != st.tape:
            raise RuntimeError("a later rollout that stores every step reused this stepper's rollout buffers "
                               "before this rollout's backward; use one stepper per pending rollout")
        lam = torch.zeros(st.shape, dtype=torch.float64, device=st.device)
        lam[st.interior] = prim_vjp(ctx.WT, gWT.to(torch.float64), st.gamma)
        if ctx.segment:
            lam = st.reverse_segments(ctx.snaps, ctx.nsteps, ctx.segment, lam, ctx.pool)
        else:
            lam = st.reverse(ctx.snaps, ctx.nsteps, ctx.snapshots, lam, ctx.pool)
        ctx.snaps = None
        ctx.pool = None
        gW0 = cons_vjp(ctx.W0.to(torch.float64), lam[st.interior], st.gamma)
        return gW0.to(ctx.W0.dtype), None, None, None, None, None, None


def rollout(W0: torch.Tensor, stepper: FixedDtStepper, nsteps: int, snapshots: Optional[int] = None,
            schedule: str = "binomial", shared: bool = True, start=None) -> torch.Tensor:
    """Differentiable fixed-dt rollout; snapshots defaults to storing every step's state.

    schedule "binomial": at most `snapshots` stored states beyond the initial one, the fewest
    recomputed steps for that memory (the backward recomputes the stage primitives per step);
    "segments": a checkpoint every ceil(nsteps / snapshots) steps, each segment recomputed once
    with its steps' states and stage primitives held (nsteps recomputed steps, one segment of
    full storage on top of the checkpoints).  A rollout that stores every step writes into the
    stepper's own buffers, which hold one pending rollout; ``shared=False`` gives it a snapshot
    pool of its own, so that several rollouts of one stepper can await their backward (chained
    rollouts).  ``start`` continues from the buffers (W, U) of a previous rollout
    (``stepper.last_state``, of which W0 is the interior) so that the conservative state is
    carried exactly, as the reference's scan carries its buffers; the derivative is the same
    (U = cons(W))."""
    if schedule not in ("binomial", "segments"):
        raise ValueError(f"unknown checkpoint schedule {schedule!r}")
    snapshots = int(nsteps if snapshots is None else snapshots)
    return RK3Rollout.apply(W0, stepper, int(nsteps), snapshots, schedule == "segments" and snapshots < nsteps - 1,
                            bool(shared), start)


def kinetic_energy(W: torch.Tensor, cell_volume: float) -> torch.Tensor:
    """Domain-integrated kinetic energy sum(0.5 rho |u|^2) dV of interior primitives."""
    return 0.5 * cell_volume * torch.sum(W[0] * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]))


def enstrophy(W: torch.Tensor, dx) -> torch.Tensor:
    """Domain-integrated enstrophy sum(0.5 |curl u|^2) dV of interior primitives on a periodic
    grid, with second-order central derivatives (the derivative_stencil_center CENTRAL2 form,
    stencils/derivative/deriv_center_2.py); inactive axes (extent 1) contribute no derivative."""
    u = W[1:4]

    def d(f, axis):
        if f.shape[axis] == 1:
            return torch.zeros_like(f)
        return (torch.roll(f, -1, axis) - torch.roll(f, 1, axis)) / (2.0 * dx[axis])

    wx = d(u[2], 1) - d(u[1], 2)
    wy = d(u[0], 2) - d(u[2], 0)
    wz = d(u[1], 0) - d(u[0], 1)
    dV = float(np.prod([h for h in dx]))
    return 0.5 * dV * torch.sum(wx * wx + wy * wy + wz * wz)
