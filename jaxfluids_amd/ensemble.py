"""Ensembles of independent 2-D domains advanced together on one GPU.

The reference advances one domain per SimulationManager (simulation_manager.py:670-1077)
and batches only equal-shaped initial conditions through ``jax.vmap``
(simulation_manager.py:1248-1294).  Here domains of *different* size, boundary types and
time step share pooled device buffers and advance through one HIP launch per RK3 stage
for all cases (``jfa_batch2d_stage``): every (case, x-chunk, 64-cell y-tile) item is one
wavefront and the hardware spreads the blocks over the chip, so small and large cases
load-balance together.  Every case follows the single-domain stage arithmetic exactly;
each keeps its own CFL time step in device memory.  One whole step can be captured in a
HIP graph and replayed.

``autotune`` times the candidate x-chunk lengths on the actual ensemble and keeps the
fastest (cached per size signature for the process).
"""
from __future__ import annotations

import ctypes
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native
from .config import precision

FACES_2D = ("east", "west", "north", "south")
_TUNE_CACHE: Dict[tuple, int] = {}


class Ensemble2D:
    """Pooled (W, U) state of independent 2-D cases and their batched RK3 stepping.

    cases: list of dicts with keys ``cells`` (nx, ny), ``ranges`` ((x0, x1), (y0, y1)),
    ``bcs`` {east, west, north, south: PERIODIC|ZEROGRADIENT|SYMMETRY}; the scheme
    (reconstruction, Riemann solver, gamma, CFL, halo cells) is shared.
    """

    def __init__(self, cases: Sequence[dict], *, stencil: str = "WENO5-JS", variable: str = "PRIMITIVE",
                 riemann: str = "HLLC", signal_speed: str = "EINFELDT", frozen_state: str = "ARITHMETIC",
                 gamma: float = 1.4, cfl: float = 0.5, nh: int = 3, dtype=torch.float64, chunk: int = 32,
                 interpolation_limiter: bool = False, limit_velocity: bool = False,
                 flux_limiter: Optional[str] = None, flux_partition: str = "UNIFORM", device=None) -> None:
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else device
        self.dtype = dtype
        self.gamma, self.cfl, self.nh = float(gamma), float(cfl), int(nh)
        # positivity limiter of the face states: epsilons of the storage precision
        lim_eps = (1e-12, 1e-10) if dtype == torch.float64 else (1e-6, 1e-6)
        self.scheme = _native.make_scheme(variable, stencil, riemann, signal_speed, frozen_state,
                                          _native.dtype_code(torch.empty(0, dtype=dtype)), gamma,
                                          precision.get_spatial_stencil_eps(), None, interpolation_limiter,
                                          limit_velocity, lim_eps, flux_limiter, flux_partition, lim_eps)
        self.eps = float(torch.finfo(dtype).eps)
        self.ncases = len(cases)
        self.table = (_native.JfaCase2D * self.ncases)()
        self.shapes: List[Tuple[int, int, int, int]] = []
        self.grids = []
        self.bcs = []
        off = 0
        for c, cs in enumerate(cases):
            nx, ny = (int(v) for v in cs["cells"])
            (x0, x1), (y0, y1) = cs.get("ranges", ((0.0, 1.0), (0.0, 1.0)))
            dx, dy = (x1 - x0) / nx, (y1 - y0) / ny
            e = self.table[c]
            e.n[0], e.n[1], e.nh = nx, ny, self.nh
            bcs = cs.get("bcs", {f: "ZEROGRADIENT" for f in FACES_2D})
            for i, f in enumerate(FACES_2D):
                e.bc[i] = _native.BC[bcs[f]]
            e.dx[0], e.dx[1] = dx, dy
            e.offset = off
            B0, B1 = nx + 2 * self.nh, ny + 2 * self.nh
            self.shapes.append((off, B0, B1, 5 * B0 * B1))
            self.grids.append(_native.make_grid((nx, ny, 1), self.nh, (dx, dy, 1.0)))
            self.bcs.append(_native.bc_array({**{f: bcs[f] for f in FACES_2D}, "top": "PERIODIC", "bottom": "PERIODIC"}))
            off += 5 * B0 * B1
        self.total = off
        mk = lambda: torch.empty(self.total, dtype=dtype, device=self.device)
        self.A = (mk(), mk())
        self.B = (mk(), mk())
        self.C = (mk(), mk())
        self.dt = torch.zeros(self.ncases, dtype=torch.float64, device=self.device)
        self._plan(int(chunk))
        self._graph = None
        self.steps_done = 0

    # ------------------------------------------------------------------ plan / views
    def _plan(self, chunk: int) -> None:
        lib = _native.load()
        nbytes = lib.jfa_batch2d_workspace_bytes(ctypes.byref(self.table), self.ncases, chunk)
        if nbytes == 0:
            raise _native.JFAError("invalid ensemble table")
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        rc = lib.jfa_batch2d_plan(ctypes.byref(self.table), self.ncases, chunk, _native.ptr(self.ws), nbytes,
                                  _native.stream_ptr())
        _native.check(rc, "jfa_batch2d_plan")
        self.chunk = chunk
        self._graph = None

    def view(self, pool: torch.Tensor, c: int) -> torch.Tensor:
        off, B0, B1, n = self.shapes[c]
        return pool[off:off + n].view(5, B0, B1, 1)

    def state(self, c: int) -> Tuple[torch.Tensor, torch.Tensor]:
        """(W, U) of case c, views into the pooled state (5, nx+2nh, ny+2nh, 1)."""
        return self.view(self.A[0], c), self.view(self.A[1], c)

    # ------------------------------------------------------------------ initial state
    def initialize(self, primitives: Sequence[np.ndarray]) -> None:
        """Per case (5, nx, ny, 1) primitives: eps-filled buffers, interior set, U = cons(W),
        face halos, first CFL dt (material_fields_initializer.py:590-690)."""
        lib = _native.load()
        eps_fill = precision.get_eps()
        ws32 = torch.empty(32, dtype=torch.uint8, device=self.device)
        for c, prim in enumerate(primitives):
            W, U = self.state(c)
            W.fill_(eps_fill)
            W[:, self.nh:-self.nh, self.nh:-self.nh, :] = torch.as_tensor(prim, dtype=self.dtype, device=self.device)
            rc = lib.jfa_cons_from_prim(_native.ptr(W), _native.ptr(U), W[0].numel(), self.gamma,
                                        self.scheme.dtype, _native.stream_ptr())
            _native.check(rc, "jfa_cons_from_prim")
            rc = lib.jfa_halo_update(_native.ptr(W), _native.ptr(U), self.grids[c], self.bcs[c], self.gamma,
                                     self.scheme.dtype, _native.stream_ptr())
            _native.check(rc, "jfa_halo_update")
            rc = lib.jfa_compute_dt(_native.ptr(W), self.grids[c], self.gamma, self.cfl, self.eps, self.scheme.dtype,
                                    None, _native.ptr(self.dt[c:c + 1]), _native.ptr(ws32), _native.stream_ptr())
            _native.check(rc, "jfa_compute_dt")
        self._graph = None
        self.steps_done = 0

    # ------------------------------------------------------------------ stepping
    def _stage(self, src, dst, stage: int, dt_next: bool) -> None:
        Un = self.A[1] if stage > 0 else None
        rc = _native.load().jfa_batch2d_stage(
            _native.ptr(src[0]), _native.ptr(src[1]), _native.ptr(Un), _native.ptr(dst[0]), _native.ptr(dst[1]),
            ctypes.byref(self.table), self.ncases, self.chunk, self.scheme, int(stage), _native.ptr(self.dt),
            self.cfl, self.eps, _native.ptr(self.dt) if dt_next else _native.ptr(None),
            _native.ptr(self.ws), self.ws.numel(), _native.stream_ptr())
        _native.check(rc, "jfa_batch2d_stage")

    def _step_kernels(self) -> None:
        self._stage(self.A, self.B, 0, False)
        self._stage(self.B, self.C, 1, False)
        self._stage(self.C, self.A, 2, True)   # U_out aliases U_n (cell-local read-before-write)

    def step(self, nsteps: int = 1, use_graph: bool = False) -> None:
        """Advance every case by nsteps RK3 steps, each with its own CFL dt."""
        if use_graph:
            g = self._capture()
            for _ in range(nsteps):
                g.replay()
        else:
            for _ in range(nsteps):
                self._step_kernels()
        self.steps_done += nsteps

    def _capture(self):
        if self._graph is None:
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                self._step_kernels()
            torch.cuda.current_stream().wait_stream(s)
            self._graph = g
        return self._graph

    # ------------------------------------------------------------------ tuning
    def signature(self) -> tuple:
        return tuple((e.n[0], e.n[1]) for e in self.table) + (self.scheme.recon_var, self.scheme.stencil,
                                                              self.scheme.riemann, self.scheme.dtype)

    def autotune(self, chunks: Sequence[int] = (8, 16, 24, 32, 48, 64), reps: int = 3) -> int:
        """Time one RK3 step for each x-chunk candidate on a copy of the state and keep the
        fastest.  The state itself is restored afterwards."""
        key = self.signature()
        if key in _TUNE_CACHE:
            ch = _TUNE_CACHE[key]
            if ch != self.chunk:
                self._plan(ch)
            return ch
        saved = (self.A[0].clone(), self.A[1].clone(), self.dt.clone())
        best = None
        for ch in chunks:
            self._plan(ch)
            self._step_kernels()  # warm up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                self._step_kernels()
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / reps
            if best is None or t < best[0]:
                best = (t, ch)
        self.A[0].copy_(saved[0])
        self.A[1].copy_(saved[1])
        self.dt.copy_(saved[2])
        ch = best[1]
        self._plan(ch)
        _TUNE_CACHE[key] = ch
        return ch

    @property
    def cells(self) -> int:
        return int(sum(e.n[0] * e.n[1] for e in self.table))
