"""Step loop of the convective-flux path on one GPU.

Mirrors SimulationManager (simulation_manager.py:49-1412): ``simulate`` /
``advance`` (:186-425) drive ``do_integration_step`` (:1178-1203) whose
``_do_integration_step`` (:536-668) runs ``do_runge_kutta_stages`` (:670-1077)
and then the CFL time step (:612-628).

Two stage implementations share the same arithmetic:
* ``fused=True`` (default): one ``jfa_rk3_stage_fused`` call per stage -- face
  fluxes of all axes, divergence, RK3 blend/update, prim(U) and face halos --
  on three ping-pong buffer pairs, the state ending back in the caller's buffers;
* ``fused=False``: the reference's call order through the plugin objects
  (SpaceSolver.compute_rhs -> RungeKutta3.perform_stage_integration ->
  HaloManager.perform_halo_update_material).
The time step stays on the device (float64 scalar); ``advance`` reads it back
once per step to test ``t < end_time`` as the reference does (:325-328).
Optionally one whole step is captured in a HIP graph and replayed.
"""
from __future__ import annotations

import time

import numpy as np
from typing import Optional

import torch

from . import _native
from .config import precision
from .data_types import (EulerIntegrationBuffers, IntegrationBuffers, JaxFluidsBuffers, MaterialFieldBuffers,
                         TimeControlVariables)
from .halos import HaloManager
from .space_solver import SpaceSolver
from .time_integration import compute_time_step_size


class SimulationManager:
    def __init__(self, input_manager, fused: bool = True, use_graph: bool = False) -> None:
        im = input_manager
        self.input_manager = im
        self.numerical_setup = im.numerical_setup
        self.case_setup = im.case_setup
        self.domain_information = im.domain_information
        self.material_manager = im.material_manager
        self.equation_manager = im.equation_manager
        self.equation_information = im.equation_information
        self.halo_manager = HaloManager(im.case_setup.boundary_conditions, im.domain_information,
                                        im.equation_manager, im.material_manager)
        cons = self.numerical_setup.conservatives
        dom = self.domain_information
        self.time_integrator = cons.time_integration.integrator(nh=dom.nh_conservatives,
                                                                inactive_axes=dom.inactive_axes)
        cf = cons.convective_fluxes
        self.convective_flux_solver = cf.convective_solver(
            convective_fluxes_setup=cf, positivity_setup=cons.positivity, diffuse_interface_setup=None,
            material_manager=self.material_manager, domain_information=dom, equation_manager=self.equation_manager,
            diffuse_interface_handler=None, positivity_handler=None)
        self.space_solver = SpaceSolver(dom, self.material_manager, self.equation_manager,
                                        self.convective_flux_solver, self.numerical_setup)
        self.fused = fused
        self.use_graph = use_graph
        self._bufs = None
        self._graph = None
        self._graph_key = None
        self.wall_clock_times = []

    # ------------------------------------------------------------------ buffers
    def _ensure_buffers(self, W: torch.Tensor):
        key = (W.dtype, W.device, tuple(W.shape))
        if self._bufs is not None and self._bufs["key"] == key:
            return self._bufs
        lib = _native.load()
        scheme = self.convective_flux_solver.scheme(_native.dtype_code(W), self.space_solver.transport)
        ws_bytes = lib.jfa_stage_workspace_bytes(self.domain_information.grid, scheme)
        self._bufs = {
            "key": key,
            "B": (torch.empty_like(W), torch.empty_like(W)),
            "C": (torch.empty_like(W), torch.empty_like(W)),
            "ws": torch.empty(ws_bytes, dtype=torch.uint8, device=W.device),
            "dt_ws": torch.empty(32, dtype=torch.uint8, device=W.device),
        }
        self._graph = None
        return self._bufs

    # ------------------------------------------------------------------ stages
    def _fused_stage(self, W_in, U_in, U_n, W_out, U_out, stage: int, dt: torch.Tensor,
                     dt_next: Optional[torch.Tensor] = None) -> None:
        b = self._bufs
        scheme = self.convective_flux_solver.scheme(_native.dtype_code(W_in), self.space_solver.transport)
        eps = precision.get_eps() if W_in.dtype == torch.float64 else float(torch.finfo(torch.float32).eps)
        cfl = float(self.numerical_setup.conservatives.time_integration.CFL)
        rc = _native.load().jfa_rk3_stage_fused(
            _native.ptr(W_in), _native.ptr(U_in), _native.ptr(U_n), _native.ptr(W_out), _native.ptr(U_out),
            self.domain_information.grid, scheme, self.halo_manager.bc_array, int(stage), _native.ptr(dt), 0.0,
            cfl, eps, _native.ptr(dt_next), _native.ptr(b["ws"]), b["ws"].numel(), _native.stream_ptr())
        _native.check(rc, "jfa_rk3_stage_fused")

    def do_runge_kutta_stages(self, primitives: torch.Tensor, conservatives: torch.Tensor, dt: torch.Tensor,
                              dt_next: Optional[torch.Tensor] = None):
        """Three RK3 stages with a device time step.  The fused path updates the given
        buffers in place and, if ``dt_next`` is given, writes the next CFL time step into it
        (it may be ``dt`` itself); the reference-order path returns new tensors."""
        if self.fused:
            b = self._ensure_buffers(primitives)
            (Wb, Ub), (Wc, Uc) = b["B"], b["C"]
            self._fused_stage(primitives, conservatives, None, Wb, Ub, 0, dt)
            self._fused_stage(Wb, Ub, conservatives, Wc, Uc, 1, dt)
            # U_out may alias U_n: each cell reads U_n[c] before writing U_out[c], and no cell
            # reads U_n in the halos.
            self._fused_stage(Wc, Uc, conservatives, primitives, conservatives, 2, dt, dt_next)
            return primitives, conservatives
        W, U = primitives, conservatives
        init = IntegrationBuffers(EulerIntegrationBuffers(U))
        integ = init
        for stage in range(self.time_integrator.no_stages):
            rhs, _, _ = self.space_solver.compute_rhs(U, W, physical_timestep_size=dt)
            W_new = torch.empty_like(W)
            W_new.copy_(W)
            integ = self.time_integrator.perform_stage_integration(
                integ, rhs, init, dt, stage, self.equation_information, self.domain_information,
                self.material_manager.get_gamma(), primitives_out=W_new)
            U = integ.euler_buffers.conservatives
            W, U = self.halo_manager.perform_halo_update_material(W_new, 0.0, self.halo_manager.fill_edge_halos_material,
                                                                  False, U, inplace=True)
            integ = IntegrationBuffers(EulerIntegrationBuffers(U))
        return W, U

    def compute_time_step(self, primitives: torch.Tensor, out: torch.Tensor = None) -> torch.Tensor:
        b = self._ensure_buffers(primitives) if self.fused else None
        return compute_time_step_size(primitives, None, None, None, None, self.domain_information,
                                      self.equation_information, self.material_manager, None,
                                      self.numerical_setup, out=out,
                                      workspace=None if b is None else b["dt_ws"])

    def _step_kernels(self, W, U, dt):
        fixed = self.numerical_setup.conservatives.time_integration.fixed_timestep
        if self.fused and not fixed:
            return self.do_runge_kutta_stages(W, U, dt, dt_next=dt)
        W, U = self.do_runge_kutta_stages(W, U, dt)
        if fixed:
            dt.fill_(float(fixed))
        else:
            self.compute_time_step(W, out=dt)
        return W, U

    def _capture(self, W, U, dt):
        self._ensure_buffers(W)
        key = (W.data_ptr(), U.data_ptr(), dt.data_ptr())
        if self._graph is not None and self._graph_key == key:
            return self._graph
        # warm up on a side stream (required before capture), then capture one step
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self._step_kernels(W, U, dt)
        torch.cuda.current_stream().wait_stream(s)
        self._graph, self._graph_key = g, key
        return g

    # ------------------------------------------------------------------ steps
    def do_integration_step(self, jxf_buffers: JaxFluidsBuffers) -> JaxFluidsBuffers:
        """One step with the current dt, then the next dt (simulation_manager.py:536-668)."""
        mf, tcv = jxf_buffers.material_fields, jxf_buffers.time_control_variables
        W, U = mf.primitives, mf.conservatives
        dt = tcv.physical_timestep_size
        if not isinstance(dt, torch.Tensor):
            dt = torch.full((1,), float(dt), dtype=torch.float64, device=W.device)
        dt_used = float(dt.item())
        if self.use_graph and self.fused:
            # the replayed step reads and writes dt in place
            self._capture(W, U, dt).replay()
        else:
            W, U = self._step_kernels(W, U, dt)
        tcv = tcv._replace(physical_simulation_time=tcv.physical_simulation_time + dt_used,
                           simulation_step=tcv.simulation_step + 1, physical_timestep_size=dt)
        return JaxFluidsBuffers(MaterialFieldBuffers(U, W, mf.temperature), tcv)

    def advance(self, jxf_buffers: JaxFluidsBuffers, output=None) -> JaxFluidsBuffers:
        """Host loop: while t < end_time and step < end_step (simulation_manager.py:325-328);
        ``output`` (an _Output) writes files after the steps its schedule selects."""
        tcv = jxf_buffers.time_control_variables
        self.wall_clock_times = []
        while tcv.physical_simulation_time < tcv.end_time and tcv.simulation_step < tcv.end_step:
            t0 = time.perf_counter()
            jxf_buffers = self.do_integration_step(jxf_buffers)
            tcv = jxf_buffers.time_control_variables
            self.wall_clock_times.append(time.perf_counter() - t0)
            if output is not None:
                output.after_step(jxf_buffers, self.wall_clock_times)
        torch.cuda.synchronize()
        return jxf_buffers

    def simulate(self, jxf_buffers: JaxFluidsBuffers, output_directory: Optional[str] = None) -> JaxFluidsBuffers:
        """Runs the step loop.  With ``output_directory``, writes HDF5 and XDMF files there as the
        reference's OutputWriter does (io_utils/output_writer.py:228-286; simulation_manager.py:
        258,391,411): at the start, after every step that reaches the next multiple of
        case_setup.general.save_dt (within 1e-12, 1e-6 in single precision), and at the end, then
        the XDMF time series."""
        if output_directory is None:
            return self.advance(jxf_buffers)
        out = _Output(self, output_directory)
        out.write(jxf_buffers, [])
        jxf_buffers = self.advance(jxf_buffers, out)
        out.write(jxf_buffers, self.wall_clock_times)
        out.xdmf.write_timeseries(output_directory)
        return jxf_buffers


    # ------------------------------------------------------------------ differentiable rollouts
    def _feed_forward(self, primes_init: torch.Tensor, physical_timestep_size: float, t_start: float,
                      feed_forward_setup) -> tuple:
        """SimulationManager._feed_forward (simulation_manager.py:1296-1412; feed_forward.py:28,231-523):
        outer_steps blocks of inner_steps fixed-dt RK3 steps from the interior primitives
        primes_init, differentiable with torch.autograd (stage VJPs of jfa_rk3_stage_vjp).
        Returns ({"primitives": [n_out, 5, ...]}, times [n_out]): the state at t_start when
        is_include_t0, then after every block; with is_include_halos the whole buffers (face
        halos, and the edge halos of viscous / heat fluxes, as the step fills them), else the
        interiors.  The time advances by physical_timestep_size per step, no CFL step is taken.
        Double precision only."""
        from .adjoint import rollout

        ffs = feed_forward_setup
        dt, t = float(physical_timestep_size), float(t_start)
        key = ("feed_forward", dt)
        if getattr(self, "_ff_key", None) != key:
            from .adjoint import FixedDtStepper

            self._ff_stepper = FixedDtStepper.from_simulation_manager(self, dt)
            self._ff_key = key
        st = self._ff_stepper
        W = primes_init.to(device=st.device, dtype=torch.float64)
        eps = precision.get_eps()

        def post(W):
            return self.halo_manager.halo_buffer(W, eps) if ffs.is_include_halos else W

        outs, times = ([post(W)], [t]) if ffs.is_include_t0 else ([], [])
        inner = int(ffs.inner_steps)
        state = None  # the (W, U) buffers carried from block to block
        for _ in range(int(ffs.outer_steps)):
            # each block keeps its snapshots in a pool of its own (the blocks' backward passes
            # run after all of them): one state per block, the block recomputed once, or every step
            if ffs.is_checkpoint_inner_step:
                W = rollout(W, st, inner, snapshots=1, schedule="segments", shared=False, start=state)
            else:
                W = rollout(W, st, inner, snapshots=inner, shared=False, start=state)
            state = st.last_state
            for _ in range(inner):
                t = t + dt
            outs.append(post(W))
            times.append(t)
        return {"primitives": torch.stack(outs)}, torch.tensor(times, dtype=torch.float64)

    def feed_forward(self, batch_primes_init: torch.Tensor, physical_timestep_size, t_start,
                     feed_forward_setup) -> tuple:
        """SimulationManager.feed_forward (simulation_manager.py:1248-1294): _feed_forward over a
        batch (leading axis) of initial states, time step sizes and start times; outputs stacked
        along a leading batch axis."""
        n = batch_primes_init.shape[0]
        # float64: a list of Python floats would otherwise become a float32 tensor
        dts = [float(x) for x in torch.as_tensor(physical_timestep_size, dtype=torch.float64).reshape(-1).tolist()]
        t0s = [float(x) for x in torch.as_tensor(t_start, dtype=torch.float64).reshape(-1).tolist()]
        if len(dts) != n or len(t0s) != n:
            raise ValueError("physical_timestep_size and t_start need one entry per initial state")
        sols, times = [], []
        for i in range(n):
            sol, tt = self._feed_forward(batch_primes_init[i], dts[i], t0s[i], feed_forward_setup)
            sols.append(sol["primitives"])
            times.append(tt)
        return {"primitives": torch.stack(sols)}, torch.stack(times)


class _Output:
    """The save_dt schedule of io_utils/output_writer.py:55-72,261-266 and the HDF5/XDMF writers."""

    def __init__(self, sim: "SimulationManager", directory: str) -> None:
        import os

        from .hdf5 import HDF5Writer, XDMFWriter

        os.makedirs(directory, exist_ok=True)
        self.dir = directory
        im = sim.input_manager
        gen = im.case_setup.general
        dom = sim.domain_information
        is_double = bool(precision.is_double)
        R = float(im.case_setup.material.specific_gas_constant)
        self.h5 = HDF5Writer(dom, gas_constant=R, is_double=is_double)
        self.xdmf = XDMFWriter(dom, self.h5.quantities, is_double)
        self.save_dt = gen.save_dt
        self.next_time = self.save_dt  # save_start 0: the t = 0 output is written anyway
        self.eps_time = 1e-12 if is_double else 1e-6
        self.last_time = None

    def write(self, jxf_buffers: JaxFluidsBuffers, wall_clock) -> None:
        mf, tcv = jxf_buffers.material_fields, jxf_buffers.time_control_variables
        t = float(tcv.physical_simulation_time)
        if t == self.last_time:
            return
        steps = list(wall_clock)
        ncell = float(np.prod(self.h5.domain.global_number_of_cells))
        wc = (steps[-1] if steps else 0.0, (steps[-1] / ncell) if steps else 0.0,
              float(np.mean(steps)) if steps else 0.0, float(np.mean(steps)) / ncell if steps else 0.0)
        path = self.h5.write_file(self.dir, mf.primitives, mf.conservatives, t, wc)
        self.xdmf.write_file(path, t)
        self.last_time = t

    def after_step(self, jxf_buffers: JaxFluidsBuffers, wall_clock) -> None:
        if not self.save_dt:
            return
        t = float(jxf_buffers.time_control_variables.physical_simulation_time)
        if t - self.next_time >= -self.eps_time:
            self.next_time += self.save_dt
            self.write(jxf_buffers, wall_clock)
