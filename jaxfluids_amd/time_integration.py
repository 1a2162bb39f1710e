"""TVD-RK3 time integration and the CFL time step on the HIP path.

Mirrors time_integration/time_integrator.py (TimeIntegrator :13-227),
time_integration/RK3.py (RungeKutta3 :10-62), time_integration/__init__.py
(DICT_TIME_INTEGRATION) and time_integration/time_step_size.py
(compute_time_step_size :15-157, convective part).
"""
from __future__ import annotations

from typing import Dict, Sequence

import torch

from . import _native
from .config import precision
from .data_types import EulerIntegrationBuffers, IntegrationBuffers

AXES = ("x", "y", "z")


class TimeIntegrator:
    """time_integrator.py:13-58."""

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0) -> None:
        self.no_stages = None
        self.n = int(nh) - int(offset)
        self.inactive_axes = tuple(inactive_axes)
        self.nhx, self.nhy, self.nhz = (
            slice(None) if a in self.inactive_axes else slice(self.n, -self.n) for a in AXES)
        self.timestep_multiplier = ()
        self.timestep_increment_factor = ()

    def integrate_buffer(self, buffer: torch.Tensor, rhs: torch.Tensor, timestep: float) -> torch.Tensor:
        """buffer[interior] += timestep * rhs (time_integrator.py:34-58); returns a new tensor."""
        out = buffer.clone()
        if buffer.shape[-3:] == rhs.shape[-3:]:
            out += timestep * rhs
        else:
            out[..., self.nhx, self.nhy, self.nhz] += timestep * rhs
        return out


class RungeKutta3(TimeIntegrator):
    """3rd-order TVD RK (RK3.py:10-62)."""

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0) -> None:
        super().__init__(nh, inactive_axes, offset)
        self.no_stages = 3
        self.timestep_multiplier = (1.0, 0.25, 2.0 / 3.0)
        self.timestep_increment_factor = (1.0, 0.5, 1.0)
        self.conservatives_multiplier = [(0.25, 0.75), (2.0 / 3.0, 1.0 / 3.0)]

    def prepare_buffer_for_integration(self, buffer: torch.Tensor, init_buffer: torch.Tensor,
                                       stage: int) -> torch.Tensor:
        """RK3.py:31-51."""
        c0, c1 = self.conservatives_multiplier[stage - 1]
        return c0 * buffer + c1 * init_buffer

    def integrate(self, buffer: torch.Tensor, rhs: torch.Tensor, timestep: float, stage: int) -> torch.Tensor:
        """RK3.py:53-62."""
        return self.integrate_buffer(buffer, rhs, timestep * self.timestep_multiplier[stage])

    def perform_stage_integration(self, integration_buffers: IntegrationBuffers, rhs_buffers: IntegrationBuffers,
                                  initial_stage_buffers: IntegrationBuffers, physical_timestep_size, stage: int,
                                  equation_information=None, domain_information=None, gamma: float = None,
                                  primitives_out: torch.Tensor = None) -> IntegrationBuffers:
        """Stage blend + update on the GPU (time_integrator.py:108-227, jfa_rk3_stage).

        ``physical_timestep_size`` may be a float or a device scalar tensor.  When
        ``primitives_out`` is given the kernel also writes prim(U) of the interior into it
        (simulation_manager.py:943-944), saving a separate conversion pass.
        """
        U = integration_buffers.euler_buffers.conservatives.clone()
        Un = initial_stage_buffers.euler_buffers.conservatives if stage > 0 else None
        rhs = rhs_buffers.euler_buffers.conservatives.contiguous()
        if domain_information is None or gamma is None:
            raise ValueError("domain_information and gamma are required on the HIP path")
        W = primitives_out if primitives_out is not None else torch.empty_like(U)
        dt_dev, dt_host = _dt_args(physical_timestep_size)
        scheme = _native.make_scheme("PRIMITIVE", "WENO5-JS", "HLLC", "EINFELDT", "ARITHMETIC",
                                     _native.dtype_code(U), gamma, precision.get_spatial_stencil_eps())
        rc = _native.load().jfa_rk3_stage(_native.ptr(U), _native.ptr(Un), _native.ptr(rhs), _native.ptr(W),
                                          domain_information.grid, scheme, int(stage), dt_dev, dt_host,
                                          _native.stream_ptr())
        _native.check(rc, "jfa_rk3_stage")
        return IntegrationBuffers(EulerIntegrationBuffers(U))


def _dt_args(dt):
    if isinstance(dt, torch.Tensor):
        if not dt.is_cuda or dt.dtype != torch.float64 or dt.numel() != 1:
            raise ValueError("device timestep must be a float64 scalar tensor on the GPU")
        return _native.ptr(dt), 0.0
    return _native.ptr(None), float(dt)


DICT_TIME_INTEGRATION: Dict[str, type] = {"RK3": RungeKutta3}


def compute_time_step_size(primitives: torch.Tensor, temperature=None, levelset=None, volume_fraction=None,
                           solid_temperature=None, domain_information=None, equation_information=None,
                           material_manager=None, solid_properties_manager=None, numerical_setup=None,
                           *, out: torch.Tensor = None, workspace: torch.Tensor = None) -> torch.Tensor:
    """CFL time step as a float64 device scalar (time_step_size.py:15-157, convective part):
    dt = CFL * dx_min / (max_cells sum_active(|u_i| + c) + eps); a fixed time step
    short-circuits (:40-43)."""
    ti = numerical_setup.conservatives.time_integration
    if ti.fixed_timestep:
        return torch.full((1,), float(ti.fixed_timestep), dtype=torch.float64, device=primitives.device)
    _native.require_device(primitives, "primitives")
    dt = out if out is not None else torch.empty(1, dtype=torch.float64, device=primitives.device)
    ws = workspace if workspace is not None else torch.empty(32, dtype=torch.uint8, device=primitives.device)
    eps = precision.get_eps() if primitives.dtype == torch.float64 else float(torch.finfo(torch.float32).eps)
    rc = _native.load().jfa_compute_dt(_native.ptr(primitives), domain_information.grid,
                                       material_manager.get_gamma(), float(ti.CFL), eps,
                                       _native.dtype_code(primitives), getattr(domain_information, "transport", None),
                                       _native.ptr(dt), _native.ptr(ws),
                                       _native.stream_ptr())
    _native.check(rc, "jfa_compute_dt")
    return dt
