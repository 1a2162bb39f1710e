"""Right-hand side of the Euler equations on the HIP path.

Mirrors solvers/space_solver.py (SpaceSolver.compute_rhs :151-453 and
compute_rhs_xi :456-674) for the convective flux:
rhs = 0.0 + rhs_x + rhs_y + rhs_z with rhs_xi = (1/dx) * (F[:-1] - F[1:]).
"""
from __future__ import annotations

from typing import Tuple

import torch

from . import _native
from .data_types import DiscretizationCounter, EulerIntegrationBuffers, IntegrationBuffers, PositivityCounter


class SpaceSolver:
    """space_solver.py:43-674 (convective flux only)."""

    def __init__(self, domain_information, material_manager, equation_manager, convective_flux_solver,
                 numerical_setup=None, **kwargs) -> None:
        self.domain_information = domain_information
        self.material_manager = material_manager
        self.equation_manager = equation_manager
        self.convective_flux_solver = convective_flux_solver
        self.numerical_setup = numerical_setup
        ap = getattr(numerical_setup, "active_physics", None)
        if ap is not None and ap.is_volume_force:
            raise NotImplementedError("volume forces are not on this path")
        # viscous / heat fluxes (source_term_solver.py:188-350) when the material carries a transport
        self.transport = getattr(material_manager, "transport", None)
        self._ws = {}

    def _workspace(self, like: torch.Tensor) -> torch.Tensor:
        key = (like.dtype, like.device)
        if key not in self._ws:
            nbytes = _native.load().jfa_rhs_workspace_bytes(self.domain_information.grid, _native.dtype_code(like))
            self._ws[key] = torch.empty(nbytes, dtype=torch.uint8, device=like.device)
        return self._ws[key]

    def compute_rhs(self, conservatives: torch.Tensor, primitives: torch.Tensor, temperature=None,
                    physical_simulation_time: float = 0.0, physical_timestep_size: float = 0.0, *args,
                    **kwargs) -> Tuple[IntegrationBuffers, PositivityCounter, DiscretizationCounter]:
        """RHS of the interior cells, (5, Nx, Ny, Nz): convective flux (positivity-limited with
        physical_timestep_size when the flux limiter is on), plus the viscous and heat fluxes when
        the material has a transport."""
        _native.require_device(primitives, "primitives")
        _native.require_device(conservatives, "conservatives")
        dom = self.domain_information
        rhs = torch.empty((5, *dom.global_number_of_cells), dtype=primitives.dtype, device=primitives.device)
        ws = self._workspace(primitives)
        scheme = self.convective_flux_solver.scheme(_native.dtype_code(primitives), self.transport)
        dt = 0.0
        if scheme.flux_limiter:
            dt = float(physical_timestep_size.reshape(-1)[0]) if torch.is_tensor(physical_timestep_size) \
                else float(physical_timestep_size)
        rc = _native.load().jfa_compute_rhs(_native.ptr(primitives), _native.ptr(conservatives), dom.grid, scheme,
                                           dt, _native.ptr(rhs), _native.ptr(ws), ws.numel(), _native.stream_ptr())
        _native.check(rc, "jfa_compute_rhs")
        return (IntegrationBuffers(EulerIntegrationBuffers(rhs)), PositivityCounter(), DiscretizationCounter())
