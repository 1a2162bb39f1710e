"""Convective flux solvers on the HIP path: high-order Godunov and flux splitting.

Mirrors solvers/convective_fluxes/convective_flux_solver.py (ABC, :17-67),
high_order_godunov.py (HighOrderGodunov, :29-419) and flux_splitting_scheme.py
(FluxSplittingScheme, :21-111).  ``compute_flux_xi`` runs one HIP kernel per call
(jfa_compute_flux_xi): per face, the 6-cell window is loaded and, for Godunov, reconstructed
in the configured variables (PRIMITIVE, CONSERVATIVE, CHAR-PRIMITIVE, CHAR-CONSERVATIVE),
converted and passed to the Riemann solver; for flux splitting, projected on the frozen
state's conservative eigenvectors, split by the eigenvalues and reconstructed -- the
reference's steps without intermediate arrays.
"""
from __future__ import annotations

from typing import Dict, Optional, Tuple

import torch

from . import _native
from .config import precision
from .riemann_solvers import AUSMP, HLL, HLLC, HLLCLM, Rusanov, signal_speed_name
from .stencils import DICT_SPATIAL_RECONSTRUCTION


class ConvectiveFluxSolver:
    """convective_flux_solver.py:17-67."""

    def __init__(self, convective_fluxes_setup, material_manager, domain_information, equation_manager, **kw):
        self.convective_fluxes_setup = convective_fluxes_setup
        self.material_manager = material_manager
        self.domain_information = domain_information
        self.equation_manager = equation_manager
        self.equation_information = equation_manager.equation_information

    def compute_flux_xi(self, primitives, conservatives, axis, **kwargs):
        raise NotImplementedError


class _DeviceFaceFlux(ConvectiveFluxSolver):
    """The native scheme of a convective flux solver and its compute_flux_xi.  Subclasses set
    _names = (reconstruction variable, stencil, face flux, signal speed, frozen state) and the
    positivity settings."""

    is_interpolation_limiter = False
    limit_velocity = False
    flux_limiter = None
    flux_partition = "UNIFORM"

    def _positivity(self, positivity_setup) -> None:
        # positivity flux limiter, applied by the stage / RHS entry points (space_solver.py:530-541)
        self.flux_limiter = getattr(positivity_setup, "flux_limiter", None) or None
        self.flux_partition = getattr(positivity_setup, "flux_partition", "UNIFORM")
        self._schemes = {}

    def scheme(self, dtype_code: int, transport: Optional[_native.JfaTransport] = None) -> _native.JfaScheme:
        """Native scheme of this flux; with a transport, the fused stage and the RHS add the
        dissipative fluxes to it."""
        key = (dtype_code, None if transport is None else id(transport))
        if key not in self._schemes:
            self._schemes[key] = _native.make_scheme(
                *self._names, dtype_code, self.material_manager.get_gamma(), precision.get_spatial_stencil_eps(),
                transport, self.is_interpolation_limiter, self.limit_velocity,
                precision.get_interpolation_limiter_eps(), self.flux_limiter, self.flux_partition,
                precision.get_flux_limiter_eps())
        return self._schemes[key]

    def compute_flux_xi(self, primitives: torch.Tensor, conservatives: torch.Tensor, axis: int,
                        curvature=None, volume_fraction=None, apertures=None, ml_setup=None
                        ) -> Tuple[torch.Tensor, None, None, Optional[torch.Tensor], None]:
        """Numerical flux on every face along axis (high_order_godunov.py:117-231,
        flux_splitting_scheme.py:65-111).

        Returns (fluxes_xi, u_hat, alpha_hat, count_interpolation_limiter,
        count_interpolation_limiter_thinc) with the single-phase Nones; the limiter count is a
        device int64 scalar when the interpolation limiter is enabled, else None.
        """
        _native.require_device(primitives, "primitives")
        _native.require_device(conservatives, "conservatives")
        dom = self.domain_information
        if axis not in dom.active_axes_indices:
            raise ValueError(f"axis {axis} is inactive")
        if tuple(primitives.shape) != (5,) + dom.buffer_shape or conservatives.shape != primitives.shape:
            raise ValueError(f"buffers must have shape {(5,) + dom.buffer_shape}")
        face = list(dom.global_number_of_cells)
        face[axis] += 1
        flux = torch.empty((5, *face), dtype=primitives.dtype, device=primitives.device)
        count = torch.zeros((), dtype=torch.int64, device=primitives.device) if self.is_interpolation_limiter \
            else None
        rc = _native.load().jfa_compute_flux_xi(
            _native.ptr(primitives), _native.ptr(conservatives), dom.grid,
            self.scheme(_native.dtype_code(primitives)), int(axis), _native.ptr(flux), _native.ptr(count),
            _native.stream_ptr())
        _native.check(rc, "jfa_compute_flux_xi")
        return flux, None, None, count, None



class HighOrderGodunov(_DeviceFaceFlux):
    """Reconstruct -> convert -> Riemann solve, per face (high_order_godunov.py:29-231)."""

    def __init__(self, convective_fluxes_setup, positivity_setup=None, diffuse_interface_setup=None,
                 material_manager=None, domain_information=None, equation_manager=None,
                 diffuse_interface_handler=None, positivity_handler=None, **kwargs) -> None:
        super().__init__(convective_fluxes_setup, material_manager, domain_information, equation_manager)
        godunov = convective_fluxes_setup.godunov
        self.reconstruction_variable = godunov.reconstruction_variable
        # positivity limiter of the face states (high_order_godunov.py:67,163-176)
        self.is_interpolation_limiter = bool(getattr(positivity_setup, "is_interpolation_limiter", False))
        self.limit_velocity = bool(getattr(positivity_setup, "limit_velocity", False))
        self._positivity(positivity_setup)
        self.riemann_solver = godunov.riemann_solver(
            material_manager=material_manager, equation_manager=equation_manager,
            signal_speed=godunov.signal_speed, catum_setup=godunov.catum_setup)
        self.reconstruction_stencil = godunov.reconstruction_stencil(
            nh=domain_information.nh_conservatives, inactive_axes=domain_information.inactive_axes)
        if self.reconstruction_variable in ("CHAR-PRIMITIVE", "CHAR-CONSERVATIVE"):
            self.reconstruction_stencil.set_slices_stencil()
        self.frozen_state = godunov.frozen_state
        stencil_name = {cls: name for name, cls in DICT_SPATIAL_RECONSTRUCTION.items()}.get(
            godunov.reconstruction_stencil)
        riemann_name = {HLLC: "HLLC", HLL: "HLL", HLLCLM: "HLLC-LM", AUSMP: "AUSMP", Rusanov: "RUSANOV"}.get(
            godunov.riemann_solver)
        if stencil_name is None or riemann_name is None:
            raise NotImplementedError(
                f"built on this path: stencils {tuple(DICT_SPATIAL_RECONSTRUCTION)} with HLLC, HLLC-LM, HLL, AUSM+ or Rusanov")
        ss_name = signal_speed_name(godunov.signal_speed)
        self._names = (self.reconstruction_variable, stencil_name, riemann_name, ss_name, self.frozen_state)

class FluxSplittingScheme(_DeviceFaceFlux):
    """Characteristic flux splitting (flux_splitting_scheme.py:21-111): per face, the window's
    conservatives and physical fluxes are projected on the conservative eigenvectors of the frozen
    state of the face's two cells, split by the eigenvalues that ``flux_splitting`` names (ROE,
    CLLF or LLF; eigendecomposition.py:672-696), the positive part reconstructed from the left and
    the negative from the right, and the sum projected back.  The interpolation limiter does not
    apply; the positivity flux limiter does, in the stage and RHS entry points."""

    def __init__(self, convective_fluxes_setup, positivity_setup=None, diffuse_interface_setup=None,
                 material_manager=None, domain_information=None, equation_manager=None, **kwargs) -> None:
        super().__init__(convective_fluxes_setup, material_manager, domain_information, equation_manager)
        setup = convective_fluxes_setup.flux_splitting
        self._positivity(positivity_setup)
        self.flux_splitting = setup.flux_splitting
        if self.flux_splitting not in _native.FLUX_SPLITTING:
            # eigendecomposition.py:698-699 raises for the other names
            raise NotImplementedError(f"flux_splitting {self.flux_splitting!r}: built ROE, CLLF and LLF")
        self.reconstruction_stencil = setup.reconstruction_stencil(
            nh=domain_information.nh_conservatives, inactive_axes=domain_information.inactive_axes)
        self.reconstruction_stencil.set_slices_stencil()
        self.frozen_state = setup.frozen_state
        stencil_name = {cls: name for name, cls in DICT_SPATIAL_RECONSTRUCTION.items()}.get(setup.reconstruction_stencil)
        if stencil_name is None:
            raise NotImplementedError(f"built on this path: stencils {tuple(DICT_SPATIAL_RECONSTRUCTION)}")
        self._names = ("PRIMITIVE", stencil_name, "FLUX-SPLITTING-" + self.flux_splitting, None, self.frozen_state)


# solvers/convective_fluxes/__init__.py:7-14 (the solver built on this path)
DICT_CONVECTIVE_SOLVER: Dict[str, type] = {"GODUNOV": HighOrderGodunov, "FLUX-SPLITTING": FluxSplittingScheme}
