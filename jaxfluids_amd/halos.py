"""Face- and edge-halo fill on the HIP path.

Mirrors halos/halo_manager.py (HaloManager.perform_halo_update_material
:146-234) -> halos/outer/material.py (face_halo_update :94-287, miscellaneous
:868-894; edge_halo_update :289-380) for PERIODIC, ZEROGRADIENT and SYMMETRY
boundaries; the conservative halo is recomputed from the primitive halo
(:248-250).  Edge halos are filled when the material has viscous or heat fluxes
(:125-129), whose central stencils read them on non-periodic domains.
"""
from __future__ import annotations

from typing import Dict

import torch

from . import _native

SUPPORTED = ("PERIODIC", "ZEROGRADIENT", "SYMMETRY")


class HaloManager:
    """halo_manager.py:24-234 (outer face halos of one block)."""

    def __init__(self, boundary_conditions: Dict[str, str], domain_information, equation_manager=None,
                 material_manager=None, **kwargs) -> None:
        self.domain_information = domain_information
        self.material_manager = material_manager
        self.boundary_conditions = dict(boundary_conditions)
        for loc in domain_information.active_face_locations:
            bc = self.boundary_conditions.get(loc)
            if bc not in SUPPORTED:
                raise NotImplementedError(f"boundary type {bc} at {loc} is not on the convective-flux path")
        for axis in domain_information.active_axes_indices:
            lo, hi = ("west", "east") if axis == 0 else (("south", "north") if axis == 1 else ("bottom", "top"))
            if (self.boundary_conditions[lo] == "PERIODIC") != (self.boundary_conditions[hi] == "PERIODIC"):
                raise AssertionError(f"PERIODIC must be set on both {lo} and {hi}")
        # faces of inactive axes are never filled; they carry a placeholder code
        self.bc_array = _native.bc_array(
            {loc: (self.boundary_conditions.get(loc) if loc in domain_information.active_face_locations
                   else "PERIODIC") for loc in _native.FACE_LOCATIONS})
        # dissipative cell-face fluxes read edge halos (halo_manager.py:125-129)
        self.fill_edge_halos_material = getattr(material_manager, "transport", None) is not None
        self.fill_vertex_halos_material = False   # levelset / diffuse-interface only

    def perform_halo_update_material(self, primitives: torch.Tensor, physical_simulation_time: float = 0.0,
                                     fill_edge_halos: bool = False, fill_vertex_halos: bool = False,
                                     conservatives: torch.Tensor = None, fill_face_halos: bool = True,
                                     ml_setup=None, inplace: bool = False):
        """Returns (W, U) when conservatives are given, else W (halo_manager.py:146-234).

        The reference returns new arrays; with ``inplace=True`` the given tensors are
        updated instead (no copy)."""
        if fill_vertex_halos:
            raise NotImplementedError("vertex halos are only needed by levelset and diffuse-interface fluxes")
        _native.require_device(primitives, "primitives")
        W = primitives if inplace else primitives.clone()
        U = None
        if conservatives is not None:
            _native.require_device(conservatives, "conservatives")
            U = conservatives if inplace else conservatives.clone()
        if fill_face_halos:
            rc = _native.load().jfa_halo_update(_native.ptr(W), _native.ptr(U), self.domain_information.grid,
                                                self.bc_array, self.material_manager.get_gamma(),
                                                _native.dtype_code(W), _native.stream_ptr())
            _native.check(rc, "jfa_halo_update")
        if fill_edge_halos:  # after the face halos they are built from (halo_manager.py:183-186)
            rc = _native.load().jfa_edge_halo_update(_native.ptr(W), _native.ptr(U), self.domain_information.grid,
                                                     self.bc_array, self.material_manager.get_gamma(),
                                                     _native.dtype_code(W), _native.stream_ptr())
            _native.check(rc, "jfa_edge_halo_update")
        return (W, U) if conservatives is not None else W
