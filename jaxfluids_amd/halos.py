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

    def halo_buffer(self, primes_interior: torch.Tensor, eps: float) -> torch.Tensor:
        """The primitive buffer of interior primitives, built with torch operations so that
        autograd carries gradients through the halo copies: eps fill
        (initialization/helper_functions.py:49-60), interior set, face halos by the boundary
        rules of face_halo_update (outer/material.py:94-287: PERIODIC copies the opposite
        interior layers, ZEROGRADIENT the boundary cell, SYMMETRY the mirrored layers with the
        normal velocity negated, :721-731), and with viscous or heat fluxes the edge halos
        (edge_halo_update :289-380), here on fully periodic domains only (periodic images).
        Vertex halos keep the fill value.  The values are those the HIP halo kernels write."""
        dom = self.domain_information
        nh = dom.nh_conservatives
        N = dom.global_number_of_cells
        act = dom.active_axes_indices
        edges = self.fill_edge_halos_material and len(act) > 1
        if edges and any(self.boundary_conditions[loc] != "PERIODIC" for loc in dom.active_face_locations):
            raise NotImplementedError("differentiable edge halos are built for fully periodic domains only")
        shape = (primes_interior.shape[0],) + tuple(N[a] + 2 * nh if a in act else 1 for a in range(3))
        inner = [slice(nh, nh + N[a]) if a in act else slice(None) for a in range(3)]
        B = torch.full(shape, float(eps), dtype=primes_interior.dtype, device=primes_interior.device)
        B[(slice(None),) + tuple(inner)] = primes_interior
        axis_of = dom.face_location_to_axis_index
        for loc in _native.FACE_LOCATIONS:
            a = axis_of[loc]
            if a not in act:
                continue
            hi = loc in ("east", "north", "top")
            bc = self.boundary_conditions[loc]
            fill, src = list(inner), list(inner)
            fill[a] = slice(nh + N[a], N[a] + 2 * nh) if hi else slice(0, nh)
            if bc == "PERIODIC":
                src[a] = slice(nh, 2 * nh) if hi else slice(N[a], N[a] + nh)
            elif bc == "ZEROGRADIENT":
                src[a] = slice(nh + N[a] - 1, nh + N[a]) if hi else slice(nh, nh + 1)
            else:  # SYMMETRY: the nh cells next to the face, mirrored
                src[a] = slice(N[a], N[a] + nh) if hi else slice(nh, 2 * nh)
            h = B[(slice(None),) + tuple(src)].clone()
            if bc == "SYMMETRY":
                sign = torch.ones(h.shape[0], dtype=h.dtype, device=h.device)
                sign[a + 1] = -1.0
                h = torch.flip(h, dims=(a + 1,)) * sign.reshape((-1, 1, 1, 1))
            B[(slice(None),) + tuple(fill)] = h  # a ZEROGRADIENT layer broadcasts over the nh cells
        if edges:
            for i, a in enumerate(act):
                for b in act[i + 1:]:
                    for ha in (False, True):
                        for hb in (False, True):
                            fill, src = list(inner), list(inner)
                            fill[a] = slice(nh + N[a], N[a] + 2 * nh) if ha else slice(0, nh)
                            src[a] = slice(nh, 2 * nh) if ha else slice(N[a], N[a] + nh)
                            fill[b] = slice(nh + N[b], N[b] + 2 * nh) if hb else slice(0, nh)
                            src[b] = slice(nh, 2 * nh) if hb else slice(N[b], N[b] + nh)
                            B[(slice(None),) + tuple(fill)] = B[(slice(None),) + tuple(src)].clone()
        return B
