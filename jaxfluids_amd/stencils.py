"""Cell-face reconstruction stencils on the HIP path.

Mirrors stencils/spatial_stencil.py, stencils/spatial_reconstruction.py and the
WENO5 classes (stencils/reconstruction/shock_capturing/weno5_base.py,
weno/weno5_js.py, weno/weno5_z.py).  ``reconstruct_xi`` evaluates the stencil on
the GPU (jfa_weno5_reconstruct_xi) and returns a new face tensor, as the
reference returns a new array.
"""
from __future__ import annotations

from typing import Dict, List, Sequence

import torch

from . import _native
from .config import precision

AXES = ("x", "y", "z")


class SpatialStencil:
    """Base of all stencils (spatial_stencil.py:8-43)."""

    required_halos = 0
    is_for_adaptive_mesh = False

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        self.eps = precision.get_spatial_stencil_eps()
        self.nh = int(nh)
        self.offset = int(offset)
        self.n = self.nh - self.offset
        self.inactive_axes = tuple(inactive_axes)
        self._stencil_size = None


class SpatialReconstruction(SpatialStencil):
    """Cell-face reconstruction (spatial_reconstruction.py:10-58)."""

    def set_slices_stencil(self) -> None:
        # spatial_reconstruction.py:148-153: switches to window slicing for characteristic
        # reconstruction.  On this path characteristic reconstruction is fused into the
        # flux kernel, so the flag is only recorded.
        self.is_window = True

    def reconstruct_xi(self, buffer: torch.Tensor, axis: int, j: int, dx=None, **kwargs) -> torch.Tensor:
        raise NotImplementedError


class WENO5Base(SpatialReconstruction):
    """weno5_base.py:10-50: d = (1/10, 6/10, 3/10), 6-cell window, 3 halos."""

    required_halos = 3
    _variant = None

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        if self.n < self.required_halos:
            raise AssertionError(
                f"Stencil with left-shift -3 is too wide for domain with {self.n} halo cells.")
        self._dr = (1 / 10, 6 / 10, 3 / 10)
        self._cr = ((1 / 3, -7 / 6, 11 / 6), (-1 / 6, 5 / 6, 1 / 3), (1 / 3, 5 / 6, -1 / 6))
        self._stencil_size = 6

    def _grid_for(self, buffer: torch.Tensor):
        if buffer.dim() != 4:
            raise ValueError("buffer must have shape (nvar, Nx+2nh, Ny+2nh, Nz+2nh)")
        cells = []
        for i, ax in enumerate(AXES):
            b = buffer.shape[1 + i]
            if ax in self.inactive_axes:
                if b != 1:
                    raise ValueError(f"inactive axis {ax} must have extent 1")
                cells.append(1)
            else:
                cells.append(b - 2 * self.n)
        return cells, _native.make_grid(cells, self.n, (1.0, 1.0, 1.0))

    def reconstruct_xi(self, buffer: torch.Tensor, axis: int, j: int, dx=None, **kwargs) -> torch.Tensor:
        """Left (j=0) or right (j=1) face states along axis (weno5_js.py:73-115)."""
        _native.require_device(buffer, "buffer")
        if self.offset != 0:
            raise NotImplementedError("offset != 0 is not used on the convective-flux path")
        cells, grid = self._grid_for(buffer)
        if AXES[axis] in self.inactive_axes:
            raise ValueError(f"axis {axis} is inactive")
        face = list(cells)
        face[axis] += 1
        out = torch.empty((buffer.shape[0], *face), dtype=buffer.dtype, device=buffer.device)
        lib = _native.load()
        rc = lib.jfa_weno5_reconstruct_xi(
            _native.ptr(buffer), buffer.shape[0], grid, int(axis), int(j), self._variant, float(self.eps),
            _native.dtype_code(buffer), _native.ptr(out), _native.stream_ptr())
        _native.check(rc, "jfa_weno5_reconstruct_xi")
        return out


class WENO5JS(WENO5Base):
    """Jiang & Shu (1996) weights (weno/weno5_js.py:10-65)."""

    _variant = _native.STENCIL["WENO5-JS"]


class WENO5Z(WENO5Base):
    """Borges et al. (2008) weights (weno/weno5_z.py:10-59)."""

    _variant = _native.STENCIL["WENO5-Z"]


# stencils/__init__.py:15-19 (the reconstruction stencils built on this path)
DICT_SPATIAL_RECONSTRUCTION: Dict[str, type] = {
    "WENO5-JS": WENO5JS,
    "WENO5-Z": WENO5Z,
}
