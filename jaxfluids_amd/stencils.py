"""Cell-face reconstruction stencils on the HIP path.

Mirrors stencils/spatial_stencil.py, stencils/spatial_reconstruction.py and the
shock-capturing classes WENO1, WENO3-JS/-Z, WENO5-JS/-Z, TENO5 and TENO5-A
(stencils/reconstruction/shock_capturing/weno5_base.py, weno3_base.py, weno/*.py, teno/*.py).
``reconstruct_xi`` evaluates the stencil on the GPU (jfa_weno5_reconstruct_xi, which serves
every stencil) and returns a new face tensor, as the reference returns a new array.  The HIP
kernels read the 6-cell window of a face for every stencil but WENO7-JS (8 cells), so buffers need
3 halo cells (WENO7-JS: 4) whatever the stencil's own requirement.
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


KERNEL_HALOS = 3  # the HIP kernels read cells -3..2 of every face


class ShockCapturingStencil(SpatialReconstruction):
    """Common host side of the shock-capturing stencils: shape checks and the GPU call."""

    required_halos = 3
    _variant = None

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        if self.n < self.required_halos:
            raise AssertionError(
                f"Stencil with left-shift -{self.required_halos} is too wide for domain with {self.n} halo cells.")
        if self.n < KERNEL_HALOS:
            raise NotImplementedError(
                f"the HIP kernels read 3 cells on each side of a face: halo_cells >= 3 (got {self.n})")

    def _grid_for(self, buffer: torch.Tensor, axis: int = 0, dx=None):
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
        # dx: the cell size along `axis` (cell_sizes[axis]; the WENO6-CUM stencils read it)
        dxs = [1.0, 1.0, 1.0]
        if dx is not None:
            dxs[axis] = float(dx)
        return cells, _native.make_grid(cells, self.n, tuple(dxs))

    def reconstruct_xi(self, buffer: torch.Tensor, axis: int, j: int, dx=None, **kwargs) -> torch.Tensor:
        """Left (j=0) or right (j=1) face states along axis (weno5_js.py:73-115)."""
        _native.require_device(buffer, "buffer")
        if self.offset != 0:
            raise NotImplementedError("offset != 0 is not used on the convective-flux path")
        cells, grid = self._grid_for(buffer, axis, dx)
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


class WENO5Base(ShockCapturingStencil):
    """weno5_base.py:10-50: d = (1/10, 6/10, 3/10), 6-cell window, 3 halos."""

    required_halos = 3

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        self._dr = (1 / 10, 6 / 10, 3 / 10)
        self._cr = ((1 / 3, -7 / 6, 11 / 6), (-1 / 6, 5 / 6, 1 / 3), (1 / 3, 5 / 6, -1 / 6))
        self._stencil_size = 6


class WENO5JS(WENO5Base):
    """Jiang & Shu (1996) weights (weno/weno5_js.py:10-65)."""

    _variant = _native.STENCIL["WENO5-JS"]


class WENO5Z(WENO5Base):
    """Borges et al. (2008) weights (weno/weno5_z.py:10-59)."""

    _variant = _native.STENCIL["WENO5-Z"]


class TENO5(WENO5Base):
    """Fu et al. (2016): d = (0.05, 0.55, 0.40), q = 6, CT = 1e-5 (teno/teno5.py:12-73)."""

    _variant = _native.STENCIL["TENO5"]

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        self._dr = (0.05, 0.55, 0.40)
        self.C, self.q, self.CT = 1.0, 6, 1e-5


class TENO5A(WENO5Base):
    """TENO5 with the adaptive cutoff of teno/teno5_a.py:12-99 (xi = 1e-3, Cr = 0.24)."""

    _variant = _native.STENCIL["TENO5-A"]


class WENO6CU(ShockCapturingStencil):
    """Central-upwind WENO6 (Hu, Wang & Adams 2010): d = (1/20, 9/20, 9/20, 1/20), C = 20, on the
    6-cell window (weno6_base.py:10-59, weno/weno6_cu.py:10-61)."""

    required_halos = 3
    _variant = _native.STENCIL["WENO6-CU"]

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        self._dr = (1 / 20, 9 / 20, 9 / 20, 1 / 20)
        self._stencil_size = 6
        self.C_ = 20


class TENO6A(ShockCapturingStencil):
    """TENO6-A (teno/teno6_a.py:10-131): TENO6 with |beta_3|, |beta_6|, the spectrally optimised
    d (fourth order) and the adaptive cutoff (xi = 1e-3, Cr = 0.17, alpha = (10.5, 4.5))."""

    required_halos = 3
    _variant = _native.STENCIL["TENO6-A"]

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        self._dr = (0.0855682281039113, 0.4294317718960898, 0.1727270875843552, 0.3122729124156450)
        self._stencil_size = 6
        self.C, self.q = 1.0, 6
        self.xi, self.Cr, self.alpha_1, self.alpha_2 = 1e-3, 0.17, 10.5, 4.5


class WENO6CUM1(WENO6CU):
    """WENO6-CUM1 (weno/weno6_cum1.py:10-67; Hu & Adams 2011, scale separation for implicit LES):
    C_q = 1000, q = 4, alpha_k = d_k (C_q + tau_6/(beta_k + eps dx^2))^q; reconstruct_xi's dx is
    the cell size along the axis."""

    _variant = _native.STENCIL["WENO6-CUM1"]

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        self.Cq_, self.q_, self.eps_ = 1000, 4, 1e-8


class WENO6CUM2(WENO6CU):
    """WENO6-CUM2 (weno/weno6_cum2.py:10-72): C_q = 1000, q = 4, eps_ = 1e-8, chi = 1/eps_, the
    ratio (beta_ave + chi dx^2)/(beta_k + chi dx^2) on tau_6/(beta_k + eps_ dx^2)."""

    _variant = _native.STENCIL["WENO6-CUM2"]

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        self.Cq_, self.q_, self.eps_ = 1000, 4, 1e-8
        self.chi = 1.0 / self.eps_


class TENO6(ShockCapturingStencil):
    """TENO6 (teno6_base.py:10-64, teno/teno6.py:10-67): d = (0.05, 0.45, 0.30, 0.20), CT = 1e-7."""

    required_halos = 3
    _variant = _native.STENCIL["TENO6"]

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        self._dr = (0.050, 0.450, 0.300, 0.200)
        self._stencil_size = 6
        self.C, self.q, self.CT = 1.0, 6, 1e-7


class WENO7JS(ShockCapturingStencil):
    """Balsara & Shu (2000) WENO7-JS (weno/weno7_js.py:10-90): d = (1, 12, 18, 4)/35, four 4-point
    candidates on the 8-cell window, 4 halo cells."""

    required_halos = 4
    _variant = _native.STENCIL["WENO7-JS"]

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        self._dr = (1 / 35, 12 / 35, 18 / 35, 4 / 35)
        self._cr = ((-1 / 4, 13 / 12, -23 / 12, 25 / 12), (1 / 12, -5 / 12, 13 / 12, 1 / 4),
                    (-1 / 12, 7 / 12, 7 / 12, -1 / 12), (1 / 4, 13 / 12, -5 / 12, 1 / 12))
        self._stencil_size = 8


class WENO3Base(ShockCapturingStencil):
    """weno3_base.py:10-45: d = (1/3, 2/3), 4-cell window, 2 halos."""

    required_halos = 2

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        self._dr = (1 / 3, 2 / 3)
        self._cr = ((-0.5, 1.5), (0.5, 0.5))
        self._stencil_size = 4


class WENO3JS(WENO3Base):
    """Jiang & Shu weights on the 3-point stencils (weno/weno3_js.py:10-42)."""

    _variant = _native.STENCIL["WENO3-JS"]


class WENO3Z(WENO3Base):
    """Don & Borges (2013) weights (weno/weno3_z.py:10-41)."""

    _variant = _native.STENCIL["WENO3-Z"]


class WENO1(ShockCapturingStencil):
    """First-order upwind states (weno/weno1_js.py:10-30)."""

    required_halos = 1
    is_for_adaptive_mesh = True
    _variant = _native.STENCIL["WENO1"]

    def __init__(self, nh: int, inactive_axes: Sequence[str], offset: int = 0, **kwargs) -> None:
        super().__init__(nh, inactive_axes, offset)
        self._stencil_size = 2


# stencils/__init__.py:15-19, reconstruction/shock_capturing/weno/__init__.py:26-41 and
# teno/__init__.py:15-25 (the reconstruction stencils built on this path)
DICT_SPATIAL_RECONSTRUCTION: Dict[str, type] = {
    "WENO1": WENO1,
    "WENO3-JS": WENO3JS,
    "WENO3-Z": WENO3Z,
    "WENO5-JS": WENO5JS,
    "WENO5-Z": WENO5Z,
    "TENO5": TENO5,
    "TENO5-A": TENO5A,
    "WENO6-CU": WENO6CU,
    "TENO6": TENO6,
    "WENO6-CUM1": WENO6CUM1,
    "WENO6-CUM2": WENO6CUM2,
    "TENO6-A": TENO6A,
    "WENO7-JS": WENO7JS,
}
