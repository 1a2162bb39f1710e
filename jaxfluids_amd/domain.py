"""Uniform-grid domain information (domain/domain_information.py:68-300)."""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np

from . import _native

AXES = ("x", "y", "z")


class DomainInformation:
    """Geometry of one uniform block: cell counts, sizes, halo count, slices.

    Mirrors the attributes of the reference's DomainInformation that the
    convective-flux path reads: ``global_number_of_cells``, ``nh_conservatives``,
    ``active_axes``/``inactive_axes`` and their indices, ``dim``,
    ``domain_slices_conservatives``, ``smallest_cell_size`` and the cell sizes
    (homogenous.py:13, domain_information.py:290,698-701).
    """

    face_location_to_axis_index = {"east": 0, "west": 0, "north": 1, "south": 1, "top": 2, "bottom": 2}

    def __init__(self, number_of_cells: Sequence[int], domain_range: Sequence[Sequence[float]], nh: int):
        self.global_number_of_cells = tuple(int(n) for n in number_of_cells)
        self.device_number_of_cells = self.global_number_of_cells
        self.domain_size = tuple((float(r[0]), float(r[1])) for r in domain_range)
        self.nh_conservatives = int(nh)
        self.nh_geometry = None
        self.active_axes_indices = tuple(i for i in range(3) if self.global_number_of_cells[i] > 1)
        self.active_axes = tuple(AXES[i] for i in self.active_axes_indices)
        self.inactive_axes = tuple(AXES[i] for i in range(3) if i not in self.active_axes_indices)
        self.inactive_axes_indices = tuple(i for i in range(3) if i not in self.active_axes_indices)
        self.dim = len(self.active_axes_indices)
        self.split_factors = (1, 1, 1)
        self.is_parallel = False
        self.is_mesh_stretching = (False, False, False)
        # homogenous.py:13
        self.cell_sizes = tuple((b - a) / n for (a, b), n in zip(self.domain_size, self.global_number_of_cells))
        # domain_information.py:290
        self.one_cell_sizes = tuple(1.0 / d for d in self.cell_sizes)
        # domain_information.py:698-701
        self.smallest_cell_size = min(self.cell_sizes[i] for i in self.active_axes_indices)
        nh_ = self.nh_conservatives
        self.domain_slices_conservatives = tuple(
            slice(nh_, -nh_) if i in self.active_axes_indices else slice(None) for i in range(3))
        self.active_face_locations = tuple(
            loc for loc, a in self.face_location_to_axis_index.items() if a in self.active_axes_indices)
        self.grid = _native.make_grid(self.global_number_of_cells, self.nh_conservatives, self.cell_sizes)

    @property
    def buffer_shape(self) -> Tuple[int, int, int]:
        # initialization/helper_functions.py:49
        nh = self.nh_conservatives
        return tuple(n + 2 * nh if n > 1 else 1 for n in self.global_number_of_cells)

    @property
    def number_of_cells_total(self) -> int:
        return int(np.prod(self.global_number_of_cells))

    def get_device_cell_sizes(self):
        return self.cell_sizes

    def get_device_one_cell_sizes(self):
        return self.one_cell_sizes

    def compute_device_mesh_grid(self):
        """Cell-center mesh of the active axes, 'ij' indexing (homogenous.py:14)."""
        centers = []
        for i in self.active_axes_indices:
            a, b = self.domain_size[i]
            d = self.cell_sizes[i]
            centers.append(np.linspace(a + d / 2, b - d / 2, self.global_number_of_cells[i]))
        return np.meshgrid(*centers, indexing="ij")

    def interior(self):
        return (Ellipsis,) + self.domain_slices_conservatives
