"""Initial material fields and time control (initialization/).

Mirrors InitializationManager.initialization (initialization/
initialization_manager.py:142) -> MaterialFieldsInitializer.
from_primitive_initial_condition (material_fields_initializer.py:590-690): the
buffer is pre-filled with eps (initialization/helper_functions.py:49-60), the
initial condition is evaluated on the cell centers, U = cons(W) on the whole
buffer, the face halos are filled, and the first dt is computed
(time_control_initializer.py:34-80).
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .config import precision
from .data_types import JaxFluidsBuffers, MaterialFieldBuffers, TimeControlVariables
from .halos import HaloManager
from .time_integration import compute_time_step_size


class InitializationManager:
    def __init__(self, input_manager, device: Optional[torch.device] = None) -> None:
        self.input_manager = input_manager
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else device
        im = input_manager
        self.halo_manager = HaloManager(im.case_setup.boundary_conditions, im.domain_information,
                                        im.equation_manager, im.material_manager)

    def primitives_on_cells(self) -> np.ndarray:
        """Initial primitives of the interior cells, float64 (material_fields_initializer.py:610-628)."""
        im = self.input_manager
        dom = im.domain_information
        mesh = dom.compute_device_mesh_grid()
        fields = [np.asarray(im.case_setup.initial_condition[k](*mesh), dtype=np.float64) *
                  np.ones(mesh[0].shape) for k in ("rho", "u", "v", "w", "p")]
        prim = np.stack(fields, axis=0)
        return prim.reshape((5,) + dom.global_number_of_cells)

    def initialization(self, primitives: Optional[np.ndarray] = None,
                       restart_file_path: Optional[str] = None) -> JaxFluidsBuffers:
        """Creates (W, U) on the device, fills halos, computes the first dt.

        ``primitives`` (5, Nx, Ny, Nz) overrides the case-setup initial condition;
        ``restart_file_path`` starts from an HDF5 output file instead, at its time
        (MaterialFieldsInitializer.from_restart_file, material_fields_initializer.py:212-333)."""
        im = self.input_manager
        dom = im.domain_information
        dtype = precision.torch_dtype
        t0 = 0.0
        if restart_file_path is not None:
            from .hdf5 import read_restart

            t0, primitives = read_restart(restart_file_path, dom)
        prim = self.primitives_on_cells() if primitives is None else np.asarray(primitives, dtype=np.float64)
        if prim.shape != (5,) + dom.global_number_of_cells:
            raise ValueError(f"initial primitives must have shape {(5,) + dom.global_number_of_cells}")
        W = torch.full((5,) + dom.buffer_shape, precision.get_eps(), dtype=dtype, device=self.device)
        W[dom.interior()] = torch.as_tensor(prim, dtype=dtype, device=self.device)
        U = im.equation_manager.get_conservatives_from_primitives(W)
        hm = self.halo_manager
        W, U = hm.perform_halo_update_material(W, 0.0, hm.fill_edge_halos_material, False, U, inplace=True)
        dt = compute_time_step_size(W, None, None, None, None, dom, im.equation_information, im.material_manager,
                                    None, im.numerical_setup)
        gs = im.case_setup.general
        tcv = TimeControlVariables(t0, 0, dt, im.numerical_setup.conservatives.time_integration.fixed_timestep,
                                   gs.end_time, gs.end_step)
        return JaxFluidsBuffers(MaterialFieldBuffers(U, W, None), tcv)
