"""Ideal-gas material and the single-phase equation manager.

Mirrors materials/single_materials/ideal_gas.py, materials/material_manager.py
(speed of sound :192-241) and equation_manager.py (prim<->cons :44-176, physical
flux :178-257).  Whole-buffer conversions run as HIP kernels.
"""
from __future__ import annotations

import math
from typing import Sequence

import torch

from . import _native


class IdealGas:
    """ideal_gas.py:14-137 (the parts on the convective path)."""

    def __init__(self, specific_heat_ratio: float = 1.4, specific_gas_constant: float = 1.0, **kwargs):
        self.gamma = float(specific_heat_ratio)
        self.R = float(specific_gas_constant)
        self.cp = self.gamma / (self.gamma - 1.0) * self.R
        self.pb = 0.0

    def get_specific_heat_ratio(self, T=None):
        return self.gamma

    def get_grueneisen(self, rho=None, T=None):
        return self.gamma - 1

    def get_speed_of_sound(self, p, rho):
        return torch.sqrt(self.gamma * p / rho)

    def get_pressure(self, e, rho):
        return (self.gamma - 1.0) * e * rho

    def get_specific_energy(self, p, rho):
        return p / (rho * (self.gamma - 1.0))

    def get_temperature(self, p, rho):
        return p / (rho * self.R)

    def get_psi(self, p, rho):
        return p / rho


class MaterialManager:
    """material_manager.py:19-719, single-phase ideal gas only."""

    equation_type = "SINGLE-PHASE"

    def __init__(self, material: IdealGas):
        self.material = material

    def get_gamma(self):
        return self.material.gamma

    def get_speed_of_sound(self, primitives=None, pressure=None, density=None, **kwargs):
        if pressure is None:
            pressure = primitives[4]
        if density is None:
            density = primitives[0]
        return self.material.get_speed_of_sound(pressure, density)

    def get_density(self, primitives):
        return primitives[0]


class EquationInformation:
    """equation_information.py:89-111 (single-phase index tables)."""

    equation_type = "SINGLE-PHASE"
    no_equations = 5
    no_primes = 5
    ids_mass = 0
    ids_velocity = (1, 2, 3)
    ids_energy = 4
    velocity_minor_axes = ((2, 3), (3, 1), (1, 2))
    primes_tuple = ("rho", "u", "v", "w", "p")
    cons_tuple = ("mass", "momentum_x", "momentum_y", "momentum_z", "energy")

    def __init__(self, active_axes_indices: Sequence[int] = (0, 1, 2)):
        self.active_axes_indices = tuple(active_axes_indices)


class EquationManager:
    """equation_manager.py:24-257 (single phase)."""

    def __init__(self, material_manager: MaterialManager, equation_information: EquationInformation = None):
        self.material_manager = material_manager
        self.equation_information = equation_information or EquationInformation()

    def _convert(self, src: torch.Tensor, to_cons: bool) -> torch.Tensor:
        _native.require_device(src, "buffer")
        if src.shape[0] != 5:
            raise ValueError("expected 5 variables on the leading axis")
        out = torch.empty_like(src)
        n = src[0].numel()
        lib = _native.load()
        fn = lib.jfa_cons_from_prim if to_cons else lib.jfa_prim_from_cons
        rc = fn(_native.ptr(src), _native.ptr(out), n, self.material_manager.get_gamma(), _native.dtype_code(src),
                _native.stream_ptr())
        _native.check(rc, "jfa_cons_from_prim" if to_cons else "jfa_prim_from_cons")
        return out

    def get_conservatives_from_primitives(self, primitives: torch.Tensor) -> torch.Tensor:
        """equation_manager.py:93-101."""
        return self._convert(primitives.contiguous(), True)

    def get_primitives_from_conservatives(self, conservatives: torch.Tensor, fluid_mask=None) -> torch.Tensor:
        """equation_manager.py:164-171."""
        return self._convert(conservatives.contiguous(), False)

    def get_fluxes_xi(self, primitives: torch.Tensor, conservatives: torch.Tensor, axis: int) -> torch.Tensor:
        """Physical flux along axis (equation_manager.py:237-252), elementwise on the device."""
        _native.require_device(primitives, "primitives")
        m = conservatives[axis + 1]
        f = [m, m * primitives[1], m * primitives[2], m * primitives[3], None]
        f[axis + 1] = f[axis + 1] + primitives[4]
        f[4] = primitives[axis + 1] * (conservatives[4] + primitives[4])
        return torch.stack(f, dim=0)
