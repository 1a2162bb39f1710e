"""Precision configuration (config/precision.py:19-231).

A process-wide singleton like the reference's ``precision``: double precision by
default; the spatial-stencil epsilon is 1e-30 in both precisions and the general
epsilon is the machine epsilon of the compute dtype.
"""
from __future__ import annotations

import numpy as np
import torch


class PrecisionConfig:
    def __init__(self) -> None:
        self.enable_double_precision()
        self.is_consistent_summation = False

    def enable_double_precision(self) -> None:
        # precision.py:44-56
        self.is_double = True
        self._eps = float(np.finfo(np.float64).eps)
        self._spatial_stencil_eps = 1e-30
        self._interpolation_limiter_eps = (1e-12, 1e-10)  # (density, pressure), precision.py:54
        self._flux_limiter_eps = (1e-12, 1e-10)  # precision.py:55

    def enable_single_precision(self) -> None:
        # precision.py:30-42
        self.is_double = False
        self._eps = float(np.finfo(np.float32).eps)
        self._spatial_stencil_eps = 1e-30
        self._interpolation_limiter_eps = (1e-6, 1e-6)  # precision.py:40
        self._flux_limiter_eps = (1e-6, 1e-6)  # precision.py:41

    def get_eps(self) -> float:
        return self._eps

    def get_spatial_stencil_eps(self) -> float:
        return self._spatial_stencil_eps

    def get_interpolation_limiter_eps(self):
        """(density, pressure) thresholds of the positivity limiter of the face states."""
        return self._interpolation_limiter_eps

    def get_flux_limiter_eps(self):
        """(density, pressure) thresholds of the positivity flux limiter."""
        return self._flux_limiter_eps

    def set_flux_limiter_eps(self, density: float, pressure: float) -> None:
        assert density > 0.0 and pressure > 0.0, "limiter epsilons must be positive"
        self._flux_limiter_eps = (float(density), float(pressure))

    def set_interpolation_limiter_eps(self, density: float, pressure: float) -> None:
        assert density > 0.0 and pressure > 0.0, "limiter epsilons must be positive"
        self._interpolation_limiter_eps = (float(density), float(pressure))

    def set_spatial_stencil_eps(self, eps: float) -> None:
        assert isinstance(eps, float) and eps > 0.0, "spatial stencil eps must be a positive float"
        self._spatial_stencil_eps = eps

    def set_eps(self, eps: float) -> None:
        assert isinstance(eps, float) and eps > 0.0, "eps must be a positive float"
        self._eps = eps

    def enable_consistent_summation(self) -> None:
        raise NotImplementedError("consistent summation is not built on this path")

    @property
    def torch_dtype(self):
        return torch.float64 if self.is_double else torch.float32


precision = PrecisionConfig()
