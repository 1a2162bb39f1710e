"""Riemann solvers and signal speeds on the HIP path.

Mirrors solvers/riemann_solvers/riemann_solver.py (RiemannSolver, :13-186),
HLLC.py (:13-126), HLL.py (:11-61), HLLCLM.py (:13-136), AUSMP.py (:11-103), Rusanov.py
(:12-47) and the
DICT_RIEMANN_SOLVER /
DICT_SIGNAL_SPEEDS registries (solvers/riemann_solvers/__init__.py:16-34).
The Riemann problem of every face is solved in one HIP kernel
(jfa_riemann_xi); signal speeds are evaluated inside that kernel and are
selected by the registry entry passed at construction.
"""
from __future__ import annotations

from typing import Callable, Dict, Tuple

import torch

from . import _native
from .config import precision


class SignalSpeed:
    """Named signal-speed estimate (signal_speeds.py).  Selects the formula the HIP
    kernel evaluates; it carries no host-side arithmetic."""

    def __init__(self, name: str, reference: str):
        self.name = name
        self.reference = reference
        self.__name__ = f"signal_speed_{name.title()}"

    def __call__(self, *args, **kwargs):
        raise NotImplementedError(
            f"{self.__name__} is evaluated inside the HIP Riemann kernels; call solve_riemann_problem_xi")

    def __repr__(self):
        return f"<signal speed {self.name} ({self.reference})>"


signal_speed_Arithmetic = SignalSpeed("ARITHMETIC", "signal_speeds.py:10-29")
signal_speed_Rusanov = SignalSpeed("RUSANOV", "signal_speeds.py:31-49")
signal_speed_Davis = SignalSpeed("DAVIS", "signal_speeds.py:51-69")
signal_speed_Einfeldt = SignalSpeed("EINFELDT", "signal_speeds.py:109-133")
signal_speed_Toro = SignalSpeed("TORO", "signal_speeds.py:135-157")

DICT_SIGNAL_SPEEDS: Dict[str, SignalSpeed] = {
    "ARITHMETIC": signal_speed_Arithmetic,
    "RUSANOV": signal_speed_Rusanov,
    "DAVIS": signal_speed_Davis,
    "EINFELDT": signal_speed_Einfeldt,
    "TORO": signal_speed_Toro,
}


class RiemannSolver:
    """riemann_solver.py:13-102: solve_riemann_problem_xi -> (flux, u_hat, alpha_hat)."""

    _kind: str = None

    def __init__(self, material_manager, equation_manager, signal_speed=None, **kwargs) -> None:
        self.eps = precision.get_eps()
        self.material_manager = material_manager
        self.equation_manager = equation_manager
        if isinstance(signal_speed, str):
            signal_speed = DICT_SIGNAL_SPEEDS[signal_speed]
        self.signal_speed = signal_speed
        self.velocity_minor = ((2, 3), (3, 1), (1, 2))

    def _scheme(self, dtype: int) -> _native.JfaScheme:
        ss = self.signal_speed.name if isinstance(self.signal_speed, SignalSpeed) else "EINFELDT"
        return _native.make_scheme("PRIMITIVE", "WENO5-JS", self._kind, ss, "ARITHMETIC", dtype,
                                   self.material_manager.get_gamma(), precision.get_spatial_stencil_eps())

    def solve_riemann_problem_xi(self, primitives_L: torch.Tensor, primitives_R: torch.Tensor,
                                 conservatives_L: torch.Tensor, conservatives_R: torch.Tensor, axis: int,
                                 **kwargs) -> Tuple[torch.Tensor, None, None]:
        ts = [t.contiguous() for t in (primitives_L, primitives_R, conservatives_L, conservatives_R)]
        for t, nm in zip(ts, ("primitives_L", "primitives_R", "conservatives_L", "conservatives_R")):
            _native.require_device(t, nm)
            if t.shape != ts[0].shape or t.dtype != ts[0].dtype or t.shape[0] != 5:
                raise ValueError("face states must share a (5, ...) shape and dtype")
        out = torch.empty_like(ts[0])
        nfaces = ts[0][0].numel()
        scheme = self._scheme(_native.dtype_code(ts[0]))
        rc = _native.load().jfa_riemann_xi(*[_native.ptr(t) for t in ts], nfaces, int(axis), scheme,
                                           _native.ptr(out), _native.stream_ptr())
        _native.check(rc, "jfa_riemann_xi")
        return out, None, None


class HLLC(RiemannSolver):
    """HLLC (Toro et al. 1994), single phase (HLLC.py:13-126)."""

    _kind = "HLLC"

    def __init__(self, material_manager, equation_manager, signal_speed, **kwargs) -> None:
        super().__init__(material_manager, equation_manager, signal_speed)
        if self.signal_speed is None:
            raise AssertionError("HLLC requires a signal_speed (read_conservatives.py:146-150)")


class HLL(RiemannSolver):
    """HLL (Harten, Lax & van Leer 1983), single phase (HLL.py:11-61)."""

    _kind = "HLL"

    def __init__(self, material_manager, equation_manager, signal_speed, **kwargs) -> None:
        super().__init__(material_manager, equation_manager, signal_speed)
        if self.signal_speed is None:
            raise AssertionError("HLL requires a signal_speed (read_conservatives.py:145-151)")


class HLLCLM(RiemannSolver):
    """HLLC with the low-Mach correction of Fleischmann et al. (2020), Ma_limit = 0.1
    (HLLCLM.py:13-136)."""

    _kind = "HLLC-LM"

    def __init__(self, material_manager, equation_manager, signal_speed, **kwargs) -> None:
        super().__init__(material_manager, equation_manager, signal_speed)
        if self.signal_speed is None:
            raise AssertionError("HLLC-LM requires a signal_speed (read_conservatives.py:145-151)")
        self.Ma_limit = 0.1


class AUSMP(RiemannSolver):
    """AUSM+ (Liou 1996) with the arithmetic interface speed of sound, alpha = 3/16, beta = 1/8
    (AUSMP.py:11-103); the signal speed is not used."""

    _kind = "AUSMP"

    def __init__(self, material_manager, equation_manager, signal_speed=None, **kwargs) -> None:
        super().__init__(material_manager, equation_manager, signal_speed)
        self.interface_speed_of_sound = "ARITHMETIC"
        self.alpha, self.beta = 3.0 / 16.0, 1.0 / 8.0


class Rusanov(RiemannSolver):
    """Rusanov / local Lax-Friedrichs (Rusanov.py:12-47)."""

    _kind = "RUSANOV"


DICT_RIEMANN_SOLVER: Dict[str, type] = {
    "HLLC": HLLC,
    "HLL": HLL,
    "HLLC-LM": HLLCLM,
    "AUSMP": AUSMP,
    "RUSANOV": Rusanov,
}
