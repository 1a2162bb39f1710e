"""Riemann solvers and signal speeds on the HIP path.

Mirrors solvers/riemann_solvers/riemann_solver.py (RiemannSolver, :13-186),
HLLC.py (:13-126), HLL.py (:11-61), HLLCLM.py (:13-136), AUSMP.py (:11-103), Rusanov.py
(:12-47) and the
DICT_RIEMANN_SOLVER /
DICT_SIGNAL_SPEEDS registries (solvers/riemann_solvers/__init__.py:16-34).
The Riemann problem of every face is solved in one HIP kernel
(jfa_riemann_xi); signal speeds are evaluated inside that kernel and are
selected by the registry entry passed at construction.  The signal speeds are also plain
functions (signal_speeds.py:10-157) evaluated elementwise by jfa_signal_speeds, for Riemann
solvers that call them.
"""
from __future__ import annotations

from typing import Callable, Dict, Tuple

import torch

from . import _native
from .config import precision


def _signal_speed(kind: str, u_L, u_R, a_L=None, a_R=None, rho_L=None, rho_R=None, p_L=None, p_R=None,
                  H_L=None, H_R=None, gamma=None) -> Tuple[torch.Tensor, torch.Tensor]:
    """(S_L, S_R) of `kind` by the elementwise HIP kernel (jfa_signal_speeds).  Inputs are device
    tensors or Python numbers, broadcast together; the result has their dtype (float64 or
    float32) and broadcast shape."""
    given = [x for x in (u_L, u_R, a_L, a_R, rho_L, rho_R, p_L, p_R, H_L, H_R) if x is not None]
    ref = next((x for x in given if isinstance(x, torch.Tensor)), None)
    if ref is None:
        raise _native.JFAError("signal speeds take device tensors (jaxfluids_amd has no CPU path)")
    _native.require_device(ref.contiguous(), "u_L")
    conv = lambda x: None if x is None else torch.as_tensor(x, dtype=ref.dtype, device=ref.device)  # noqa: E731
    names = ("u_L", "u_R", "a_L", "a_R", "rho_L", "rho_R", "p_L", "p_R", "H_L", "H_R")
    vals = [conv(x) for x in (u_L, u_R, a_L, a_R, rho_L, rho_R, p_L, p_R, H_L, H_R)]
    for v, nm in zip(vals, names):
        if v is not None and (v.device != ref.device or v.dtype != ref.dtype):
            raise _native.JFAError(f"{nm} must share the device and dtype of the other inputs")
    present = [v for v in vals if v is not None]
    shape = torch.broadcast_shapes(*[v.shape for v in present])
    vals = [None if v is None else v.expand(shape).contiguous() for v in vals]
    S_L = torch.empty(shape, dtype=ref.dtype, device=ref.device)
    S_R = torch.empty_like(S_L)
    g = 0.0 if gamma is None else float(gamma)
    rc = _native.load().jfa_signal_speeds(_native.SIGNAL_SPEED[kind], *[_native.ptr(v) for v in vals], g,
                                          S_L.numel(), _native.dtype_code(S_L), _native.ptr(S_L), _native.ptr(S_R),
                                          _native.stream_ptr())
    _native.check(rc, f"signal speed {kind}")
    return S_L, S_R


def signal_speed_Arithmetic(u_L, u_R, a_L, a_R, *args, **kwargs):
    """signal_speeds.py:10-29: min(u_mean - a_mean, u_L - a_L), max(u_mean + a_mean, u_R + a_R)."""
    return _signal_speed("ARITHMETIC", u_L, u_R, a_L, a_R)


def signal_speed_Rusanov(u_L, u_R, a_L, a_R, *args, **kwargs):
    """signal_speeds.py:31-49: -/+ max(|u_L| + a_L, |u_R| + a_R)."""
    return _signal_speed("RUSANOV", u_L, u_R, a_L, a_R)


def signal_speed_Davis(u_L, u_R, a_L, a_R, *args, **kwargs):
    """signal_speeds.py:51-69 (Toro 10.48)."""
    return _signal_speed("DAVIS", u_L, u_R, a_L, a_R)


def signal_speed_Davis_2(u_L, u_R, a_L, a_R, rho_L, rho_R, H_L, H_R, gamma, *args, **kwargs):
    """signal_speeds.py:71-107: u_Roe -/+ a_Roe from Roe's averages of u and of the total
    enthalpy H.  The Riemann solvers call signal speeds with p_L/p_R, not H_L/H_R, so there (as in
    the reference) this raises TypeError for the missing H_L, H_R."""
    return _signal_speed("DAVIS2", u_L, u_R, a_L, a_R, rho_L, rho_R, H_L=H_L, H_R=H_R, gamma=gamma)


def signal_speed_Einfeldt(u_L, u_R, a_L, a_R, rho_L, rho_R, *args, **kwargs):
    """signal_speeds.py:109-133 (Toro 10.52-10.54)."""
    return _signal_speed("EINFELDT", u_L, u_R, a_L, a_R, rho_L, rho_R)


def signal_speed_Toro(u_L, u_R, a_L, a_R, rho_L, rho_R, p_L, p_R, gamma, *args, **kwargs):
    """signal_speeds.py:135-157 (Toro 10.59-10.60) with the PVRS star pressure (:201-231)."""
    return _signal_speed("TORO", u_L, u_R, a_L, a_R, rho_L, rho_R, p_L, p_R, gamma=gamma)


DICT_SIGNAL_SPEEDS: Dict[str, Callable] = {
    "ARITHMETIC": signal_speed_Arithmetic,
    "RUSANOV": signal_speed_Rusanov,
    "DAVIS": signal_speed_Davis,
    "DAVIS2": signal_speed_Davis_2,
    "EINFELDT": signal_speed_Einfeldt,
    "TORO": signal_speed_Toro,
}
_SIGNAL_SPEED_NAME = {fn: name for name, fn in DICT_SIGNAL_SPEEDS.items()}


def signal_speed_name(signal_speed) -> str:
    """Registry name of a signal speed given by name or as one of the registry's functions."""
    if signal_speed is None:
        return "EINFELDT"
    if isinstance(signal_speed, str):
        if signal_speed not in DICT_SIGNAL_SPEEDS:
            raise AssertionError(f"signal speed {signal_speed} is not in DICT_SIGNAL_SPEEDS")
        return signal_speed
    name = _SIGNAL_SPEED_NAME.get(signal_speed)
    if name is None:
        raise NotImplementedError("only the functions of DICT_SIGNAL_SPEEDS run inside the HIP Riemann kernels")
    return name


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
        ss = signal_speed_name(self.signal_speed)
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
