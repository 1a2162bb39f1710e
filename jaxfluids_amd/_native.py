"""ctypes binding of libjaxfluids_amd.so (the C ABI declared in include/jaxfluids_amd.h).

The product path has no CPU fallback: if the library is missing or no HIP device
is present, calls raise.  Tensors are passed as raw device pointers together with
torch's current HIP stream.
"""
from __future__ import annotations

import ctypes
import os
from typing import Tuple, Optional, Sequence

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjaxfluids_amd.so")

# constants (mirror include/jaxfluids_amd.h)
JFA_OK, JFA_EINVAL, JFA_EUNSUPPORTED, JFA_ELAUNCH = 0, 1, 2, 3
NOALLOC_SCRATCH_BYTES = 64 << 20  # JFA_NOALLOC_SCRATCH_BYTES (include/jaxfluids_amd.h)
RECON_VAR = {"PRIMITIVE": 0, "CONSERVATIVE": 1, "CHAR-PRIMITIVE": 2, "CHAR-CONSERVATIVE": 3}
STENCIL = {"WENO5-JS": 0, "WENO5-Z": 1, "WENO1": 2, "WENO3-JS": 3, "WENO3-Z": 4, "TENO5": 5, "TENO5-A": 6,
           "WENO6-CU": 7, "TENO6": 8, "WENO6-CUM1": 9, "WENO6-CUM2": 10,
           "TENO6-A": 11, "WENO7-JS": 12}
RIEMANN = {"HLLC": 0, "RUSANOV": 1, "HLL": 2, "HLLC-LM": 3, "AUSMP": 4}
# convective_solver FLUX-SPLITTING with its eigenvalues (flux_splitting), in place of a Riemann solver
FLUX_SPLITTING = {"ROE": 5, "CLLF": 6, "LLF": 7}
RIEMANN.update({"FLUX-SPLITTING-" + k: v for k, v in FLUX_SPLITTING.items()})
SIGNAL_SPEED = {"ARITHMETIC": 0, "RUSANOV": 1, "DAVIS": 2, "EINFELDT": 3, "TORO": 4, "DAVIS2": 5}
# Riemann solvers that call their signal speed with (u, a, rho_L=, rho_R=, p_L=, p_R=, gamma=)
# (HLLC.py:93-102, HLL.py:38-47, HLLCLM.py:72-81)
_SIGNAL_SPEED_CALLERS = ("HLLC", "HLL", "HLLC-LM")
FROZEN = {"ARITHMETIC": 0, "ROE": 1}
BC = {"PERIODIC": 0, "ZEROGRADIENT": 1, "SYMMETRY": 2}
F64, F32 = 0, 1
FACE_LOCATIONS = ("east", "west", "north", "south", "top", "bottom")


class JfaGrid(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 3), ("nh", ctypes.c_int32), ("dx", ctypes.c_double * 3)]


class JfaTransport(ctypes.Structure):
    _fields_ = [("viscous", ctypes.c_int32), ("heat", ctypes.c_int32), ("R", ctypes.c_double),
                ("mu_law", ctypes.c_int32), ("kappa_law", ctypes.c_int32),
                ("mu_ref", ctypes.c_double), ("T_ref", ctypes.c_double), ("mu_param", ctypes.c_double),
                ("bulk_viscosity", ctypes.c_double), ("prandtl", ctypes.c_double),
                ("kappa_ref", ctypes.c_double), ("kappa_T_ref", ctypes.c_double), ("kappa_param", ctypes.c_double),
                ("halos", ctypes.c_int32)]


class JfaScheme(ctypes.Structure):
    _fields_ = [
        ("recon_var", ctypes.c_int32),
        ("stencil", ctypes.c_int32),
        ("riemann", ctypes.c_int32),
        ("signal_speed", ctypes.c_int32),
        ("frozen_state", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("gamma", ctypes.c_double),
        ("eps_stencil", ctypes.c_double),
        ("transport", ctypes.POINTER(JfaTransport)),
        ("interpolation_limiter", ctypes.c_int32),
        ("limit_velocity", ctypes.c_int32),
        ("limiter_eps_density", ctypes.c_double),
        ("limiter_eps_pressure", ctypes.c_double),
        ("flux_limiter", ctypes.c_int32),
        ("flux_partition", ctypes.c_int32),
        ("flux_limiter_eps_density", ctypes.c_double),
        ("flux_limiter_eps_pressure", ctypes.c_double),
    ]


class JfaCase2D(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 2), ("nh", ctypes.c_int32), ("bc", ctypes.c_int32 * 4),
                ("reserved", ctypes.c_int32), ("dx", ctypes.c_double * 2), ("offset", ctypes.c_int64)]


class JFAError(RuntimeError):
    pass


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double
_SZ = ctypes.c_size_t
_PG = ctypes.POINTER(JfaGrid)
_PS = ctypes.POINTER(JfaScheme)
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PD = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every symbol declared in include/jaxfluids_amd.h
SIGNATURES = {
    "jfa_version": (ctypes.c_char_p, []),
    "jfa_last_error": (ctypes.c_char_p, []),
    "jfa_weno5_reconstruct_xi": (_I32, [_P, _I32, _PG, _I32, _I32, _I32, _D, _I32, _P, _P]),
    "jfa_riemann_xi": (_I32, [_P, _P, _P, _P, _I64, _I32, _PS, _P, _P]),
    "jfa_signal_speeds": (_I32, [_I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _I64, _I32, _P, _P, _P]),
    "jfa_compute_flux_xi": (_I32, [_P, _P, _PG, _PS, _I32, _P, _P, _P]),
    "jfa_compute_rhs": (_I32, [_P, _P, _PG, _PS, _D, _P, _P, _SZ, _P]),
    "jfa_rhs_workspace_bytes": (_SZ, [_PG, _I32]),
    "jfa_rk3_stage": (_I32, [_P, _P, _P, _P, _PG, _PS, _I32, _P, _D, _P]),
    "jfa_halo_update": (_I32, [_P, _P, _PG, _PI32, _D, _I32, _P]),
    "jfa_edge_halo_update": (_I32, [_P, _P, _PG, _PI32, _D, _I32, _P]),
    "jfa_compute_dt": (_I32, [_P, _PG, _D, _D, _D, _I32, ctypes.POINTER(JfaTransport), _P, _P, _P]),
    "jfa_cons_from_prim": (_I32, [_P, _P, _I64, _D, _I32, _P]),
    "jfa_prim_from_cons": (_I32, [_P, _P, _I64, _D, _I32, _P]),
    "jfa_rk3_stage_fused": (_I32, [_P, _P, _P, _P, _P, _PG, _PS, _PI32, _I32, _P, _D, _D, _D, _P, _P, _SZ, _P]),
    "jfa_stage_workspace_bytes": (_SZ, [_PG, _PS]),
    "jfa_rk3_stage_fused_noalloc": (_I32, [_P, _P, _P, _P, _P, _PG, _PS, _PI32, _I32, _P, _D, _D, _D, _P, _P]),
    "jfa_stage_vjp_workspace_bytes": (_SZ, [_PG, _PS]),
    "jfa_adjoint_set_march_min_waves": (_I32, [_I64]),
    "jfa_rk3_stage_vjp": (_I32, [_P, _P, _P, _P, _PG, _PS, _PI32, _I32, _D, _P, _SZ, _P]),
    "jfa_batch2d_workspace_bytes": (_SZ, [_P, _I32, _I32]),
    "jfa_batch2d_plan": (_I32, [_P, _I32, _I32, _P, _SZ, _P]),
    "jfa_batch2d_stage": (_I32, [_P, _P, _P, _P, _P, _P, _I32, _I32, _PS, _I32, _P, _D, _D, _P, _P, _SZ, _P]),
}

_lib: Optional[ctypes.CDLL] = None


def load() -> ctypes.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise JFAError(
            f"{LIB_PATH} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(jaxfluids_amd has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != JFA_OK:
        msg = {JFA_EINVAL: "invalid argument", JFA_EUNSUPPORTED: "unsupported option",
               JFA_ELAUNCH: "HIP launch failure"}.get(rc, f"status {rc}")
        detail = load().jfa_last_error().decode() if rc == JFA_ELAUNCH else ""
        raise JFAError(f"{what}: {msg} {detail}".strip())


def require_device(t: torch.Tensor, name: str = "tensor") -> None:
    if not t.is_cuda:
        raise JFAError(f"{name} must be a HIP device tensor (jaxfluids_amd has no CPU path)")
    if not t.is_contiguous():
        raise JFAError(f"{name} must be contiguous")
    if t.dtype not in (torch.float64, torch.float32):
        raise JFAError(f"{name} must be float64 or float32")


def dtype_code(t: torch.Tensor) -> int:
    return F64 if t.dtype == torch.float64 else F32


def stream_ptr() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t: Optional[torch.Tensor]) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def make_grid(cells: Sequence[int], nh: int, dx: Sequence[float]) -> JfaGrid:
    g = JfaGrid()
    for i in range(3):
        g.n[i] = int(cells[i])
        g.dx[i] = float(dx[i])
    g.nh = int(nh)
    return g


LAW_POWER, LAW_SUTHERLAND, LAW_PRANDTL = 0, 1, 2
FLUX_LIMITER = {None: 0, "SIMPLE": 1, "NASA": 2, "HAS": 3}
FLUX_PARTITION = {"UNIFORM": 0, "CELLSIZE": 1, "WAVESPEED": 2}


def make_transport(R: float = 1.0, mu_ref: float = 0.0, T_ref: float = 1.0, mu_exp: float = 0.0,
                   bulk_viscosity: float = 0.0, prandtl: Optional[float] = None, viscous: bool = True,
                   mu_sutherland: Optional[float] = None, kappa_ref: Optional[float] = None,
                   kappa_T_ref: float = 1.0, kappa_exp: float = 0.0,
                   kappa_sutherland: Optional[float] = None, halos: bool = False) -> JfaTransport:
    """Dissipative-flux transport (material.py:82-124).

    Viscosity: mu_ref (T/T_ref)^mu_exp, or SUTHERLAND with constant mu_sutherland.  Heat flux when
    prandtl (kappa = cp mu / prandtl) or kappa_ref (kappa_ref (T/kappa_T_ref)^kappa_exp, or
    SUTHERLAND with kappa_sutherland) is given.  halos: jfa_compute_rhs reads the buffer's face
    and edge halos (a non-periodic domain) instead of wrapping periodically."""
    t = JfaTransport()
    t.halos = 1 if halos else 0
    t.viscous = 1 if viscous else 0
    t.heat = 1 if (prandtl is not None or kappa_ref is not None) else 0
    t.R, t.mu_ref, t.T_ref = float(R), float(mu_ref), float(T_ref)
    t.mu_law = LAW_SUTHERLAND if mu_sutherland is not None else LAW_POWER
    t.mu_param = float(mu_sutherland if mu_sutherland is not None else mu_exp)
    t.bulk_viscosity = float(bulk_viscosity)
    t.prandtl = float(prandtl) if prandtl is not None else 1.0
    if prandtl is not None:
        t.kappa_law = LAW_PRANDTL
        t.kappa_ref, t.kappa_T_ref, t.kappa_param = 0.0, 1.0, 0.0
    else:
        t.kappa_law = LAW_SUTHERLAND if kappa_sutherland is not None else LAW_POWER
        t.kappa_ref = float(kappa_ref) if kappa_ref is not None else 0.0
        t.kappa_T_ref = float(kappa_T_ref)
        t.kappa_param = float(kappa_sutherland if kappa_sutherland is not None else kappa_exp)
    return t


def make_scheme(recon_var: str, stencil: str, riemann: str, signal_speed: Optional[str], frozen_state: str,
                dtype: int, gamma: float, eps_stencil: float, transport: Optional[JfaTransport] = None,
                interpolation_limiter: bool = False, limit_velocity: bool = False,
                limiter_eps: Tuple[float, float] = (1e-12, 1e-10), flux_limiter: Optional[str] = None,
                flux_partition: str = "UNIFORM",
                flux_limiter_eps: Tuple[float, float] = (1e-12, 1e-10)) -> JfaScheme:
    """Native scheme; ``limiter_eps`` = (density, pressure) epsilons of the positivity limiter of
    the reconstructed face states (used when ``interpolation_limiter``); ``flux_limiter`` "SIMPLE"
    / "NASA" / "HAS" with ``flux_partition`` "UNIFORM" / "CELLSIZE" / "WAVESPEED" and its epsilons, the positivity flux
    limiter of the stage and RHS entry points."""
    s = JfaScheme()
    if signal_speed == "DAVIS2":
        if riemann in _SIGNAL_SPEED_CALLERS:
            # what the reference raises when its Riemann solver calls signal_speed_Davis_2
            # (signal_speeds.py:71-73) with p_L/p_R instead of the total enthalpies
            raise TypeError("signal_speed_Davis_2() missing 2 required positional arguments: 'H_L' and 'H_R'")
        signal_speed = None  # the Riemann solver does not call a signal speed
    try:
        s.recon_var = RECON_VAR[recon_var]
        s.stencil = STENCIL[stencil]
        s.riemann = RIEMANN[riemann]
        s.signal_speed = SIGNAL_SPEED[signal_speed or "EINFELDT"]
        s.frozen_state = FROZEN[frozen_state or "ARITHMETIC"]
    except KeyError as e:
        raise NotImplementedError(f"option {e} is not built into jaxfluids_amd") from None
    s.dtype = int(dtype)
    s.gamma = float(gamma)
    s.eps_stencil = float(eps_stencil)
    if transport is not None:
        s._transport_ref = transport  # keeps the struct alive as long as the scheme
        s.transport = ctypes.pointer(transport)
    s.interpolation_limiter = 1 if interpolation_limiter else 0
    s.limit_velocity = 1 if limit_velocity else 0
    s.limiter_eps_density, s.limiter_eps_pressure = (float(v) for v in limiter_eps)
    try:
        s.flux_limiter = FLUX_LIMITER[flux_limiter or None]
        s.flux_partition = FLUX_PARTITION[flux_partition]
    except KeyError as e:
        raise NotImplementedError(f"positivity option {e} is not built into jaxfluids_amd") from None
    s.flux_limiter_eps_density, s.flux_limiter_eps_pressure = (float(v) for v in flux_limiter_eps)
    return s


def bc_array(bcs) -> ctypes.Array:
    arr = (ctypes.c_int32 * 6)()
    for i, loc in enumerate(FACE_LOCATIONS):
        arr[i] = BC.get(bcs.get(loc, "PERIODIC"), -1) if isinstance(bcs, dict) else BC[bcs[i]]
        if arr[i] < 0:
            raise NotImplementedError(f"boundary type {bcs.get(loc)} at {loc} is not on the hot path")
    return arr
