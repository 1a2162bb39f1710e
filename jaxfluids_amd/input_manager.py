"""Case / numerical setup readers for the convective-flux path.

Accepts the reference's two setup dicts (or paths to their JSON files) with the
same keys (input/input_manager.py:24-120, input/numerical_setup/
read_conservatives.py:20-203, input/case_setup/*) and resolves the registry
strings to the classes built here.  Initial-condition strings of the form
"lambda x, ...: expr" are accepted through a whitelisting parser instead of
Python eval of arbitrary text (the reference evals them,
input/setup_reader.py:171).
"""
from __future__ import annotations

import ast
import json
import math
from typing import Any, Callable, Dict, Union

import numpy as np

from . import _native
from .config import precision
from .convective_fluxes import DICT_CONVECTIVE_SOLVER
from .data_types import (ActivePhysicsSetup, CaseSetup, ConservativesSetup, ConvectiveFluxesSetup, FluxSplittingSetup,
                         DissipativeFluxesSetup, DomainSetup, GeneralSetup, HighOrderGodunovSetup, MaterialSetup,
                         NumericalSetup, PositivitySetup, PrecisionSetup, TimeIntegrationSetup, TransportSetup)
from .riemann_solvers import DICT_RIEMANN_SOLVER, DICT_SIGNAL_SPEEDS
from .stencils import DICT_SPATIAL_RECONSTRUCTION, KERNEL_HALOS
from .time_integration import DICT_TIME_INTEGRATION

TUPLE_RECONSTRUCTION_VARIABLES = ("PRIMITIVE", "CONSERVATIVE", "CHAR-PRIMITIVE", "CHAR-CONSERVATIVE")
TUPLE_FROZEN_STATE = ("ARITHMETIC", "ROE")
FACE_LOCATIONS = ("east", "west", "north", "south", "top", "bottom")
PRIMES = ("rho", "u", "v", "w", "p")

# ------------------------------------------------------------------ safe lambda parsing
_ALLOWED_FUNCS = {
    "sin": np.sin, "cos": np.cos, "tan": np.tan, "exp": np.exp, "log": np.log, "sqrt": np.sqrt,
    "tanh": np.tanh, "sinh": np.sinh, "cosh": np.cosh, "arctan": np.arctan, "arctan2": np.arctan2,
    "abs": np.abs, "where": np.where, "minimum": np.minimum, "maximum": np.maximum, "heaviside": np.heaviside,
    "ones_like": np.ones_like, "zeros_like": np.zeros_like, "power": np.power, "sign": np.sign,
}
_ALLOWED_CONSTS = {"pi": math.pi, "e": math.e}
_MODULE_NAMES = ("jnp", "np", "numpy")
_ALLOWED_NODES = (
    ast.Expression, ast.Lambda, ast.arguments, ast.arg, ast.BinOp, ast.UnaryOp, ast.Compare, ast.BoolOp,
    ast.Call, ast.Name, ast.Load, ast.Constant, ast.Attribute, ast.IfExp, ast.Tuple,
    ast.Add, ast.Sub, ast.Mult, ast.Div, ast.Pow, ast.Mod, ast.FloorDiv, ast.USub, ast.UAdd,
    ast.Lt, ast.LtE, ast.Gt, ast.GtE, ast.Eq, ast.NotEq, ast.And, ast.Or, ast.BitAnd, ast.BitOr, ast.Not,
)


class _Namespace:
    def __init__(self):
        for k, v in {**_ALLOWED_FUNCS, **_ALLOWED_CONSTS}.items():
            setattr(self, k, v)


def parse_lambda(text: str) -> Callable:
    """Compile "lambda x, y, z: <expr>" after checking every AST node against a whitelist
    of arithmetic, comparisons and numpy math functions."""
    tree = ast.parse(text.strip(), mode="eval")
    if not isinstance(tree.body, ast.Lambda):
        raise ValueError(f"expected a lambda expression, got {text!r}")
    params = {a.arg for a in tree.body.args.args}
    for node in ast.walk(tree):
        if not isinstance(node, _ALLOWED_NODES):
            raise ValueError(f"disallowed syntax {type(node).__name__} in {text!r}")
        if isinstance(node, ast.Attribute):
            if not (isinstance(node.value, ast.Name) and node.value.id in _MODULE_NAMES
                    and (node.attr in _ALLOWED_FUNCS or node.attr in _ALLOWED_CONSTS)):
                raise ValueError(f"disallowed attribute access in {text!r}")
        if isinstance(node, ast.Name) and isinstance(node.ctx, ast.Load):
            if node.id not in params and node.id not in _MODULE_NAMES and node.id not in _ALLOWED_FUNCS \
                    and node.id not in _ALLOWED_CONSTS:
                raise ValueError(f"unknown name {node.id!r} in {text!r}")
        if isinstance(node, ast.Constant) and not isinstance(node.value, (int, float, bool)):
            raise ValueError(f"only numeric constants are allowed in {text!r}")
    ns = _Namespace()
    env = {"__builtins__": {}, **{m: ns for m in _MODULE_NAMES}, **_ALLOWED_FUNCS, **_ALLOWED_CONSTS}
    return eval(compile(tree, "<initial_condition>", "eval"), env)  # noqa: S307 - AST whitelisted above


def _as_callable(value: Any) -> Callable:
    if callable(value):
        return value
    if isinstance(value, (int, float)):
        c = float(value)
        return lambda *xs: c * np.ones_like(xs[0], dtype=float)
    if isinstance(value, str):
        return parse_lambda(value)
    raise ValueError(f"cannot interpret initial condition {value!r}")


def _load(setup: Union[str, Dict]) -> Dict:
    if isinstance(setup, str):
        with open(setup) as f:
            return json.load(f)
    return dict(setup)


# ------------------------------------------------------------------ readers
def read_dissipative_fluxes(d: Dict) -> DissipativeFluxesSetup:
    """read_conservatives.py:374-430 with the reference's defaults (CENTRAL4); the stencils built
    here are second-order central: reconstruction central_2, face derivative deriv_face_2,
    centre derivative deriv_center_2, no Laplacian form."""
    setup = DissipativeFluxesSetup(
        reconstruction_stencil=d.get("reconstruction_stencil", "CENTRAL4"),
        derivative_stencil_center=d.get("derivative_stencil_center", "CENTRAL4"),
        derivative_stencil_face=d.get("derivative_stencil_face", "CENTRAL4"),
        is_laplacian=bool(d.get("is_laplacian", False)),
        second_derivative_stencil_center=d.get("second_derivative_stencil_center", "CENTRAL4"))
    first = (setup.reconstruction_stencil, setup.derivative_stencil_center, setup.derivative_stencil_face)
    if any(v != "CENTRAL2" for v in first) or setup.is_laplacian:
        raise NotImplementedError(
            "dissipative_fluxes: the second-order central stencils are built (reconstruction_stencil, "
            f"derivative_stencil_center, derivative_stencil_face = 'CENTRAL2', is_laplacian false); got {first}"
            f"{', is_laplacian' if setup.is_laplacian else ''}")
    return setup


def _check_stencil(st: str, halo_cells: int):
    if st not in DICT_SPATIAL_RECONSTRUCTION:
        raise NotImplementedError(f"reconstruction_stencil {st!r} is not built (built: {tuple(DICT_SPATIAL_RECONSTRUCTION)})")
    stencil_cls = DICT_SPATIAL_RECONSTRUCTION[st]
    assert halo_cells >= stencil_cls.required_halos, \
        f"halo_cells={halo_cells} < {stencil_cls.required_halos} required by {st}"
    if halo_cells < KERNEL_HALOS:
        raise NotImplementedError(f"the HIP kernels need halo_cells >= {KERNEL_HALOS} for every stencil ({st}: {halo_cells})")
    return stencil_cls


def _read_godunov(cf: Dict, halo_cells: int) -> ConvectiveFluxesSetup:
    """read_godunov (read_conservatives.py:125-203)."""
    g = cf.get("godunov", {})
    rs = g.get("riemann_solver", "HLLC")
    if rs not in DICT_RIEMANN_SOLVER:
        raise NotImplementedError(f"riemann_solver {rs!r} is not built (built: {tuple(DICT_RIEMANN_SOLVER)})")
    ss = g.get("signal_speed", "EINFELDT")
    if ss not in DICT_SIGNAL_SPEEDS:
        raise NotImplementedError(f"signal_speed {ss!r} is not built (built: {tuple(DICT_SIGNAL_SPEEDS)})")
    stencil_cls = _check_stencil(g.get("reconstruction_stencil", "WENO5-Z"), halo_cells)
    rv = g.get("reconstruction_variable", "PRIMITIVE")
    assert rv in TUPLE_RECONSTRUCTION_VARIABLES, f"reconstruction_variable must be one of {TUPLE_RECONSTRUCTION_VARIABLES}"
    fz = g.get("frozen_state", "ARITHMETIC")
    assert fz in TUPLE_FROZEN_STATE, f"frozen_state must be one of {TUPLE_FROZEN_STATE}"
    godunov = HighOrderGodunovSetup(DICT_RIEMANN_SOLVER[rs], DICT_SIGNAL_SPEEDS[ss], stencil_cls, None, rv, fz, None)
    return ConvectiveFluxesSetup(DICT_CONVECTIVE_SOLVER["GODUNOV"], godunov)


def _read_flux_splitting(cf: Dict, halo_cells: int) -> ConvectiveFluxesSetup:
    """read_flux_splitting (read_conservatives.py:205-243): flux_splitting one of
    TUPLE_FLUX_SPLITTING (solvers/__init__.py:1-3; CLF has no eigenvalues in
    eigendecomposition.py:672-699), the reconstruction stencil and the frozen state."""
    d = cf.get("flux_splitting")
    assert isinstance(d, dict), "convective_fluxes: key 'flux_splitting' is required for FLUX-SPLITTING"
    fsp = d.get("flux_splitting", "ROE")
    assert fsp in ("ROE", "CLLF", "LLF", "CLF"), f"flux_splitting must be one of ('ROE', 'CLLF', 'LLF', 'CLF')"
    if fsp not in _native.FLUX_SPLITTING:
        raise NotImplementedError(f"flux_splitting {fsp!r} has no eigenvalues in the reference (built: ROE, CLLF, LLF)")
    stencil_cls = _check_stencil(d.get("reconstruction_stencil", "WENO5-Z"), halo_cells)
    if stencil_cls.__name__ == "WENO7JS":
        raise NotImplementedError("flux splitting with WENO7-JS is not built (the Godunov flux is)")
    fz = d.get("frozen_state", "ARITHMETIC")
    assert fz in TUPLE_FROZEN_STATE, f"frozen_state must be one of {TUPLE_FROZEN_STATE}"
    return ConvectiveFluxesSetup(DICT_CONVECTIVE_SOLVER["FLUX-SPLITTING"], None,
                                 FluxSplittingSetup(fsp, stencil_cls, None, fz))


def read_numerical_setup(d: Dict) -> NumericalSetup:
    """read_conservatives.py:20-203 (convective-flux keys)."""
    assert "conservatives" in d, "numerical_setup: key 'conservatives' is required"
    c = d["conservatives"]
    assert "halo_cells" in c, "numerical_setup/conservatives: key 'halo_cells' is required"
    halo_cells = int(c["halo_cells"])
    ti = c.get("time_integration")
    assert ti is not None, "numerical_setup/conservatives: key 'time_integration' is required"
    integrator = ti.get("integrator")
    assert integrator in DICT_TIME_INTEGRATION, f"time_integration/integrator must be one of {tuple(DICT_TIME_INTEGRATION)}"
    cfl = float(ti.get("CFL", 0.5))
    assert cfl > 0.0, "CFL must be > 0"
    fixed = ti.get("fixed_timestep", False)
    if fixed:
        fixed = float(fixed)
        assert fixed > 0.0, "fixed_timestep must be > 0"
    cf = c.get("convective_fluxes", {})
    solver = cf.get("convective_solver", "GODUNOV")
    if solver not in DICT_CONVECTIVE_SOLVER:
        raise NotImplementedError(f"convective_solver {solver!r} is not on this path (built: {tuple(DICT_CONVECTIVE_SOLVER)})")
    if solver == "FLUX-SPLITTING":
        conv = _read_flux_splitting(cf, halo_cells)
    else:
        conv = _read_godunov(cf, halo_cells)
    pos = c.get("positivity", {})  # read_positivity.py:8-64
    flux_limiter = pos.get("flux_limiter") or None
    if flux_limiter not in (None, "SIMPLE", "NASA", "HAS"):
        raise AssertionError(f"positivity/flux_limiter must be one of ('SIMPLE', 'NASA', 'HAS'), got {flux_limiter!r}")
    flux_partition = pos.get("flux_partition", "UNIFORM")
    if flux_partition not in ("UNIFORM", "CELLSIZE", "WAVESPEED"):
        raise AssertionError(f"positivity/flux_partition must be one of ('UNIFORM', 'CELLSIZE', 'WAVESPEED')")
    for key in ("is_thinc_interpolation_limiter", "is_volume_fraction_limiter", "is_acdi_flux_limiter"):
        if pos.get(key):
            raise NotImplementedError(f"positivity/{key} belongs to multiphase models, not on this path")
    positivity = PositivitySetup(is_interpolation_limiter=bool(pos.get("is_interpolation_limiter", False)),
                                 flux_limiter=flux_limiter or False,
                                 limit_velocity=bool(pos.get("limit_velocity", False)),
                                 flux_partition=flux_partition)
    ap = d.get("active_physics", {})
    active = ActivePhysicsSetup(
        is_convective_flux=bool(ap.get("is_convective_flux", True)),
        is_viscous_flux=bool(ap.get("is_viscous_flux", False)),
        is_heat_flux=bool(ap.get("is_heat_flux", False)),
        is_volume_force=bool(ap.get("is_volume_force", False)))
    if active.is_volume_force or not active.is_convective_flux:
        raise NotImplementedError("the convective flux (with optional viscous and heat fluxes) is built on this path")
    dissipative = read_dissipative_fluxes(c.get("dissipative_fluxes", {})) \
        if (active.is_viscous_flux or active.is_heat_flux) else None
    cons = ConservativesSetup(halo_cells, TimeIntegrationSetup(DICT_TIME_INTEGRATION[integrator], cfl, fixed), conv,
                              dissipative, positivity)
    pr = d.get("precision", {})
    prec = PrecisionSetup(bool(pr.get("is_double_precision_compute", True)),
                          bool(pr.get("is_double_precision_output", True)))
    return NumericalSetup(cons, active, prec)


_PROBE_T = np.array([0.05, 0.3, 0.7, 1.0, 1.3, 2.0, 3.7, 11.0, 260.0])


def power_law(value: Any, what: str):
    """(ref, T_ref, exponent) of a CUSTOM property value (read_material_manager.py:239-252,
    :304-316): a number is the constant ref; a lambda of T is accepted when it is a power law
    ref T^n, checked at probe temperatures from 0.05 to 260 (relative 1e-12), as the built
    transport laws evaluate it."""
    if isinstance(value, (int, float)) and not isinstance(value, bool):
        if not value >= 0.0:
            raise ValueError(f"{what}: value must be >= 0")
        return float(value), 1.0, 0.0
    f = parse_lambda(value) if isinstance(value, str) else value
    y = np.asarray(f(_PROBE_T), dtype=np.float64) * np.ones_like(_PROBE_T)
    ref = float(np.asarray(f(np.array([1.0])), dtype=np.float64).reshape(-1)[0])
    if not (ref >= 0.0 and np.all(np.isfinite(y))):
        raise ValueError(f"{what}: {value!r} must be finite and >= 0")
    if ref == 0.0:
        n = 0.0
    else:
        n = math.log2(float(np.asarray(f(np.array([2.0])), dtype=np.float64).reshape(-1)[0]) / ref)
        n = float(f"{n:.12g}")  # the exponent the lambda was written with
    if not np.allclose(y, ref * _PROBE_T ** n, rtol=1e-12, atol=0.0):
        raise NotImplementedError(
            f"{what}: CUSTOM lambdas are built as power laws ref * T**n (or constants); {value!r} is not one")
    return ref, 1.0, n


def read_transport(t: Any) -> Any:
    """material_properties/transport (read_material_manager.py:201-355) resolved to laws; None
    when absent.  dynamic_viscosity: CUSTOM (number or power-law lambda of T) or SUTHERLAND
    [mu_ref, T_ref, C]; bulk_viscosity; thermal_conductivity: CUSTOM (number or power-law lambda),
    PRANDTL (prandtl_number) or SUTHERLAND [kappa_ref, T_ref, C]."""
    if not t:
        return None
    kw = {}
    dv = t.get("dynamic_viscosity", {})
    model = dv.get("model", "CUSTOM")
    if model == "CUSTOM":
        kw["mu_ref"], kw["T_ref"], kw["mu_exp"] = power_law(dv.get("value", 0.0), "dynamic_viscosity")
    elif model == "SUTHERLAND":
        mu_ref, T_ref, C = (float(v) for v in dv["sutherland_parameters"])
        kw.update(mu_ref=mu_ref, T_ref=T_ref, mu_sutherland=C)
    else:
        raise ValueError(f"dynamic_viscosity model {model!r}: one of ('CUSTOM', 'SUTHERLAND')")
    kw["bulk_viscosity"] = float(t.get("bulk_viscosity", 0.0))
    if kw["bulk_viscosity"] < 0.0:
        raise ValueError("bulk_viscosity must be >= 0")
    tc = t.get("thermal_conductivity", {})
    model = tc.get("model", "CUSTOM")
    if model == "CUSTOM":
        kw["kappa_ref"], kw["kappa_T_ref"], kw["kappa_exp"] = power_law(tc.get("value", 0.0), "thermal_conductivity")
    elif model == "PRANDTL":
        kw["prandtl"] = float(tc["prandtl_number"])
        if not kw["prandtl"] > 0.0:
            raise ValueError("prandtl_number must be > 0")
    elif model == "SUTHERLAND":
        k_ref, T_ref, C = (float(v) for v in tc["sutherland_parameters"])
        kw.update(kappa_ref=k_ref, kappa_T_ref=T_ref, kappa_sutherland=C)
    else:
        raise ValueError(f"thermal_conductivity model {model!r}: one of ('CUSTOM', 'PRANDTL', 'SUTHERLAND')")
    return TransportSetup(**kw)


def read_case_setup(d: Dict) -> CaseSetup:
    gen = d.get("general", {})
    general = GeneralSetup(
        case_name=gen.get("case_name", "case"),
        end_time=float(gen.get("end_time", float("inf"))),
        end_step=int(gen.get("end_step", 2 ** 62)),
        save_path=gen.get("save_path", "./results"),
        save_dt=gen.get("save_dt"))
    dom = d["domain"]
    cells = tuple(int(dom[a]["cells"]) for a in ("x", "y", "z"))
    ranges = tuple(tuple(float(v) for v in dom[a]["range"]) for a in ("x", "y", "z"))
    dec = dom.get("decomposition", {})
    if any(int(dec.get(k, 1)) != 1 for k in ("split_x", "split_y", "split_z")):
        raise NotImplementedError("domain decomposition across devices is not part of the one-GPU build")
    bcs = {}
    for loc in FACE_LOCATIONS:
        entry = d["boundary_conditions"].get(loc, {"type": "INACTIVE"})
        bcs[loc] = entry["type"] if isinstance(entry, dict) else entry
    ic = d["initial_condition"]
    ic_calls = {k: _as_callable(ic[k]) for k in PRIMES}
    mp = d.get("material_properties", {}).get("equation_of_state", {})
    model = mp.get("model", "IdealGas")
    if model != "IdealGas":
        raise NotImplementedError(f"equation of state {model!r} is not on this path")
    transport = read_transport(d.get("material_properties", {}).get("transport"))
    material = MaterialSetup(model, float(mp.get("specific_heat_ratio", 1.4)),
                             float(mp.get("specific_gas_constant", 1.0)), transport)
    return CaseSetup(general, DomainSetup(cells, ranges), bcs, ic_calls, material)


class InputManager:
    """input/input_manager.py:24-120: reads both setups and builds the physics objects."""

    def __init__(self, case_setup: Union[str, Dict], numerical_setup: Union[str, Dict]) -> None:
        from .domain import DomainInformation
        from .equation_manager import EquationInformation, EquationManager, IdealGas, MaterialManager

        self.case_setup_dict = _load(case_setup)
        self.numerical_setup_dict = _load(numerical_setup)
        self.numerical_setup = read_numerical_setup(self.numerical_setup_dict)
        self.case_setup = read_case_setup(self.case_setup_dict)
        if self.numerical_setup.precision.is_double_precision_compute:
            precision.enable_double_precision()
        else:
            precision.enable_single_precision()
        cs = self.case_setup
        self.domain_information = DomainInformation(cs.domain.cells, cs.domain.ranges,
                                                    self.numerical_setup.conservatives.halo_cells)
        for loc in self.domain_information.active_face_locations:
            assert cs.boundary_conditions[loc] != "INACTIVE", f"boundary {loc} of an active axis cannot be INACTIVE"
        self.material_manager = MaterialManager(IdealGas(cs.material.specific_heat_ratio,
                                                         cs.material.specific_gas_constant))
        self.material_manager.transport = self._native_transport()
        # compute_time_step_size reads the dissipative limits from here
        self.domain_information.transport = self.material_manager.transport
        self.equation_information = EquationInformation(self.domain_information.active_axes_indices)
        self.equation_manager = EquationManager(self.material_manager, self.equation_information)

    def _native_transport(self):
        """jfa_transport of the active dissipative fluxes (None: inviscid).  The transport
        properties are required for them (read_material_manager.py:211-221).  On periodic domains
        the central stencils wrap onto the interior; with a non-periodic active face they read the
        face and edge halos (the transport's halos flag, for the unfused RHS)."""
        from . import _native

        ap = self.numerical_setup.active_physics
        if not (ap.is_viscous_flux or ap.is_heat_flux):
            return None
        cs = self.case_setup
        tr = cs.material.transport
        assert tr is not None, "material_properties/transport is required for viscous or heat fluxes"
        assert not ap.is_heat_flux or tr.prandtl is not None or tr.kappa_ref is not None, \
            "material_properties/transport/thermal_conductivity is required for the heat flux"
        halos = any(cs.boundary_conditions[loc] != "PERIODIC" for loc in self.domain_information.active_face_locations)
        kw = tr._asdict()
        if not ap.is_heat_flux:
            kw["prandtl"] = kw["kappa_ref"] = None
        return _native.make_transport(R=cs.material.specific_gas_constant, viscous=ap.is_viscous_flux, halos=halos,
                                      **kw)
