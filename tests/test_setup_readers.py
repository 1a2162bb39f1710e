"""Setup-dict readers of the dissipative fluxes (host logic, no GPU): the numerical setup's
active_physics / dissipative_fluxes keys and the case setup's material_properties/transport,
resolved to the laws the native transport evaluates."""
import copy

import numpy as np
import pytest

from jaxfluids_amd import _native
from jaxfluids_amd.input_manager import InputManager, power_law, read_dissipative_fluxes, read_numerical_setup, \
    read_transport

NUM = {"conservatives": {"halo_cells": 3, "time_integration": {"integrator": "RK3", "CFL": 0.5},
                         "convective_fluxes": {"convective_solver": "GODUNOV",
                                               "godunov": {"riemann_solver": "HLLC", "signal_speed": "EINFELDT",
                                                           "reconstruction_stencil": "WENO5-Z",
                                                           "reconstruction_variable": "CHAR-PRIMITIVE"}},
                         "dissipative_fluxes": {"reconstruction_stencil": "CENTRAL2",
                                                "derivative_stencil_center": "CENTRAL2",
                                                "derivative_stencil_face": "CENTRAL2"}},
       "active_physics": {"is_convective_flux": True, "is_viscous_flux": True, "is_heat_flux": True}}


def _case(bc="PERIODIC", transport=None):
    return {
        "general": {"case_name": "tgv", "end_step": 3},
        "domain": {a: {"cells": 8, "range": [0.0, 2 * np.pi]} for a in "xyz"},
        "boundary_conditions": {f: {"type": bc} for f in ("east", "west", "north", "south", "top", "bottom")},
        "initial_condition": {"rho": 1.0, "u": "lambda x,y,z: jnp.sin(x)*jnp.cos(y)*jnp.cos(z)",
                              "v": "lambda x,y,z: -jnp.cos(x)*jnp.sin(y)*jnp.cos(z)", "w": 0.0, "p": 100.0},
        "material_properties": {
            "equation_of_state": {"model": "IdealGas", "specific_heat_ratio": 1.4, "specific_gas_constant": 1.0},
            "transport": transport if transport is not None else {
                "dynamic_viscosity": {"model": "CUSTOM", "value": "lambda T: 0.0028 * T**0.76"},
                "bulk_viscosity": 0.0,
                "thermal_conductivity": {"model": "PRANDTL", "prandtl_number": 0.71}}},
    }


def test_dissipative_stencils_second_order_only():
    ns = read_numerical_setup(NUM)
    assert ns.active_physics.is_viscous_flux and ns.conservatives.dissipative_fluxes.derivative_stencil_face == "CENTRAL2"
    # the reference's defaults are fourth order: not built, so an error rather than a silent change
    with pytest.raises(NotImplementedError):
        read_dissipative_fluxes({})
    bad = copy.deepcopy(NUM)
    bad["conservatives"]["dissipative_fluxes"]["derivative_stencil_face"] = "CENTRAL4"
    with pytest.raises(NotImplementedError):
        read_numerical_setup(bad)
    bad = copy.deepcopy(NUM)
    bad["conservatives"]["dissipative_fluxes"]["is_laplacian"] = True
    with pytest.raises(NotImplementedError):
        read_numerical_setup(bad)
    inviscid = copy.deepcopy(NUM)
    inviscid["active_physics"] = {"is_convective_flux": True}
    assert read_numerical_setup(inviscid).conservatives.dissipative_fluxes is None


def test_custom_lambdas_resolve_to_power_laws():
    ref, T_ref, n = power_law("lambda T: 0.0028 * T**0.76", "mu")
    assert (ref, T_ref, n) == (0.0028, 1.0, 0.76)
    assert power_law(5e-4, "mu") == (5e-4, 1.0, 0.0)
    assert power_law("lambda T: 1e-3 * jnp.ones_like(T)", "mu") == (1e-3, 1.0, 0.0)
    assert power_law("lambda T: 2e-3 * jnp.sqrt(T)", "mu")[2] == 0.5
    for not_power in ("lambda T: 1e-3 + 1e-4 * T", "lambda T: 1e-3 * jnp.exp(T)"):
        with pytest.raises(NotImplementedError):
            power_law(not_power, "mu")
    with pytest.raises(ValueError):
        power_law(-1.0, "mu")


def test_transport_models():
    t = read_transport({"dynamic_viscosity": {"model": "SUTHERLAND", "sutherland_parameters": [1.716e-5, 273.15, 110.4]},
                        "bulk_viscosity": 1e-6,
                        "thermal_conductivity": {"model": "CUSTOM", "value": 0.026}})
    assert (t.mu_ref, t.T_ref, t.mu_sutherland, t.bulk_viscosity) == (1.716e-5, 273.15, 110.4, 1e-6)
    assert (t.kappa_ref, t.kappa_exp, t.prandtl) == (0.026, 0.0, None)
    t = read_transport({"dynamic_viscosity": {"model": "CUSTOM", "value": 1e-3},
                        "thermal_conductivity": {"model": "SUTHERLAND", "sutherland_parameters": [0.02, 300.0, 194.0]}})
    assert (t.kappa_ref, t.kappa_T_ref, t.kappa_sutherland) == (0.02, 300.0, 194.0)
    assert read_transport(None) is None
    with pytest.raises(ValueError):
        read_transport({"dynamic_viscosity": {"model": "POLYNOMIAL"}})


def test_input_manager_builds_native_transport():
    im = InputManager(_case(), NUM)
    t = im.material_manager.transport
    assert im.domain_information.transport is t
    assert (t.viscous, t.heat, t.mu_law, t.kappa_law) == (1, 1, _native.LAW_POWER, _native.LAW_PRANDTL)
    assert (t.mu_ref, t.T_ref, t.mu_param, t.prandtl, t.R) == (0.0028, 1.0, 0.76, 0.71, 1.0)
    # heat flux off: no conductivity is passed even when the case gives one
    num = copy.deepcopy(NUM)
    num["active_physics"]["is_heat_flux"] = False
    t = InputManager(_case(), num).material_manager.transport
    assert (t.viscous, t.heat) == (1, 0)
    # inviscid setups carry none
    num["active_physics"] = {"is_convective_flux": True}
    assert InputManager(_case(), num).material_manager.transport is None


def test_input_manager_dissipative_halos_flag():
    """Periodic domains wrap the dissipative stencils; any other active face makes the unfused RHS
    read the halos (the transport's halos flag), and the halo manager fills the edge halos for
    every dissipative run (halo_manager.py:125-129)."""
    from jaxfluids_amd.halos import HaloManager

    assert InputManager(_case(), NUM).material_manager.transport.halos == 0
    for bc in ("ZEROGRADIENT", "SYMMETRY"):
        im = InputManager(_case(bc=bc), NUM)
        assert im.material_manager.transport.halos == 1
        hm = HaloManager(im.case_setup.boundary_conditions, im.domain_information, im.equation_manager,
                         im.material_manager)
        assert hm.fill_edge_halos_material
    num = copy.deepcopy(NUM)
    num["active_physics"] = {"is_convective_flux": True}
    im = InputManager(_case(bc="ZEROGRADIENT"), num)
    assert not HaloManager(im.case_setup.boundary_conditions, im.domain_information, im.equation_manager,
                           im.material_manager).fill_edge_halos_material


def test_input_manager_requires_transport_for_dissipative_runs():
    case = _case()
    del case["material_properties"]["transport"]
    with pytest.raises(AssertionError):
        InputManager(case, NUM)
