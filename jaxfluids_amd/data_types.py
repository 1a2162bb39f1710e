"""Setup and buffer containers of the convective-flux path.

Field names follow the reference NamedTuples so code written against them reads
the same: data_types/numerical_setup/conservatives.py:10-90,
data_types/buffers.py:9-115, data_types/case_setup/*.
"""
from __future__ import annotations

from typing import Any, Callable, Dict, NamedTuple, Optional, Tuple

import torch


# --------------------------------------------------------------- numerical setup
class HighOrderGodunovSetup(NamedTuple):
    """conservatives.py:35-42"""
    riemann_solver: type
    signal_speed: Any
    reconstruction_stencil: type
    split_reconstruction: Any = None
    reconstruction_variable: str = "PRIMITIVE"
    frozen_state: str = "ARITHMETIC"
    catum_setup: Any = None


class FluxSplittingSetup(NamedTuple):
    """conservatives.py:44-48"""
    flux_splitting: str
    reconstruction_stencil: type
    split_reconstruction: Any = None
    frozen_state: str = "ARITHMETIC"


class ConvectiveFluxesSetup(NamedTuple):
    convective_solver: type
    godunov: Optional[HighOrderGodunovSetup]
    flux_splitting: Optional[FluxSplittingSetup] = None


class TimeIntegrationSetup(NamedTuple):
    """conservatives.py:10-13"""
    integrator: type
    CFL: float = 0.5
    fixed_timestep: Any = False


class PositivitySetup(NamedTuple):
    """conservatives.py PositivitySetup (the single-phase fields)"""
    is_interpolation_limiter: bool = False
    is_thinc_interpolation_limiter: bool = False
    flux_limiter: Any = False
    limit_velocity: bool = False
    flux_partition: str = "UNIFORM"


class DissipativeFluxesSetup(NamedTuple):
    """conservatives.py:69-74 (stencil names; the built path is second-order central)"""
    reconstruction_stencil: str = "CENTRAL2"
    derivative_stencil_center: str = "CENTRAL2"
    derivative_stencil_face: str = "CENTRAL2"
    is_laplacian: bool = False
    second_derivative_stencil_center: str = "CENTRAL2"


class ConservativesSetup(NamedTuple):
    halo_cells: int
    time_integration: TimeIntegrationSetup
    convective_fluxes: ConvectiveFluxesSetup
    dissipative_fluxes: Any = None
    positivity: PositivitySetup = PositivitySetup()


class ActivePhysicsSetup(NamedTuple):
    is_convective_flux: bool = True
    is_viscous_flux: bool = False
    is_heat_flux: bool = False
    is_volume_force: bool = False
    is_geometric_source: bool = False
    is_surface_tension: bool = False


class PrecisionSetup(NamedTuple):
    is_double_precision_compute: bool = True
    is_double_precision_output: bool = True


class NumericalSetup(NamedTuple):
    conservatives: ConservativesSetup
    active_physics: ActivePhysicsSetup
    precision: PrecisionSetup


# --------------------------------------------------------------- case setup
class GeneralSetup(NamedTuple):
    case_name: str = "case"
    end_time: float = float("inf")
    end_step: int = 2 ** 62
    save_path: str = "./results"
    save_dt: Optional[float] = None


class DomainSetup(NamedTuple):
    cells: Tuple[int, int, int]
    ranges: Tuple[Tuple[float, float], ...]


class TransportSetup(NamedTuple):
    """Transport properties (data_types/case_setup/material_properties.py, TransportPropertiesSetup),
    resolved to laws: viscosity mu_ref (T/T_ref)^mu_exp or Sutherland (mu_sutherland = C);
    conductivity cp mu / prandtl, or kappa_ref (T/kappa_T_ref)^kappa_exp or Sutherland
    (kappa_sutherland = C); None where a property is not given."""
    mu_ref: float = 0.0
    T_ref: float = 1.0
    mu_exp: float = 0.0
    mu_sutherland: Optional[float] = None
    bulk_viscosity: float = 0.0
    prandtl: Optional[float] = None
    kappa_ref: Optional[float] = None
    kappa_T_ref: float = 1.0
    kappa_exp: float = 0.0
    kappa_sutherland: Optional[float] = None


class MaterialSetup(NamedTuple):
    model: str = "IdealGas"
    specific_heat_ratio: float = 1.4
    specific_gas_constant: float = 1.0
    transport: Optional[TransportSetup] = None


class CaseSetup(NamedTuple):
    general: GeneralSetup
    domain: DomainSetup
    boundary_conditions: Dict[str, str]
    initial_condition: Dict[str, Callable]
    material: MaterialSetup


# --------------------------------------------------------------- buffers
class MaterialFieldBuffers(NamedTuple):
    """buffers.py:9"""
    conservatives: torch.Tensor
    primitives: torch.Tensor
    temperature: Optional[torch.Tensor] = None


class TimeControlVariables(NamedTuple):
    """buffers.py:88"""
    physical_simulation_time: float
    simulation_step: int
    physical_timestep_size: Any = None     # device scalar tensor (float64) or float
    fixed_timestep_size: Any = None
    end_time: float = float("inf")
    end_step: int = 2 ** 62


class EulerIntegrationBuffers(NamedTuple):
    """buffers.py:103"""
    conservatives: torch.Tensor
    levelset: Any = None
    solid_velocity: Any = None
    solid_energy: Any = None


class IntegrationBuffers(NamedTuple):
    """buffers.py:109"""
    euler_buffers: EulerIntegrationBuffers


class PositivityCounter(NamedTuple):
    interpolation_limiter: Any = None
    thinc_limiter: Any = None
    flux_limiter: Any = None
    acdi_limiter: Any = None
    volume_fraction_limiter: Any = None


class DiscretizationCounter(NamedTuple):
    acdi: Any = None
    thinc: Any = None


class JaxFluidsBuffers(NamedTuple):
    material_fields: MaterialFieldBuffers
    time_control_variables: TimeControlVariables


class FeedForwardSetup(NamedTuple):
    """feed_forward/data_types.py:5-12.  The scan / fori_loop choice and the per-step
    checkpoint flag select how the reference traces the loop; here they change nothing in the
    results.  is_checkpoint_inner_step keeps one state per block of inner_steps and recomputes
    the block once in the backward pass (jax.checkpoint around the block), otherwise every
    step's state is kept."""
    outer_steps: int
    inner_steps: int = 1
    is_scan: bool = True
    is_checkpoint_inner_step: bool = True
    is_checkpoint_integration_step: bool = False
    is_include_t0: bool = True
    is_include_halos: bool = True
