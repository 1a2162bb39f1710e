"""jaxfluids_amd: MI355X-native convective-flux engine with the JAX-Fluids plugin API.

The hot path -- WENO5-JS/-Z (and WENO1, WENO3-JS/-Z, TENO5, TENO5-A) reconstruction (primitive, conservative or
characteristic variables), HLLC / Rusanov Riemann solvers, flux divergence,
TVD-RK3 stage update, face-halo fill and the CFL time step -- runs as
hand-written HIP kernels for gfx950 in ``libjaxfluids_amd.so``, called through a
ctypes C ABI (include/jaxfluids_amd.h).  The Python classes mirror the
reference's plugin interfaces (names, arguments, registries) and hold the state
in PyTorch-ROCm device tensors.
"""
from . import _native
from .config import precision
from .convective_fluxes import DICT_CONVECTIVE_SOLVER, FluxSplittingScheme, HighOrderGodunov
from .data_types import FeedForwardSetup
from .domain import DomainInformation
from .equation_manager import EquationInformation, EquationManager, IdealGas, MaterialManager
from .halos import HaloManager
from .initialization import InitializationManager
from .input_manager import InputManager
from .riemann_solvers import AUSMP, DICT_RIEMANN_SOLVER, DICT_SIGNAL_SPEEDS, HLL, HLLC, HLLCLM, Rusanov
from .simulation_manager import SimulationManager
from .space_solver import SpaceSolver
from .stencils import (DICT_SPATIAL_RECONSTRUCTION, TENO5, TENO5A, TENO6, WENO1, WENO3JS, WENO3Z, WENO5JS, WENO5Z,
                       WENO6CU)
from .time_integration import DICT_TIME_INTEGRATION, RungeKutta3, compute_time_step_size

__version__ = "0.1.0"
