"""Host-side checks of the signal-speed plugin functions (solvers/riemann_solvers/signal_speeds.py):
the registry holds the reference's six entries with the reference's call signatures, and DAVIS2
behaves as in the reference -- callable as a plain function, but a TypeError for the Riemann
solvers that call their signal speed with p_L/p_R (HLLC, HLL, HLLC-LM), since it needs the total
enthalpies H_L/H_R.  No GPU needed."""
import inspect

import pytest

from jaxfluids_amd import _native
from jaxfluids_amd import riemann_solvers as rs

REFERENCE_PARAMS = {  # signal_speeds.py:10,31,51,71,109,135
    "ARITHMETIC": ["u_L", "u_R", "a_L", "a_R", "args", "kwargs"],
    "RUSANOV": ["u_L", "u_R", "a_L", "a_R", "args", "kwargs"],
    "DAVIS": ["u_L", "u_R", "a_L", "a_R", "args", "kwargs"],
    "DAVIS2": ["u_L", "u_R", "a_L", "a_R", "rho_L", "rho_R", "H_L", "H_R", "gamma", "args", "kwargs"],
    "EINFELDT": ["u_L", "u_R", "a_L", "a_R", "rho_L", "rho_R", "args", "kwargs"],
    "TORO": ["u_L", "u_R", "a_L", "a_R", "rho_L", "rho_R", "p_L", "p_R", "gamma", "args", "kwargs"],
}


def test_registry_and_signatures():
    assert list(rs.DICT_SIGNAL_SPEEDS) == ["ARITHMETIC", "RUSANOV", "DAVIS", "DAVIS2", "EINFELDT", "TORO"]
    for name, fn in rs.DICT_SIGNAL_SPEEDS.items():
        assert list(inspect.signature(fn).parameters) == REFERENCE_PARAMS[name], name
        assert rs.signal_speed_name(fn) == name and rs.signal_speed_name(name) == name


@pytest.mark.parametrize("riemann", ["HLLC", "HLL", "HLLC-LM"])
def test_davis2_in_a_riemann_solver_raises_like_the_reference(riemann):
    with pytest.raises(TypeError, match="H_L"):
        _native.make_scheme("PRIMITIVE", "WENO5-JS", riemann, "DAVIS2", "ARITHMETIC", 0, 1.4, 1e-30)
    # the way HLLC calls its signal speed (HLLC.py:93-102) misses the enthalpies in the reference too
    with pytest.raises(TypeError):
        rs.signal_speed_Davis_2(1.0, 1.0, 1.0, 1.0, rho_L=1.0, rho_R=1.0, p_L=1.0, p_R=1.0, gamma=1.4)


@pytest.mark.parametrize("riemann", ["RUSANOV", "AUSMP"])
def test_davis2_is_unused_by_solvers_without_signal_speeds(riemann):
    s = _native.make_scheme("PRIMITIVE", "WENO5-JS", riemann, "DAVIS2", "ARITHMETIC", 0, 1.4, 1e-30)
    assert s.riemann == _native.RIEMANN[riemann]


def test_signal_speeds_need_device_tensors():
    with pytest.raises(_native.JFAError):
        rs.signal_speed_Davis(1.0, 2.0, 1.0, 1.0)
