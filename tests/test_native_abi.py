"""CPU-side checks of the C ABI: the library loads and exports every symbol that
include/jaxfluids_amd.h declares; the ctypes table covers the header exactly.
No compute call is made (there is no GPU here)."""
import ctypes
import os
import re

import pytest

from jaxfluids_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "jaxfluids_amd.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jfa_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_ctypes_table():
    assert header_symbols() == sorted(_native.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason="library not built")
def test_library_exports_all_header_symbols():
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in header_symbols():
        assert hasattr(lib, name), name
    lib = _native.load()
    assert lib.jfa_version().decode().startswith("jaxfluids_amd")


def _c_layout(struct, fields):
    """sizeof and offsetof of a header struct, as the C compiler lays it out."""
    import shutil
    import subprocess
    import tempfile

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ["#include <stddef.h>", "#include <stdio.h>", f'#include "{HEADER}"', "int main(void) {",
             f'    printf("%zu\\n", sizeof({struct}));']
    lines += [f'    printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    lines += ["    return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-std=c99", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), [int(v) for v in out[1:]]


@pytest.mark.parametrize("struct,cls", [("jfa_grid", "JfaGrid"), ("jfa_scheme", "JfaScheme"),
                                        ("jfa_transport", "JfaTransport"), ("jfa_case2d", "JfaCase2D")])
def test_struct_layouts(struct, cls):
    """Every ctypes structure has the size and field offsets the C compiler gives the header's."""
    ct = getattr(_native, cls)
    fields = [f[0] for f in ct._fields_]
    size, offsets = _c_layout(struct, fields)
    assert ctypes.sizeof(ct) == size
    assert [getattr(ct, f).offset for f in fields] == offsets


def test_safe_lambda_parser():
    from jaxfluids_amd.input_manager import parse_lambda
    import numpy as np

    f = parse_lambda("lambda x, y: 1.0*(x <= 0.5) + jnp.sin(2*jnp.pi*y)")
    assert np.isclose(f(0.25, 0.25), 2.0)
    for bad in ("lambda x: __import__('os').system('true')", "lambda x: x.__class__", "lambda x: open('f')",
                "(lambda x: x)(1)"):
        with pytest.raises(ValueError):
            parse_lambda(bad)


def test_numerical_setup_reader_rejects_options_off_the_path():
    from jaxfluids_amd.input_manager import read_numerical_setup

    base = {"conservatives": {"halo_cells": 3, "time_integration": {"integrator": "RK3", "CFL": 0.5},
                              "convective_fluxes": {"convective_solver": "GODUNOV",
                                                    "godunov": {"riemann_solver": "HLLC", "signal_speed": "EINFELDT",
                                                                "reconstruction_stencil": "WENO5-JS",
                                                                "reconstruction_variable": "PRIMITIVE"}}}}
    ns = read_numerical_setup(base)
    assert ns.conservatives.time_integration.CFL == 0.5
    import copy

    bad = copy.deepcopy(base)
    bad["conservatives"]["convective_fluxes"]["godunov"]["riemann_solver"] = "CATUM"
    with pytest.raises(NotImplementedError):
        read_numerical_setup(bad)
    bad = copy.deepcopy(base)
    bad["conservatives"]["halo_cells"] = 2
    with pytest.raises(AssertionError):
        read_numerical_setup(bad)


def test_setup_reader_accepts_further_stencils():
    """WENO1, WENO3-JS/-Z, TENO5 and TENO5-A are read from the setup dicts; the kernels read
    3 cells on each side of a face, so fewer halo cells are refused even where the stencil's
    own requirement is smaller."""
    import copy

    from jaxfluids_amd.input_manager import read_numerical_setup
    from jaxfluids_amd.stencils import DICT_SPATIAL_RECONSTRUCTION

    base = {"conservatives": {"halo_cells": 3, "time_integration": {"integrator": "RK3", "CFL": 0.5},
                              "convective_fluxes": {"convective_solver": "GODUNOV",
                                                    "godunov": {"riemann_solver": "HLLC", "signal_speed": "EINFELDT",
                                                                "reconstruction_stencil": "WENO5-JS",
                                                                "reconstruction_variable": "CHAR-PRIMITIVE"}}}}
    for name in ("WENO1", "WENO3-JS", "WENO3-Z", "TENO5", "TENO5-A", "WENO6-CU", "TENO6", "WENO6-CUM1", "WENO6-CUM2", "TENO6-A"):
        d = copy.deepcopy(base)
        d["conservatives"]["convective_fluxes"]["godunov"]["reconstruction_stencil"] = name
        ns = read_numerical_setup(d)
        assert ns.conservatives.convective_fluxes.godunov.reconstruction_stencil is DICT_SPATIAL_RECONSTRUCTION[name]
    d["conservatives"]["halo_cells"] = 2
    d["conservatives"]["convective_fluxes"]["godunov"]["reconstruction_stencil"] = "WENO3-Z"
    with pytest.raises(NotImplementedError):
        read_numerical_setup(d)


@pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason="library not built")
def test_stage_vjp_refuses_stencils_without_a_reverse_mode():
    """jfa_rk3_stage_vjp validates the scheme before any device call: WENO5-JS/-Z, WENO1,
    WENO3-JS/-Z, TENO5 and TENO5-A have reverse modes (JFA_EINVAL for the NULL buffers); the
    six- and seven-point stencils return JFA_EUNSUPPORTED."""
    lib = _native.load()
    grid = _native.make_grid((8, 8, 8), 3, (0.1, 0.1, 0.1))
    bc = _native.bc_array({f: "PERIODIC" for f in ("east", "west", "north", "south", "top", "bottom")})
    for name, want in (("WENO1", _native.JFA_EINVAL), ("WENO3-JS", _native.JFA_EINVAL),
                       ("WENO3-Z", _native.JFA_EINVAL), ("TENO5", _native.JFA_EINVAL),
                       ("TENO5-A", _native.JFA_EINVAL), ("WENO6-CU", _native.JFA_EUNSUPPORTED),
                       ("TENO6", _native.JFA_EUNSUPPORTED), ("WENO7-JS", _native.JFA_EUNSUPPORTED)):
        scheme = _native.make_scheme("PRIMITIVE", name, "HLLC", "EINFELDT", "ARITHMETIC", _native.F64, 1.4, 1e-30)
        rc = lib.jfa_rk3_stage_vjp(None, None, None, None, grid, scheme, bc, 0, 1e-3, None, 0, None)
        assert rc == want, (name, rc)


@pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason="library not built")
def test_stage_vjp_reconstruction_variables():
    """Reverse modes exist for every reconstruction variable (the characteristic ones with the
    ARITHMETIC or ROE frozen state): each gets as far as the argument check, JFA_EINVAL for the
    NULL buffers."""
    lib = _native.load()
    grid = _native.make_grid((8, 8, 8), 3, (0.1, 0.1, 0.1))
    bc = _native.bc_array({f: "PERIODIC" for f in ("east", "west", "north", "south", "top", "bottom")})
    for var, frozen, want in (("PRIMITIVE", "ROE", _native.JFA_EINVAL),
                              ("CHAR-PRIMITIVE", "ARITHMETIC", _native.JFA_EINVAL),
                              ("CHAR-PRIMITIVE", "ROE", _native.JFA_EINVAL),
                              ("CONSERVATIVE", "ARITHMETIC", _native.JFA_EINVAL),
                              ("CHAR-CONSERVATIVE", "ARITHMETIC", _native.JFA_EINVAL),
                              ("CHAR-CONSERVATIVE", "ROE", _native.JFA_EINVAL)):
        scheme = _native.make_scheme(var, "WENO5-Z", "HLLC", "EINFELDT", frozen, _native.F64, 1.4, 1e-30)
        rc = lib.jfa_rk3_stage_vjp(None, None, None, None, grid, scheme, bc, 0, 1e-3, None, 0, None)
        assert rc == want, (var, frozen, rc)


@pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason="library not built")
def test_stage_vjp_refuses_further_riemann_solvers():
    """The stage VJP has reverse modes for every face flux: HLLC, HLL, HLLC-LM, AUSM+, Rusanov
    and the flux-splitting scheme (JFA_EINVAL for the NULL buffers)."""
    lib = _native.load()
    grid = _native.make_grid((8, 8, 8), 3, (0.1, 0.1, 0.1))
    bc = _native.bc_array({f: "PERIODIC" for f in ("east", "west", "north", "south", "top", "bottom")})
    for rs in ("HLL", "HLLC-LM", "AUSMP", "FLUX-SPLITTING-ROE", "FLUX-SPLITTING-CLLF", "FLUX-SPLITTING-LLF"):
        scheme = _native.make_scheme("PRIMITIVE", "WENO5-JS", rs, "EINFELDT", "ARITHMETIC", _native.F64, 1.4, 1e-30)
        assert lib.jfa_rk3_stage_vjp(None, None, None, None, grid, scheme, bc, 0, 1e-3, None, 0, None) == \
            _native.JFA_EINVAL, rs


@pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason="library not built")
def test_riemann_xi_refuses_flux_splitting():
    """A flux-splitting scheme has no face states: jfa_riemann_xi returns JFA_EUNSUPPORTED
    before any device call; a face-flux value past the last one is refused by every entry point."""
    lib = _native.load()
    for fs in _native.FLUX_SPLITTING:
        scheme = _native.make_scheme("PRIMITIVE", "WENO5-Z", "FLUX-SPLITTING-" + fs, None, "ROE", _native.F64, 1.4,
                                     1e-30)
        assert lib.jfa_riemann_xi(None, None, None, None, 4, 0, scheme, None, None) == _native.JFA_EUNSUPPORTED
    scheme = _native.make_scheme("PRIMITIVE", "WENO5-Z", "HLLC", None, "ROE", _native.F64, 1.4, 1e-30)
    scheme.riemann = max(_native.FLUX_SPLITTING.values()) + 1
    assert lib.jfa_riemann_xi(None, None, None, None, 4, 0, scheme, None, None) == _native.JFA_EUNSUPPORTED
    grid = _native.make_grid((8, 8, 8), 3, (0.1, 0.1, 0.1))
    assert lib.jfa_compute_flux_xi(None, None, grid, scheme, 0, None, None, None) == _native.JFA_EUNSUPPORTED


def test_setup_reader_flux_splitting():
    """convective_solver FLUX-SPLITTING (read_conservatives.py:205-243) builds FluxSplittingScheme;
    CLF is an accepted name without eigenvalues in the reference and is refused."""
    import copy

    from jaxfluids_amd import FluxSplittingScheme
    from jaxfluids_amd.input_manager import read_numerical_setup
    from jaxfluids_amd.stencils import DICT_SPATIAL_RECONSTRUCTION

    base = {"conservatives": {"halo_cells": 3, "time_integration": {"integrator": "RK3", "CFL": 0.5},
                              "convective_fluxes": {"convective_solver": "FLUX-SPLITTING",
                                                    "flux_splitting": {"flux_splitting": "CLLF",
                                                                       "reconstruction_stencil": "TENO5"}}}}
    cf = read_numerical_setup(base).conservatives.convective_fluxes
    assert cf.convective_solver is FluxSplittingScheme and cf.godunov is None
    assert cf.flux_splitting.flux_splitting == "CLLF" and cf.flux_splitting.frozen_state == "ARITHMETIC"
    assert cf.flux_splitting.reconstruction_stencil is DICT_SPATIAL_RECONSTRUCTION["TENO5"]
    bad = copy.deepcopy(base)
    bad["conservatives"]["convective_fluxes"]["flux_splitting"]["flux_splitting"] = "CLF"
    with pytest.raises(NotImplementedError):
        read_numerical_setup(bad)
    bad["conservatives"]["convective_fluxes"]["flux_splitting"]["flux_splitting"] = "GLF"
    with pytest.raises(AssertionError):
        read_numerical_setup(bad)
