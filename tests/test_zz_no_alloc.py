"""jfa_rk3_stage_fused_noalloc (library-owned scratch) against jfa_rk3_stage_fused (caller
workspace): the same kernels, so the results must be bit-identical.  Two RK3 steps in a row, the
second on a smaller grid, so the second call runs on scratch the first one left behind."""
import numpy as np
import pytest
import torch

from oracle import jxf_oracle as o

from .test_gpu_parity import FACES, dev, native_grid, native_scheme, random_state, say

pytestmark = pytest.mark.gpu


def _step(lib, entry, dom, s, W0, U0, with_ws):
    from jaxfluids_amd import _native

    grid = native_grid(dom)
    scheme = native_scheme(s)
    bcs = _native.bc_array({f: (s.bcs[f] if o.FACE_AXIS[f] in dom.active else "PERIODIC") for f in FACES})
    Wa, Ua = dev(W0), dev(U0)
    Wb, Ub, Wc, Uc = (torch.full_like(Wa, float("nan")) for _ in range(4))
    dt = torch.tensor([0.25 * o.compute_dt(W0, dom, s)], dtype=torch.float64, device="cuda")
    dtn = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    nbytes = lib.jfa_stage_workspace_bytes(grid, scheme)
    ws = torch.zeros(nbytes, dtype=torch.uint8, device="cuda") if with_ws else None
    P = _native.ptr
    st = _native.stream_ptr()
    for k, (wi, ui, wo, uo) in enumerate([(Wa, Ua, Wb, Ub), (Wb, Ub, Wc, Uc), (Wc, Uc, Wa, Ua)]):
        args = (P(wi), P(ui), P(Ua if k else None), P(wo), P(uo), grid, scheme, bcs, k, P(dt), 0.0, s.cfl,
                o.EPS64, P(dtn if k == 2 else None))
        rc = entry(*args, P(ws), nbytes, st) if with_ws else entry(*args, st)
        _native.check(rc, "stage")
    torch.cuda.synchronize()
    return Wa.cpu().numpy(), Ua.cpu().numpy(), dtn.cpu().numpy()


@pytest.mark.parametrize("variable,stencil", [("CHAR-PRIMITIVE", "WENO5-Z"), ("PRIMITIVE", "WENO5-JS")])
def test_noalloc_stage_matches_workspace_stage(variable, stencil):
    from jaxfluids_amd import _native

    lib = _native.load()
    s = o.Setup(gamma=1.4, stencil=stencil, variable=variable, riemann="HLLC", signal_speed="EINFELDT",
                bcs={"east": "ZEROGRADIENT", "west": "ZEROGRADIENT", "north": "PERIODIC", "south": "PERIODIC",
                     "top": "SYMMETRY", "bottom": "SYMMETRY"})
    for cells, seed in (((40, 36, 70), 11), ((20, 18, 24), 12)):  # the second call is smaller
        say(f"noalloc stage {variable} {stencil} cells={cells} seed={seed}")
        dom, W = random_state(cells, 5, seed)
        Wb, Ub = o.create_buffers(W, dom, s)
        ref = _step(lib, lib.jfa_rk3_stage_fused, dom, s, Wb, Ub, True)
        got = _step(lib, lib.jfa_rk3_stage_fused_noalloc, dom, s, Wb, Ub, False)
        I = dom.interior
        assert np.all(np.isfinite(ref[0][I]))
        for a, b in zip(ref, got):
            assert np.array_equal(a, b, equal_nan=True)


def test_noalloc_rejects_oversize_grid():
    from jaxfluids_amd import _native

    lib = _native.load()
    s = o.Setup(gamma=1.4, stencil="WENO5-Z", variable="CHAR-PRIMITIVE", riemann="HLLC", signal_speed="EINFELDT")
    dom = o.Domain((128, 128, 128), ((0, 1),) * 3, nh=5)
    grid = native_grid(dom)
    scheme = native_scheme(s)
    assert lib.jfa_stage_workspace_bytes(grid, scheme) > _native.NOALLOC_SCRATCH_BYTES
    bcs = _native.bc_array({f: "PERIODIC" for f in FACES})
    W = torch.ones(1, dtype=torch.float64, device="cuda")  # never touched: the size check comes first
    rc = lib.jfa_rk3_stage_fused_noalloc(_native.ptr(W), _native.ptr(W), None, _native.ptr(W), _native.ptr(W), grid,
                                         scheme, bcs, 0, None, 1e-3, 0.5, o.EPS64, None, _native.stream_ptr())
    assert rc == _native.JFA_EINVAL
    assert b"limit" in lib.jfa_last_error()
