"""CPU checks of the batched 2-D stage kernel (jfa_batch.h) through a host build of the library
sources (tests/emu): the real plan/stage C ABI and kernel code, run on host buffers, against
the oracle case by case.

Case blocks sit in the pooled buffers with unregistered gaps between them, and every state
load and store of the kernel is checked against the registered blocks, so an out-of-bounds
index shows up as a violation here instead of a fault on the GPU.  Emulated wavefronts run
64 host threads in lock step at every shuffle, so these cases are small.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from oracle import jxf_oracle as o

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "emu"))

TOL = 1e-12
GAP = 977  # unregistered elements between case blocks

CASES = [
    ((37, 70), ((0.0, 1.0), (0.0, 1.3)), {"east": "PERIODIC", "west": "PERIODIC", "north": "PERIODIC", "south": "PERIODIC"}),
    ((5, 9), ((0.0, 0.5), (0.0, 0.9)), {"east": "PERIODIC", "west": "PERIODIC", "north": "ZEROGRADIENT", "south": "SYMMETRY"}),
    ((21, 64), ((0.0, 2.0), (0.0, 2.0)), {f: "ZEROGRADIENT" for f in ("east", "west", "north", "south")}),
    ((9, 67), ((0.0, 1.0), (0.0, 1.0)), {f: "SYMMETRY" for f in ("east", "west", "north", "south")}),
    ((70, 3), ((0.0, 1.0), (0.0, 0.1)), {"east": "ZEROGRADIENT", "west": "SYMMETRY", "north": "ZEROGRADIENT", "south": "ZEROGRADIENT"}),
]


@pytest.fixture(scope="module")
def emu():
    import build as emu_build
    from jaxfluids_amd import _native

    lib = ctypes.CDLL(emu_build.build())
    for name in ("jfa_batch2d_workspace_bytes", "jfa_batch2d_plan", "jfa_batch2d_stage"):
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = _native.SIGNATURES[name]
    lib.emu_add_range.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    lib.emu_violations.restype = ctypes.c_longlong
    lib.emu_accesses.restype = ctypes.c_longlong
    lib.emu_first_violation.restype = ctypes.c_ulonglong
    return lib


def case_state(cells, ranges, seed):
    rng = np.random.default_rng(seed)
    nx, ny = cells
    x = ranges[0][0] + (np.arange(nx) + 0.5) * (ranges[0][1] - ranges[0][0]) / nx
    y = ranges[1][0] + (np.arange(ny) + 0.5) * (ranges[1][1] - ranges[1][0]) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 6)
    k = 2 * np.pi
    rho = 1.0 + 0.3 * np.sin(k * X + ph[0]) * np.cos(k * Y + ph[1])
    u = 0.3 * np.sin(k * Y + ph[2]) + rng.uniform(-0.3, 0.3)
    v = 0.3 * np.cos(k * X + ph[3]) + rng.uniform(-0.3, 0.3)
    p = 1.0 + 0.2 * np.cos(k * X + k * Y + ph[4])
    m = (X + 0.3 * Y) > 0.55 * (ranges[0][1] + 0.3 * ranges[1][1])
    rho = np.where(m, 0.5 * rho, rho)
    p = np.where(m, 0.4 * p, p)
    W = np.stack([rho, u, v, np.zeros_like(X), p])[..., None]
    return W * (1.0 + 1e-3 * rng.standard_normal(W.shape))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


class HostEnsemble:
    """Pooled host buffers + case table in the layout Ensemble2D uses, with gaps."""

    def __init__(self, lib, scheme, nh=3, chunk=16, dtype=np.float64):
        from jaxfluids_amd import _native

        self.lib, self.nh, self.chunk, self.dtype = lib, nh, chunk, dtype
        st, var, rs, ss, fz = scheme
        self.scheme = _native.make_scheme(var, st, rs, ss, fz, _native.F64 if dtype == np.float64 else _native.F32,
                                          1.4, o.SPATIAL_STENCIL_EPS)
        self.table = (_native.JfaCase2D * len(CASES))()
        self.blocks = []
        off = GAP
        for c, (cells, ranges, bcs) in enumerate(CASES):
            e = self.table[c]
            e.n[0], e.n[1], e.nh = cells[0], cells[1], nh
            for i, f in enumerate(("east", "west", "north", "south")):
                e.bc[i] = _native.BC[bcs[f]]
            e.dx[0] = (ranges[0][1] - ranges[0][0]) / cells[0]
            e.dx[1] = (ranges[1][1] - ranges[1][0]) / cells[1]
            e.offset = off
            n = 5 * (cells[0] + 2 * nh) * (cells[1] + 2 * nh)
            self.blocks.append((off, n, (5, cells[0] + 2 * nh, cells[1] + 2 * nh, 1)))
            off += n + GAP
        self.total = off
        self.pools = [np.full(self.total, np.nan, dtype=dtype) for _ in range(6)]  # A.W A.U B.W B.U C.W C.U
        lib.emu_clear_ranges()
        for pool in self.pools:
            for boff, n, _ in self.blocks:
                lib.emu_add_range(pool.ctypes.data + boff * pool.itemsize, n * pool.itemsize)
        nbytes = lib.jfa_batch2d_workspace_bytes(ctypes.byref(self.table), len(CASES), chunk)
        assert nbytes > 0
        self.ws = np.zeros(nbytes + 256, dtype=np.uint8)
        self.ws_ptr = (self.ws.ctypes.data + 255) // 256 * 256
        self.nbytes = nbytes
        assert lib.jfa_batch2d_plan(ctypes.byref(self.table), len(CASES), chunk, self.ws_ptr, nbytes, None) == 0
        self.dt = np.zeros(len(CASES))

    def view(self, pool, c):
        off, n, shape = self.blocks[c]
        return self.pools[pool][off:off + n].reshape(shape)

    def stage(self, src, dst, stage, dt_next):
        P = lambda i: self.pools[i].ctypes.data
        rc = self.lib.jfa_batch2d_stage(P(2 * src), P(2 * src + 1), P(1) if stage else None, P(2 * dst),
                                        P(2 * dst + 1), ctypes.byref(self.table), len(CASES), self.chunk, self.scheme,
                                        stage, self.dt.ctypes.data, 0.5, float(np.finfo(self.dtype).eps),
                                        self.dt.ctypes.data if dt_next else None, self.ws_ptr, self.nbytes, None)
        assert rc == 0

    def step(self):
        self.stage(0, 1, 0, False)
        self.stage(1, 2, 1, False)
        self.stage(2, 0, 2, True)


def oracle_case(c, scheme, prim, nh=3):
    cells, ranges, bcs = CASES[c]
    dom = o.Domain((cells[0], cells[1], 1), (ranges[0], ranges[1], (0.0, 1.0)), nh=nh)
    st, var, rs, ss, fz = scheme
    s = o.Setup(gamma=1.4, stencil=st, variable=var, riemann=rs, signal_speed=ss, frozen_state=fz,
                bcs={**bcs, "top": "PERIODIC", "bottom": "PERIODIC"})
    W, U = o.create_buffers(prim, dom, s)
    return dom, s, W, U


def face_regions(dom):
    out = [dom.interior]
    for a in range(2):
        sl = [dom.dom(0), dom.dom(1), slice(None)]
        for part in (slice(0, dom.nh), slice(-dom.nh, None)):
            sl[a] = part
            out.append((Ellipsis,) + tuple(sl))
    return out


SCHEMES = [
    ("WENO5-JS", "PRIMITIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
    ("WENO5-Z", "CHAR-PRIMITIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
    ("WENO5-Z", "CONSERVATIVE", "RUSANOV", "EINFELDT", "ARITHMETIC"),
    ("WENO5-JS", "CHAR-CONSERVATIVE", "HLLC", "TORO", "ROE"),
]


FURTHER = [  # the run-time stencil switch of the kernel
    ("TENO5", "CHAR-PRIMITIVE", "HLLC", "EINFELDT", "ROE"),
    ("TENO5-A", "PRIMITIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
    ("WENO3-Z", "CHAR-PRIMITIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
    ("WENO3-JS", "CONSERVATIVE", "RUSANOV", "EINFELDT", "ARITHMETIC"),
    ("WENO1", "PRIMITIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
    ("WENO6-CU", "CHAR-PRIMITIVE", "HLLC", "EINFELDT", "ROE"),
    ("TENO6", "PRIMITIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
    ("WENO5-JS", "CHAR-PRIMITIVE", "HLL", "EINFELDT", "ARITHMETIC"),  # HLL Riemann solver
    ("TENO5", "PRIMITIVE", "HLLC-LM", "DAVIS", "ARITHMETIC"),  # HLLC-LM Riemann solver
    ("WENO5-Z", "CONSERVATIVE", "AUSMP", "EINFELDT", "ARITHMETIC"),  # AUSM+ Riemann solver
    ("WENO5-Z", "PRIMITIVE", "FLUX-SPLITTING-ROE", "EINFELDT", "ROE"),  # flux splitting
    ("TENO5", "CHAR-CONSERVATIVE", "FLUX-SPLITTING-LLF", "EINFELDT", "ARITHMETIC"),
    ("WENO5-JS", "PRIMITIVE", "FLUX-SPLITTING-CLLF", "EINFELDT", "ARITHMETIC"),
    ("WENO6-CUM1", "CHAR-PRIMITIVE", "HLLC", "EINFELDT", "ROE"),  # the cell size per case
    ("WENO6-CUM2", "CONSERVATIVE", "HLLC", "EINFELDT", "ARITHMETIC"),
    ("TENO6-A", "CHAR-CONSERVATIVE", "HLLC", "EINFELDT", "ROE"),
]


@pytest.mark.parametrize("scheme,nh", [(s, 3) for s in SCHEMES + FURTHER] + [(SCHEMES[1], 5), (SCHEMES[0], 4)],
                         ids=lambda v: "-".join(v) if isinstance(v, tuple) else f"nh{v}")
def test_emulated_batch_step_matches_oracle(emu, scheme, nh):
    """One RK3 step of 5 uneven cases (ragged x chunks and y tiles, every boundary type):
    every load/store in bounds, state and next dt per case equal to the oracle's."""
    prims = [case_state(c[0], c[1], seed=700 + i) for i, c in enumerate(CASES)]
    ens = HostEnsemble(emu, scheme, nh=nh, chunk=16)
    refs = [oracle_case(i, scheme, p, nh) for i, p in enumerate(prims)]
    for i, (dom, s, W, U) in enumerate(refs):
        ens.view(0, i)[...] = W
        ens.view(1, i)[...] = U
        ens.dt[i] = o.compute_dt(W, dom, s)
    dt0 = ens.dt.copy()
    ens.step()
    assert emu.emu_violations() == 0, hex(emu.emu_first_violation())
    assert emu.emu_accesses() > 0
    for i, (dom, s, W, U) in enumerate(refs):
        Wr, Ur = o.rk3_step(W, U, dt0[i], dom, s)
        Wg, Ug = ens.view(0, i), ens.view(1, i)
        for r in face_regions(dom):
            assert rel(Wg[r], Wr[r]) < TOL, (i, rel(Wg[r], Wr[r]))
            assert rel(Ug[r], Ur[r]) < TOL, (i, rel(Ug[r], Ur[r]))
        dtr = o.compute_dt(Wr, dom, s)
        assert abs(ens.dt[i] - dtr) <= 1e-14 * dtr, (i, ens.dt[i], dtr)


@pytest.mark.parametrize("fs", ["ROE", "CLLF"])
def test_emulated_batch_flux_splitting_fp32_storage(emu, fs):
    """The FP32-storage batched stage with flux splitting (float pools, FP64 arithmetic) against
    the FP64 oracle after one step: the float32 rounding of the state, < 1e-6 relative L-inf."""
    scheme = ("WENO5-Z", "PRIMITIVE", "FLUX-SPLITTING-" + fs, "EINFELDT", "ROE")
    prims = [case_state(c[0], c[1], seed=800 + i) for i, c in enumerate(CASES)]
    ens = HostEnsemble(emu, scheme, nh=3, chunk=16, dtype=np.float32)
    refs = [oracle_case(i, scheme, p) for i, p in enumerate(prims)]
    for i, (dom, s, W, U) in enumerate(refs):
        ens.view(0, i)[...] = W
        ens.view(1, i)[...] = U
        ens.dt[i] = o.compute_dt(W, dom, s)
    dt0 = ens.dt.copy()
    ens.step()
    assert emu.emu_violations() == 0, hex(emu.emu_first_violation())
    for i, (dom, s, W, U) in enumerate(refs):
        Wr, _ = o.rk3_step(W, U, dt0[i], dom, s)
        err = rel(ens.view(0, i)[dom.interior].astype(np.float64), Wr[dom.interior])
        assert 1e-9 < err < 1e-6, (i, err)
        assert abs(ens.dt[i] - o.compute_dt(Wr, dom, s)) <= 1e-6 * dt0[i], i


def test_emulated_batch_chunks_agree(emu):
    """Different x-chunk lengths give the same state bit for bit."""
    scheme = SCHEMES[1]
    prims = [case_state(c[0], c[1], seed=800 + i) for i, c in enumerate(CASES)]
    outs = []
    for chunk in (64, 7):
        ens = HostEnsemble(emu, scheme, chunk=chunk)
        for i, p in enumerate(prims):
            dom, s, W, U = oracle_case(i, scheme, p)
            ens.view(0, i)[...] = W
            ens.view(1, i)[...] = U
            ens.dt[i] = o.compute_dt(W, dom, s)
        ens.step()
        assert emu.emu_violations() == 0, hex(emu.emu_first_violation())
        outs.append([ens.view(0, i).copy() for i in range(len(CASES))])
    for a, b in zip(*outs):
        I = (slice(None), slice(3, -3), slice(3, -3))
        assert np.array_equal(a[I], b[I])


def test_emulated_batch_rejects_bad_tables(emu):
    """Host-side validation: chunk range, too-small halos, periodic/symmetric axes shorter than
    the halo, and blocks beyond the 32-bit lane-offset limit."""
    from jaxfluids_amd import _native

    t = (_native.JfaCase2D * 1)()
    t[0].n[0], t[0].n[1], t[0].nh = 10, 10, 3
    t[0].dx[0] = t[0].dx[1] = 0.1
    for f in range(4):
        t[0].bc[f] = _native.BC["PERIODIC"]
    assert emu.jfa_batch2d_workspace_bytes(ctypes.byref(t), 1, 16) > 0
    assert emu.jfa_batch2d_workspace_bytes(ctypes.byref(t), 1, 0) == 0
    assert emu.jfa_batch2d_workspace_bytes(ctypes.byref(t), 1, 65) == 0
    t[0].nh = 2
    assert emu.jfa_batch2d_workspace_bytes(ctypes.byref(t), 1, 16) == 0
    t[0].nh, t[0].n[1] = 3, 2
    assert emu.jfa_batch2d_workspace_bytes(ctypes.byref(t), 1, 16) == 0
    t[0].n[0], t[0].n[1] = 20000, 20000
    assert emu.jfa_batch2d_workspace_bytes(ctypes.byref(t), 1, 16) == 0


def test_emulated_batch_time_step_propagates_nan(emu):
    """A NaN cell in one case makes that case's next dt NaN (the reference's jnp.max propagates
    NaN, time_step_size.py:109); the other cases keep the oracle's dt."""
    scheme = SCHEMES[0]
    prims = [case_state(c[0], c[1], seed=900 + i) for i, c in enumerate(CASES)]
    ens = HostEnsemble(emu, scheme, chunk=16)
    refs = [oracle_case(i, scheme, p) for i, p in enumerate(prims)]
    bad = 2
    for i, (dom, s, W, U) in enumerate(refs):
        ens.view(0, i)[...] = W
        ens.view(1, i)[...] = U
        ens.dt[i] = o.compute_dt(W, dom, s)
    dom = refs[bad][0]
    ens.view(0, bad)[0, dom.nh + 3, dom.nh + 5, 0] = np.nan
    ens.view(1, bad)[0, dom.nh + 3, dom.nh + 5, 0] = np.nan
    dt0 = ens.dt.copy()
    ens.step()
    assert emu.emu_violations() == 0, hex(emu.emu_first_violation())
    for i, (dom, s, W, U) in enumerate(refs):
        if i == bad:
            assert np.isnan(ens.dt[i])
            continue
        dtr = o.compute_dt(o.rk3_step(W, U, dt0[i], dom, s)[0], dom, s)
        assert abs(ens.dt[i] - dtr) <= 1e-14 * dtr, (i, ens.dt[i], dtr)
