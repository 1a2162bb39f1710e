"""The HDF5 reader and writer (jaxfluids_amd.hdf5) without h5py.

The reader is pinned on the reference's own HDF5 file, tests/golden/blasius_inlet.h5 (the inlet
profile of notebooks/simulations/laminar_boundarylayer, written by h5py): its layout and values
are checked against the example's case_setup.json and the physics of a boundary-layer profile.
The writer is checked by round trips through the reader and against the structures of that
file (superblock, message types and versions, datatype encoding).  libhdf5 itself is not in this
image, so that a file written here opens in h5py is not tested: parity unpinned beyond those
structures.
"""
import os

import numpy as np
import pytest

from jaxfluids_amd import hdf5

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "blasius_inlet.h5")


def test_reader_on_the_reference_inlet_profile():
    """primitives (5, 502) on y (502,): a no-slip wall at y = 0 (u = v = w = 0), p = 1 everywhere,
    and the free stream of case_setup.json's west boundary (rho, u, v, w, p) =
    (1, 2.662235902394827, 0, 0, 1) at the top, y = 10."""
    t = hdf5.read(GOLDEN)
    assert set(t) == {"primitives", "y"}
    P, y = t["primitives"], t["y"]
    assert P.shape == (5, 502) and y.shape == (502,) and P.dtype == np.float64
    assert y[0] == 0.0 and abs(y[-1] - 10.0) < 1e-12 and np.all(np.diff(y) > 0)
    assert np.all(P[1:4, 0] == 0.0)
    assert np.all(P[4] == 1.0)
    assert np.allclose(P[:, -1], [1.0, 2.662235902394827, 0.0, 0.0, 1.0], rtol=0, atol=1e-7)
    assert np.all(np.diff(P[1]) >= -1e-9)            # u rises through the layer (to the profile's rounding)
    assert np.all((P[0] > 0.5) & (P[0] <= 1.0))


def _tree(rng):
    return {
        "time": np.float64(0.125),
        "count": np.int64(7),
        "flag": True,
        "name": "FLUID",
        "names": np.array(["density", "velocity", "pressure"]),
        "f32": rng.standard_normal((3, 4)).astype(np.float32),
        "i32": np.arange(6, dtype=np.int32).reshape(2, 3),
        "bools": np.array([True, False]),
        "empty_group": {},
        "deep": {"a": {"b": {"c": np.linspace(0, 1, 9)}}},
        # more members than one symbol-table node holds (8): several nodes under one B-tree
        "many": {f"q{i:02d}": rng.standard_normal((2, 3, 4)) for i in range(21)},
    }


def _same(a, b, path=""):
    assert set(a) == set(b), (path, set(a), set(b))
    for k, v in a.items():
        if isinstance(v, dict):
            _same(v, b[k], path + k + "/")
            continue
        x, y = np.asarray(v), np.asarray(b[k])
        if x.dtype.kind == "U":
            x = np.char.encode(x, "utf-8")
        assert x.shape == y.shape and x.dtype.kind == y.dtype.kind, (path + k, x.shape, y.shape, x.dtype, y.dtype)
        assert np.array_equal(x, y), path + k


def test_round_trip(tmp_path):
    tree = _tree(np.random.default_rng(0))
    p = str(tmp_path / "t.h5")
    hdf5.write(p, tree)
    _same(tree, hdf5.read(p))


def test_writer_uses_the_structures_of_h5py_files(tmp_path):
    """Superblock version 0 with the golden file's node sizes; a float64 dataset's header holds
    the same message types in the same versions, with a byte-identical float64 datatype and fill
    value message; the group B-tree's keys are the largest name of each symbol-table node (what
    libhdf5's lookup by name bisects on)."""
    g = open(GOLDEN, "rb").read()
    p = str(tmp_path / "s.h5")
    hdf5.write(p, {"primitives": np.zeros((5, 7)), "grp": {f"n{i:02d}": np.float64(i) for i in range(13)}})
    w = open(p, "rb").read()
    assert w[:16] == g[:16]                          # signature, superblock version 0, 8-byte sizes
    assert w[16:20] == g[16:20]                      # leaf K = 4, internal K = 16
    rg, rw = hdf5._Reader(g), hdf5._Reader(w)

    def dataset_msgs(r, name):
        body = next(b for t, b in r.messages(r.root["header"]) if t == hdf5.MSG_SYMBOL_TABLE)
        ents = dict(r.group_entries(int.from_bytes(body[:8], "little"), int.from_bytes(body[8:16], "little")))
        return [(t, b) for t, b in r.messages(ents[name]) if t != hdf5.MSG_NIL]

    mg, mw = dataset_msgs(rg, "primitives"), dataset_msgs(rw, "primitives")
    assert [t for t, _ in mg] == [t for t, _ in mw]
    for (t, bg), (_, bw) in zip(mg, mw):
        assert bg[0] == bw[0], (t, bg[:4].hex(), bw[:4].hex())   # message versions
        if t in (hdf5.MSG_DATATYPE, hdf5.MSG_FILL):
            assert bg == bw
    assert mw[0][1][:8] == mg[0][1][:8] and mw[0][1][8:24] == (5).to_bytes(8, "little") + (7).to_bytes(8, "little")
    # B-tree keys of a group with two symbol-table nodes
    body = next(b for t, b in rw.messages(rw.root["header"]) if t == hdf5.MSG_SYMBOL_TABLE)
    ents = dict(rw.group_entries(int.from_bytes(body[:8], "little"), int.from_bytes(body[8:16], "little")))
    gbody = next(b for t, b in rw.messages(ents["grp"]) if t == hdf5.MSG_SYMBOL_TABLE)
    bt, heap = int.from_bytes(gbody[:8], "little"), int.from_bytes(gbody[8:16], "little")
    used = int.from_bytes(w[bt + 6:bt + 8], "little")
    keys = [rw.heap_string(heap, int.from_bytes(w[bt + 24 + 16 * i:bt + 32 + 16 * i], "little")) for i in range(used + 1)]
    assert used == 2 and keys == ["", "n07", "n12"]


def _domain():
    from jaxfluids_amd.domain import DomainInformation

    return DomainInformation((12, 9, 1), ((0.0, 1.0), (0.0, 0.75), (0.0, 1.0)), 3)


def test_output_layout_and_restart(tmp_path):
    """HDF5Writer's file: domain grids as homogenous.py builds them, fields transposed to
    (Nz, Ny, Nx) with velocity's active components last, temperature p/(rho R); read_restart
    inverts it and checks the cell counts."""
    d = _domain()
    rng = np.random.default_rng(2)
    W = np.full((5, 18, 15, 1), 2.2e-16)
    W[:, 3:-3, 3:-3] = rng.uniform(0.5, 1.5, (5, 12, 9, 1))
    W[3] = 0.0
    U = W * 2.0
    q = {"primitives": ["density", "velocity", "pressure", "temperature"],
         "conservatives": ["mass", "momentum", "energy"]}
    w = hdf5.HDF5Writer(d, gas_constant=0.5, quantities=q)
    path = w.write_file(str(tmp_path), W, U, 0.25)
    assert os.path.basename(path) == "data_0.2500000000.h5"
    t = hdf5.read(path)
    I = (slice(3, -3), slice(3, -3), slice(None))
    assert np.array_equal(t["domain"]["gridX"], np.linspace(1 / 24, 1 - 1 / 24, 12))
    assert np.array_equal(t["domain"]["gridFY"], np.linspace(0.0, 0.75, 10))
    assert t["domain"]["gridZ"].shape == () and t["domain"]["cellsizeY"] == 0.75 / 9
    assert list(t["domain"]["active_axes_indices"]) == [0, 1] and t["domain"]["dim"] == 2
    assert t["primitives"]["density"].shape == (1, 9, 12)
    assert np.array_equal(t["primitives"]["density"], W[0][I].T)
    assert np.array_equal(t["primitives"]["velocity"][..., 1], W[2][I].T)
    assert np.array_equal(t["primitives"]["temperature"], (W[4][I] / (W[0][I] * 0.5)).T)
    assert np.array_equal(t["conservatives"]["momentum"][..., 0], U[1][I].T)
    assert t["metadata"]["levelset_model"] is np.False_ or t["metadata"]["levelset_model"] == False  # noqa: E712
    assert t["time"] == 0.25
    time, prim = hdf5.read_restart(path, d)
    assert time == 0.25 and np.array_equal(prim, W[(slice(None),) + I])
    from jaxfluids_amd.domain import DomainInformation

    with pytest.raises(AssertionError, match="Number of cells"):
        hdf5.read_restart(path, DomainInformation((12, 10, 1), ((0.0, 1.0), (0.0, 0.75), (0.0, 1.0)), 3))


def test_xdmf_companion(tmp_path):
    import xml.dom.minidom

    d = _domain()
    x = hdf5.XDMFWriter(d, {"primitives": ["density", "velocity"]})
    p = x.write_file(str(tmp_path / "data_0.5000000000.h5"), 0.5)
    doc = xml.dom.minidom.parse(p)
    items = [n.firstChild.data for n in doc.getElementsByTagName("DataItem")]
    assert items == ["data_0.5000000000.h5:domain/gridFX", "data_0.5000000000.h5:domain/gridFY",
                     "data_0.5000000000.h5:domain/gridFZ", "data_0.5000000000.h5:primitives/density",
                     "data_0.5000000000.h5:primitives/velocity"]
    dims = [n.getAttribute("Dimensions") for n in doc.getElementsByTagName("DataItem")]
    assert dims[3:] == ["1 9 12", "1 9 12 2"]
    assert doc.getElementsByTagName("Topology")[0].getAttribute("Dimensions") == "13 10 2"
    xml.dom.minidom.parse(x.write_timeseries(str(tmp_path)))
