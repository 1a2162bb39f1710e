"""HDF5 files of the simulation output and restart, without h5py.

The reference writes its output with h5py (``io_utils/hdf5_writer.py:79-266``): groups
``metadata/``, ``domain/``, ``wall_clock/``, ``primitives/`` (and ``conservatives/``) and the
dataset ``time``; it restarts from such a file (``initialization/material_fields_initializer.py:
212``) and reads boundary profiles from HDF5 files (``blasius_inlet.h5`` of the laminar
boundary-layer example).  h5py and libhdf5 are not part of this build, so this module reads and
writes the file format itself, limited to what those files use:

* superblock version 0 with 8-byte offsets and lengths;
* version-1 object headers (with continuation blocks);
* groups as symbol tables: a version-1 group B-tree, symbol-table nodes and a local heap;
* datasets with a simple or scalar dataspace (versions 1 and 2), little-endian IEEE floats,
  integers, h5py's boolean enumeration and strings (fixed-length; variable-length through the
  global heap on reading only), in compact, contiguous or chunked (uncompressed) layout;
* attributes are skipped on reading and not written.

``read(path)`` returns a nested dict of numpy arrays; ``write(path, tree)`` writes a nested dict
of numpy arrays and scalars in the same structures h5py produces for them.  Field arrays are
stored as the reference stores them: the interior cells in (Nz, Ny, Nx) order
(``hdf5_writer.py:295-333`` transposes (x, y, z) buffers with ``jnp.swapaxes(buffer, -1, -3)``).
"""
from __future__ import annotations

import struct
from typing import Any, Dict, List, Optional, Tuple

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF

# object header message types
MSG_NIL, MSG_DATASPACE, MSG_DATATYPE, MSG_FILL_OLD, MSG_FILL = 0x0000, 0x0001, 0x0003, 0x0004, 0x0005
MSG_LAYOUT, MSG_ATTRIBUTE, MSG_CONTINUATION, MSG_SYMBOL_TABLE, MSG_MTIME = 0x0008, 0x000C, 0x0010, 0x0011, 0x0012


class HDF5Error(ValueError):
    """The file uses a part of the HDF5 format this module does not read."""


# ===================================================================== reading
class _Reader:
    def __init__(self, data: bytes) -> None:
        if data[:8] != SIGNATURE:
            raise HDF5Error("not an HDF5 file (signature)")
        self.d = data
        version = data[8]
        if version != 0:
            raise HDF5Error(f"superblock version {version} (only version 0 is read)")
        self.so, self.sl = data[13], data[14]
        if self.so != 8 or self.sl != 8:
            raise HDF5Error("only 8-byte offsets and lengths are read")
        self.base = self._u(24, 8)
        # root group symbol table entry at byte 56
        self.root = self._entry(56)

    def _u(self, at: int, n: int) -> int:
        return int.from_bytes(self.d[at:at + n], "little")

    def _entry(self, at: int) -> Dict[str, int]:
        e = {"name_off": self._u(at, 8), "header": self._u(at + 8, 8), "cache": self._u(at + 16, 4)}
        if e["cache"] == 1:
            e["btree"] = self._u(at + 24, 8)
            e["heap"] = self._u(at + 32, 8)
        return e

    # -- structures
    def heap_data(self, addr: int) -> Tuple[int, int]:
        if self.d[addr:addr + 4] != b"HEAP":
            raise HDF5Error("local heap signature")
        size = self._u(addr + 8, 8)
        data_addr = self._u(addr + 24, 8)
        return data_addr, size

    def heap_string(self, heap_addr: int, off: int) -> str:
        data_addr, _ = self.heap_data(heap_addr)
        a = data_addr + off
        b = self.d.index(b"\0", a)
        return self.d[a:b].decode("utf-8")

    def group_entries(self, btree: int, heap: int) -> List[Tuple[str, int]]:
        out: List[Tuple[str, int]] = []
        self._walk(btree, heap, out)
        return out

    def _walk(self, node: int, heap: int, out: List[Tuple[str, int]]) -> None:
        if self.d[node:node + 4] != b"TREE":
            raise HDF5Error("group B-tree signature")
        if self.d[node + 4] != 0:
            raise HDF5Error("not a group B-tree")
        level = self.d[node + 5]
        used = self._u(node + 6, 2)
        at = node + 24  # after left/right sibling addresses
        children = []
        for _ in range(used):
            at += 8  # key (heap offset)
            children.append(self._u(at, 8))
            at += 8
        for ch in children:
            if level > 0:
                self._walk(ch, heap, out)
            else:
                self._snod(ch, heap, out)

    def _snod(self, addr: int, heap: int, out: List[Tuple[str, int]]) -> None:
        if self.d[addr:addr + 4] != b"SNOD":
            raise HDF5Error("symbol table node signature")
        n = self._u(addr + 6, 2)
        for i in range(n):
            e = self._entry(addr + 8 + 40 * i)
            out.append((self.heap_string(heap, e["name_off"]), e["header"]))

    def messages(self, addr: int) -> List[Tuple[int, bytes]]:
        d = self.d
        if d[addr] != 1:
            raise HDF5Error(f"object header version {d[addr]} (only version 1 is read)")
        nmsg = self._u(addr + 2, 2)
        size = self._u(addr + 8, 4)
        blocks = [(addr + 16, size)]
        msgs: List[Tuple[int, bytes]] = []
        while blocks and len(msgs) < nmsg:
            at, length = blocks.pop(0)
            end = at + length
            while at + 8 <= end and len(msgs) < nmsg:
                mtype, msize = self._u(at, 2), self._u(at + 2, 2)
                body = d[at + 8:at + 8 + msize]
                at += 8 + msize
                if mtype == MSG_CONTINUATION:
                    blocks.append((int.from_bytes(body[:8], "little"), int.from_bytes(body[8:16], "little")))
                msgs.append((mtype, body))
        return msgs

    # -- objects
    def read_object(self, addr: int) -> Any:
        msgs = self.messages(addr)
        types = {t for t, _ in msgs}
        if MSG_SYMBOL_TABLE in types:
            body = next(b for t, b in msgs if t == MSG_SYMBOL_TABLE)
            btree, heap = int.from_bytes(body[:8], "little"), int.from_bytes(body[8:16], "little")
            return {name: self.read_object(h) for name, h in self.group_entries(btree, heap)}
        if MSG_DATASPACE in types and MSG_DATATYPE in types and MSG_LAYOUT in types:
            shape = _dataspace(next(b for t, b in msgs if t == MSG_DATASPACE))
            dtype = _datatype(next(b for t, b in msgs if t == MSG_DATATYPE))
            return self._dataset(next(b for t, b in msgs if t == MSG_LAYOUT), shape, dtype)
        raise HDF5Error(f"object at {addr} is neither a group nor a dataset")

    def _dataset(self, layout: bytes, shape: Tuple[int, ...], dtype: np.dtype) -> np.ndarray:
        n = int(np.prod(shape)) if shape else 1
        nbytes = n * (16 if dtype.kind == "O" else dtype.itemsize)
        version = layout[0]
        if version == 3:
            cls = layout[1]
            if cls == 0:  # compact
                size = int.from_bytes(layout[2:4], "little")
                raw = layout[4:4 + size]
            elif cls == 1:  # contiguous
                addr = int.from_bytes(layout[2:10], "little")
                raw = b"" if addr == UNDEF else self.d[self.base + addr:self.base + addr + nbytes]
                if addr == UNDEF:
                    raw = b"\0" * nbytes
            elif cls == 2:  # chunked
                rank = layout[2]
                addr = int.from_bytes(layout[3:11], "little")
                cdims = [int.from_bytes(layout[11 + 4 * i:15 + 4 * i], "little") for i in range(rank)]
                return self._chunked(addr, shape, cdims[:-1], dtype)
            else:
                raise HDF5Error(f"layout class {cls}")
        elif version in (1, 2):
            rank = layout[1]
            cls = layout[2]
            at = 8
            if cls == 0:
                dims = [int.from_bytes(layout[at + 4 * i:at + 4 * i + 4], "little") for i in range(rank)]
                at += 4 * rank
                size = int.from_bytes(layout[at:at + 4], "little")
                raw = layout[at + 4:at + 4 + size]
            else:
                addr = int.from_bytes(layout[at:at + 8], "little")
                if cls == 2:
                    cdims = [int.from_bytes(layout[at + 8 + 4 * i:at + 12 + 4 * i], "little") for i in range(rank)]
                    return self._chunked(addr, shape, cdims[:-1], dtype)
                raw = self.d[self.base + addr:self.base + addr + nbytes]
        else:
            raise HDF5Error(f"layout message version {version}")
        if dtype.kind == "O":  # variable-length strings: (length, collection address, index) each
            arr = np.array([self._vlen(raw[16 * i:16 * i + 16]) for i in range(n)], dtype=object)
        else:
            arr = np.frombuffer(raw[:nbytes], dtype=dtype).copy()
        return arr.reshape(shape) if shape else arr.reshape(())[()]

    def _vlen(self, ref: bytes) -> bytes:
        length = int.from_bytes(ref[0:4], "little")
        coll = int.from_bytes(ref[4:12], "little")
        index = int.from_bytes(ref[12:16], "little")
        if self.d[coll:coll + 4] != b"GCOL":
            raise HDF5Error("global heap collection signature")
        end = coll + int.from_bytes(self.d[coll + 8:coll + 16], "little")
        at = coll + 16
        while at + 16 <= end:
            idx = int.from_bytes(self.d[at:at + 2], "little")
            size = int.from_bytes(self.d[at + 8:at + 16], "little")
            if idx == index:
                return bytes(self.d[at + 16:at + 16 + min(size, length)])
            if idx == 0:
                break
            at += 16 + (size + 7) // 8 * 8
        raise HDF5Error("global heap object not found")

    def _chunked(self, btree: int, shape, cdims, dtype: np.dtype) -> np.ndarray:
        out = np.zeros(shape, dtype=dtype)
        rank = len(shape)
        self._chunks(btree, rank, cdims, dtype, out)
        return out

    def _chunks(self, node: int, rank: int, cdims, dtype: np.dtype, out: np.ndarray) -> None:
        d = self.d
        if d[node:node + 4] != b"TREE" or d[node + 4] != 1:
            raise HDF5Error("raw-data chunk B-tree")
        level, used = d[node + 5], self._u(node + 6, 2)
        at = node + 24
        keysize = 8 + 8 * (rank + 1)
        for _ in range(used):
            csize, mask = self._u(at, 4), self._u(at + 4, 4)
            offs = [self._u(at + 8 + 8 * i, 8) for i in range(rank)]
            child = self._u(at + keysize, 8)
            at += keysize + 8
            if level > 0:
                self._chunks(child, rank, cdims, dtype, out)
                continue
            if mask:
                raise HDF5Error("filtered (compressed) chunks are not read")
            chunk = np.frombuffer(d[self.base + child:self.base + child + csize], dtype=dtype)
            chunk = chunk[:int(np.prod(cdims))].reshape(cdims)
            sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, cdims, out.shape))
            out[sl] = chunk[tuple(slice(0, s.stop - s.start) for s in sl)]


def _dataspace(b: bytes) -> Tuple[int, ...]:
    version, rank, flags = b[0], b[1], b[2]
    if version == 1:
        at = 8
    elif version == 2:
        if b[3] == 2:  # null dataspace
            return (0,)
        at = 4
    else:
        raise HDF5Error(f"dataspace version {version}")
    return tuple(int.from_bytes(b[at + 8 * i:at + 8 * i + 8], "little") for i in range(rank))


def _datatype(b: bytes) -> np.dtype:
    cls = b[0] & 0x0F
    bits0 = b[1]
    size = int.from_bytes(b[4:8], "little")
    order = ">" if bits0 & 1 else "<"
    if cls == 1:  # floating point
        if size not in (2, 4, 8):
            raise HDF5Error(f"float of {size} bytes")
        return np.dtype(f"{order}f{size}")
    if cls == 0:  # fixed point
        signed = bool(bits0 & 0x08)
        return np.dtype(f"{order}{'i' if signed else 'u'}{size}")
    if cls == 3:  # fixed-length string
        return np.dtype(f"S{size}")
    if cls == 8:  # enumeration: h5py stores numpy bools as int8 {FALSE: 0, TRUE: 1}
        nmem = b[1] | (b[2] << 8)
        base = _datatype(b[8:])
        at = 8 + 12  # an integer base type: 8-byte header + bit offset and precision
        names = []
        for _ in range(nmem):
            end = b.index(b"\0", at)
            names.append(b[at:end].decode("utf-8"))
            at += (end - at + 1 + 7) // 8 * 8
        if names == ["FALSE", "TRUE"] and size == 1:
            return np.dtype(np.bool_)
        return base
    if cls == 9 and (bits0 & 0x0F) == 1:  # variable-length string
        return np.dtype("O")
    raise HDF5Error(f"datatype class {cls} (sequences and compound types are not read)")


def read(path: str) -> Dict[str, Any]:
    """The whole file as a nested dict: groups are dicts, datasets numpy arrays (scalars as numpy
    scalars, fixed-length strings as bytes)."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    return r.read_object(r.root["header"])


# ===================================================================== writing
LEAF_K, INTERNAL_K = 4, 16  # group symbol-table node and B-tree node sizes (as h5py's default files)
_FILL_V2 = bytes.fromhex("0202020100000000")  # fill value message as h5py writes it (late allocation, undefined)


def _pad8(b: bytes) -> bytes:
    return b + b"\0" * (-len(b) % 8)


class _Writer:
    def __init__(self) -> None:
        self.buf = bytearray(96)  # superblock, written last

    def alloc(self, data: bytes, align: int = 8) -> int:
        self.buf += b"\0" * (-len(self.buf) % align)
        at = len(self.buf)
        self.buf += data
        return at

    # -- object header version 1 from (type, body) messages
    def header(self, msgs: List[Tuple[int, bytes]]) -> int:
        body = b""
        for mtype, data in msgs:
            data = _pad8(data)
            body += struct.pack("<HHB3x", mtype, len(data), 0) + data
        head = struct.pack("<BBHII", 1, 0, len(msgs), 1, len(body)) + b"\0" * 4
        return self.alloc(head + body)

    # -- a dataset
    def dataset(self, value) -> int:
        arr = np.asarray(value)
        is_bool = arr.dtype == np.bool_
        if is_bool:
            arr = arr.astype(np.int8)
        if arr.dtype.kind == "U":
            arr = np.char.encode(arr, "utf-8")
        if arr.dtype.kind in "iu" and arr.dtype.itemsize not in (1, 2, 4, 8):
            arr = arr.astype(np.int64)
        # copy(order="C") keeps 0-d arrays 0-d (ascontiguousarray would make them 1-d)
        arr = (arr.astype(arr.dtype.newbyteorder("<")) if arr.dtype.kind in "fiu" else arr).copy(order="C")
        shape = arr.shape
        space = struct.pack("<BBB5x", 1, len(shape), 1 if shape else 0)
        space += b"".join(struct.pack("<Q", n) for n in shape) * (2 if shape else 1)  # dims and max dims
        if not shape:
            space = struct.pack("<BBB5x", 1, 0, 0)
        dt = arr.dtype
        if dt.kind == "f":
            if dt.itemsize == 8:
                dtype = struct.pack("<BBBBI", 0x11, 0x20, 63, 0, 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
            elif dt.itemsize == 4:
                dtype = struct.pack("<BBBBI", 0x11, 0x20, 31, 0, 4) + struct.pack("<HHBBBBI", 0, 32, 23, 8, 0, 23, 127)
            else:
                raise HDF5Error(f"float of {dt.itemsize} bytes")
        elif is_bool:  # h5py's boolean: an enumeration of int8 with FALSE = 0, TRUE = 1
            base = struct.pack("<BBBBI", 0x10, 0x08, 0, 0, 1) + struct.pack("<HH", 0, 8)
            dtype = struct.pack("<BBBBI", 0x18, 2, 0, 0, 1) + base + _pad8(b"FALSE\0") + _pad8(b"TRUE\0") + bytes([0, 1])
        elif dt.kind in "iu":
            dtype = struct.pack("<BBBBI", 0x10, 0x08 if dt.kind == "i" else 0, 0, 0, dt.itemsize)
            dtype += struct.pack("<HH", 0, 8 * dt.itemsize)
        elif dt.kind == "S":
            dtype = struct.pack("<BBBBI", 0x13, 0x00, 0, 0, max(dt.itemsize, 1))  # null-terminated ASCII
            if dt.itemsize == 0:
                arr = arr.astype("S1")
        else:
            raise HDF5Error(f"dtype {dt} is not written")
        raw = arr.tobytes()
        addr = self.alloc(raw) if raw else UNDEF
        layout = struct.pack("<BBQQ", 3, 1, addr, len(raw))
        return self.header([(MSG_DATASPACE, space), (MSG_DATATYPE, dtype), (MSG_FILL, _FILL_V2),
                            (MSG_LAYOUT, layout)])

    # -- a group: children first, then its heap, symbol-table nodes, B-tree and header
    def group(self, tree: Dict[str, Any]) -> Tuple[int, int, int]:
        entries = []
        for name in sorted(tree):  # symbol-table nodes hold names in ascending order
            v = tree[name]
            if isinstance(v, dict):
                hdr, bt, hp = self.group(v)
                entries.append((name, hdr, (bt, hp)))
            else:
                entries.append((name, self.dataset(v), None))
        heap = bytearray(b"\0" * 8)  # offset 0: the empty name
        offs = []
        for name, _, _ in entries:
            offs.append(len(heap))
            heap += _pad8(name.encode("utf-8") + b"\0")
        heap_data = self.alloc(bytes(heap))
        heap_addr = self.alloc(b"HEAP" + bytes(4) + struct.pack("<QQQ", len(heap), UNDEF, heap_data))
        nodes, keys = [], [0]
        for i in range(0, max(len(entries), 1), 2 * LEAF_K):
            chunk = list(zip(entries, offs))[i:i + 2 * LEAF_K]
            body = b"SNOD" + struct.pack("<BBH", 1, 0, len(chunk))
            for (name, hdr, scratch), off in chunk:
                if scratch is None:
                    body += struct.pack("<QQI4x16x", off, hdr, 0)
                else:
                    body += struct.pack("<QQI4xQQ", off, hdr, 1, scratch[0], scratch[1])
            body += b"\0" * (40 * (2 * LEAF_K - len(chunk)))
            nodes.append(self.alloc(body))
            keys.append(chunk[-1][1] if chunk else 0)  # the largest name in the node
        if len(nodes) > 2 * INTERNAL_K:
            raise HDF5Error(f"group with more than {2 * INTERNAL_K * 2 * LEAF_K} members")
        bt = b"TREE" + struct.pack("<BBHQQ", 0, 0, len(nodes), UNDEF, UNDEF)
        for k, n in zip(keys, nodes):
            bt += struct.pack("<QQ", k, n)
        bt += struct.pack("<Q", keys[len(nodes)])
        bt += b"\0" * (8 * (2 * INTERNAL_K + 1) + 8 * 2 * INTERNAL_K - (len(bt) - 24))
        bt_addr = self.alloc(bt)
        hdr = self.header([(MSG_SYMBOL_TABLE, struct.pack("<QQ", bt_addr, heap_addr))])
        return hdr, bt_addr, heap_addr

    def finish(self, root: Tuple[int, int, int]) -> bytes:
        hdr, bt, hp = root
        sb = SIGNATURE + struct.pack("<BBBBBBBB", 0, 0, 0, 0, 0, 8, 8, 0)
        sb += struct.pack("<HHI", LEAF_K, INTERNAL_K, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, len(self.buf), UNDEF)
        sb += struct.pack("<QQI4xQQ", 0, hdr, 1, bt, hp)
        assert len(sb) == 96
        self.buf[:96] = sb
        return bytes(self.buf)


def write(path: str, tree: Dict[str, Any]) -> None:
    """Writes a nested dict (groups) of numpy arrays / scalars / strings (datasets).  Booleans are
    stored as h5py stores them (an int8 enumeration FALSE/TRUE); strings as fixed-length
    null-terminated strings, which h5py returns as bytes (it writes Python str as
    variable-length strings, which this module reads but does not write)."""
    w = _Writer()
    data = w.finish(w.group(tree))
    with open(path, "wb") as f:
        f.write(data)


# ===================================================================== the simulation's output layout
def _cells_faces_sizes(domain_information, axis: int):
    """homogenous.py:13-21: cell centers linspace(a + dx/2, b - dx/2, n), faces linspace(a, b, n+1),
    the cell size; squeezed as hdf5_writer.py:125-131 squeezes them (a 1-cell axis gives scalars)."""
    (a, b), n = domain_information.domain_size[axis], domain_information.global_number_of_cells[axis]
    dx = (b - a) / n
    centers = np.squeeze(np.linspace(a + dx / 2, b - dx / 2, n))
    faces = np.squeeze(np.linspace(a, b, n + 1))
    return centers, faces, np.float64(dx)


class HDF5Writer:
    """Output files of io_utils/hdf5_writer.py:31-266 for a single-phase, single-device run:
    ``metadata/``, ``domain/``, ``wall_clock/``, ``time`` and the requested ``primitives/`` and
    ``conservatives/`` quantities (density, velocity, pressure, temperature; mass, momentum,
    energy), each the interior cells transposed to (Nz, Ny, Nx) and, for vectors, the active
    components last (``hdf5_writer.py:330-333``).  Non-dimensional units (the unit handler's
    reference values of 1, as on this path)."""

    PRIMITIVE_INDEX = {"density": 0, "pressure": 4}
    CONSERVATIVE_INDEX = {"mass": 0, "energy": 4}

    def __init__(self, domain_information, gas_constant: float = 1.0, is_double: bool = True,
                 quantities: Optional[Dict[str, List[str]]] = None, is_metadata: bool = True,
                 is_domain: bool = True, is_wall_clock_times: bool = True, is_time: bool = True) -> None:
        self.domain = domain_information
        self.R = float(gas_constant)
        self.is_double = bool(is_double)
        self.quantities = quantities or {"primitives": ["density", "velocity", "pressure"]}
        self.is_metadata, self.is_domain = is_metadata, is_domain
        self.is_wall_clock_times, self.is_time = is_wall_clock_times, is_time
        self.num_digits_output = 10

    def _field(self, key: str, quantity: str, W: np.ndarray, U: np.ndarray) -> np.ndarray:
        d = self.domain
        nh = d.nh_conservatives
        sl = tuple(slice(nh, -nh) if n > 1 else slice(None) for n in d.global_number_of_cells)
        act = list(d.active_axes_indices)
        if key == "primitives":
            if quantity == "velocity":
                buf = W[[1 + a for a in act]]
            elif quantity == "temperature":
                buf = W[4] / (W[0] * self.R)              # ideal_gas.py:81-83
            else:
                buf = W[self.PRIMITIVE_INDEX[quantity]]
        else:
            buf = U[[1 + a for a in act]] if quantity == "momentum" else U[self.CONSERVATIVE_INDEX[quantity]]
        buf = buf[(Ellipsis,) + sl]
        return np.ascontiguousarray(buf.T).astype("f8" if self.is_double else "f4")

    def tree(self, primitives, conservatives, time: float, wall_clock=None) -> Dict[str, Any]:
        """The file's contents as a nested dict (what write_file writes)."""
        W = np.asarray(primitives.detach().cpu() if hasattr(primitives, "detach") else primitives, dtype=np.float64)
        U = np.asarray(conservatives.detach().cpu() if hasattr(conservatives, "detach") else conservatives,
                       dtype=np.float64)
        d = self.domain
        ft = "f8" if self.is_double else "f4"
        t: Dict[str, Any] = {}
        if self.is_metadata:
            t["metadata"] = {
                "is_parallel": False, "is_multihost": False, "process_id": np.int64(0),
                "local_device_count": np.int64(1), "global_device_count": np.int64(1),
                "fluid_names": np.array(["fluid_0"]), "number_fluids": np.int64(1),
                "levelset_model": False, "diffuse_interface_model": False, "cavitation_model": False,
                "is_double_precision": self.is_double,
                "available_quantities": {k: np.array(v) for k, v in self.quantities.items() if v},
            }
        if self.is_domain:
            dom: Dict[str, Any] = {"dim": np.int64(d.dim),
                                   "active_axes_indices": np.array(d.active_axes_indices, dtype=np.int64)}
            for i, X in enumerate("XYZ"):
                c, f, s = _cells_faces_sizes(d, i)
                dom[f"grid{X}"], dom[f"gridF{X}"], dom[f"cellsize{X}"] = (np.asarray(c, ft), np.asarray(f, ft),
                                                                         np.asarray(s, ft))
            dom["split_factors"] = np.array(d.split_factors, dtype=np.int64)
            t["domain"] = dom
        if self.is_wall_clock_times:
            wc = wall_clock or (0.0, 0.0, 0.0, 0.0)
            t["wall_clock"] = {k: np.float64(v) for k, v in
                               zip(("step", "step_per_cell", "mean_step", "mean_step_per_cell"), wc)}
        if self.is_time:
            t["time"] = np.asarray(time, dtype=ft)
        for key in ("conservatives", "primitives"):
            if self.quantities.get(key):
                t[key] = {q: self._field(key, q, W, U) for q in self.quantities[key]}
        return t

    def filename(self, time: float, step: Optional[int] = None) -> str:
        """data_<time with 10 decimals>.h5, or data_<step>.h5 (hdf5_writer.py:100-112)."""
        return f"data_{step:d}.h5" if step is not None else f"data_{time:.{self.num_digits_output}f}.h5"

    def write_file(self, directory: str, primitives, conservatives, time: float, wall_clock=None,
                   step: Optional[int] = None) -> str:
        import os

        path = os.path.join(directory, self.filename(time, step))
        write(path, self.tree(primitives, conservatives, time, wall_clock))
        return path


def read_restart(path: str, domain_information) -> Tuple[float, np.ndarray]:
    """MaterialFieldsInitializer.from_restart_file for a single-phase, single-device file
    (initialization/material_fields_initializer.py:212-333, load_buffer of
    initialization/helper_functions.py:197-226): checks dim, cell counts and the absence of
    level-set / diffuse-interface models, loads primitives/density, velocity, pressure
    (velocity: its components moved first, then (z, y, x) swapped back), and returns
    (time, primitives (5, Nx, Ny, Nz)) with zero velocity on inactive axes."""
    f = read(path)
    d = domain_information
    prim_q = [q.decode() if isinstance(q, bytes) else str(q) for q in np.atleast_1d(f["metadata"]["available_quantities"]["primitives"])]
    for q in ("density", "velocity", "pressure"):
        assert q in prim_q, f"{q} not in restart file {path}."
    assert int(f["domain"]["dim"]) == d.dim, f"Dimension of restart file {path} does not match case setup file."
    act = [int(a) for a in np.atleast_1d(f["domain"]["active_axes_indices"])]
    split = [int(s) for s in np.atleast_1d(f["domain"]["split_factors"])]
    cells = []
    for i, X in enumerate("XYZ"):
        g = np.atleast_1d(f["domain"][f"grid{X}"])
        cells.append(g.shape[-1] * split[i] if i in act else 1)
    assert tuple(cells) == tuple(d.global_number_of_cells), \
        f"Number of cells of restart file {path} do not match case setup file."
    for key in ("levelset_model", "diffuse_interface_model"):
        v = f["metadata"][key]
        v = v.decode("utf-8") if isinstance(v, (bytes, np.bytes_)) else v
        assert v in (False, None, "None", "False"), f"{key} of restart file {path} is not single phase"

    def load(q, vector=False):
        b = np.asarray(f["primitives"][q], dtype=np.float64)
        if vector:
            b = np.moveaxis(b, -1, -4)
        return np.swapaxes(b, -1, -3)

    rho, p, vel = load("density"), load("pressure"), load("velocity", True)
    W = np.zeros((5,) + tuple(d.global_number_of_cells))
    W[0], W[4] = rho, p
    for k, a in enumerate(d.active_axes_indices):
        W[1 + a] = vel[k]
    return float(f["time"]), W


class XDMFWriter:
    """XDMF companions of the HDF5 files (io_utils/xdmf_writer.py:37-211): a temporal collection
    with one uniform grid per output time -- the cell faces gridFX/FY/FZ as a 3DRectMesh (VXVYVZ
    geometry) and one cell-centred attribute per written quantity (vectors with the active
    components last) -- per file, and data_time_series.xdmf over all files written."""

    VECTORS = ("velocity", "momentum")

    def __init__(self, domain_information, quantities: Optional[Dict[str, List[str]]] = None,
                 is_double: bool = True) -> None:
        self.domain = domain_information
        self.quantities = quantities or {"primitives": ["density", "velocity", "pressure"]}
        self.precision = 8 if is_double else 4
        self.series: List[str] = []
        self.times: List[float] = []

    def _grid(self, h5name: str, time: float) -> str:
        nx, ny, nz = self.domain.global_number_of_cells
        fx, fy, fz = nx + 1, ny + 1, nz + 1
        p = self.precision
        item = lambda dims, ref: (f'<DataItem Format="HDF" NumberType="Float" Precision="{p}" '  # noqa: E731
                                  f'Dimensions="{dims}">{h5name}:{ref}</DataItem>')
        parts = [f'<Grid Name="SpatialData_{time:e}" GridType="Uniform">',
                 f'<Time TimeType="Single" Value="{time:e}" />',
                 '<Geometry Type="VXVYVZ">',
                 item(fx, "domain/gridFX"), item(fy, "domain/gridFY"), item(fz, "domain/gridFZ"),
                 '</Geometry>',
                 f'<Topology Dimensions="{fx} {fy} {fz}" Type="3DRectMesh"/>']
        for key in ("conservatives", "primitives"):
            for q in self.quantities.get(key) or []:
                if q in self.VECTORS:
                    parts += [f'<Attribute Name="{q}" AttributeType="Vector" Center="Cell">',
                              item(f"{nz} {ny} {nx} {self.domain.dim}", f"{key}/{q}"), '</Attribute>']
                else:
                    parts += [f'<Attribute Name="{q}" AttributeType="Scalar" Center="Cell">',
                              item(f"{nz} {ny} {nx}", f"{key}/{q}"), '</Attribute>']
        parts.append('</Grid>')
        return "\n".join(parts)

    @staticmethod
    def _document(name: str, grids: List[str]) -> str:
        return "\n".join(['<?xml version="1.0" ?>', '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>', '<Xdmf Version="3.0">',
                          '<Domain>', f'<Grid Name="{name}" GridType="Collection" CollectionType="Temporal">',
                          *grids, '</Grid>', '</Domain>', '</Xdmf>'])

    def write_file(self, h5path: str, time: float) -> str:
        import os

        stem = os.path.splitext(h5path)[0]
        grid = self._grid(os.path.basename(h5path), time)
        if time not in self.times:
            self.times.append(time)
            self.series.append(grid)
        with open(stem + ".xdmf", "w") as f:
            f.write(self._document("TimeStep", [grid]))
        return stem + ".xdmf"

    def write_timeseries(self, directory: str) -> str:
        import os

        path = os.path.join(directory, "data_time_series.xdmf")
        with open(path, "w") as f:
            f.write(self._document("TimeSeries", self.series))
        return path
