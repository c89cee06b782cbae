"""Photon-table files in HDF5 without an HDF5 library.

The reference writes its training tables with HDF5.jl: `save_hdf!(fname,
group, dataset, attributes)` (scripts/photon_tables/utils.jl:24-55) appends a
matrix as `group/dataset_<n>` with the simulation's attributes and keeps the
count in the group attribute `nsims`. Neither h5py nor libhdf5 exists here, so
this module writes and reads the subset of the format such files need, laid
out like the files HDF5.jl itself writes (the reference's
assets/pmt_acc.hd5 is one):

* superblock version 0, 8-byte offsets and lengths;
* version-1 object headers; groups stored compactly (Link Info, Group Info
  and one Link message per member, hard links only);
* datasets: version-1 dataspace, IEEE little-endian float64 or int64, fill
  value message version 2, contiguous layout (version 3);
* attributes (version-1 messages) on groups and datasets: scalar float64,
  int64, or fixed-length null-terminated ASCII/UTF-8 strings.

A Julia matrix is column-major, so HDF5.jl stores an N x 10 photon table with
dimensions (10, N); `save_hdf` does the same for an (N, 10) array, and
`read_hdf` returns the array as stored, (10, N).
"""
from __future__ import annotations

import os
import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
_SIG = b"\x89HDF\r\n\x1a\n"


def _pad8(b: bytes) -> bytes:
    return b + b"\0" * (-len(b) % 8)


# ---------------------------------------------------------------- datatypes
def _dtype_float64() -> bytes:
    # class 1 (floating point), version 1; little-endian, implied mantissa
    # normalisation, sign at bit 63; precision 64, exponent 52/11, mantissa
    # 0/52, bias 1023
    return bytes([0x11, 0x20, 0x3F, 0x00]) + struct.pack("<I", 8) + struct.pack(
        "<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)


def _dtype_int64() -> bytes:
    # class 0 (fixed point), version 1; little-endian, signed
    return bytes([0x10, 0x08, 0x00, 0x00]) + struct.pack("<I", 8) + struct.pack("<HH", 0, 64)


def _dtype_string(n: int, utf8: bool) -> bytes:
    # class 3 (string), version 1; null-terminated, ASCII or UTF-8
    return bytes([0x13, 0x10 if utf8 else 0x00, 0x00, 0x00]) + struct.pack("<I", n)


def _dataspace(shape) -> bytes:
    if len(shape) == 0:
        return bytes([1, 0, 0, 0]) + b"\0" * 4
    body = bytes([1, len(shape), 1, 0]) + b"\0" * 4
    body += b"".join(struct.pack("<Q", int(d)) for d in shape)
    return body + b"".join(struct.pack("<Q", int(d)) for d in shape)


def _attr_value(v):
    """(datatype bytes, data bytes) of a scalar attribute value."""
    if isinstance(v, (bool, np.bool_)):
        v = int(v)
    if isinstance(v, (int, np.integer)):
        return _dtype_int64(), struct.pack("<q", int(v))
    if isinstance(v, (float, np.floating)):
        return _dtype_float64(), struct.pack("<d", float(v))
    if isinstance(v, str):
        raw = v.encode("utf-8")
        return _dtype_string(len(raw) + 1, not raw.isascii()), raw + b"\0"
    raise TypeError(f"attribute values are int, float or str, not {type(v).__name__}")


# --------------------------------------------------------------- messages
def _msg(mtype: int, body: bytes, flags: int = 0) -> bytes:
    body = _pad8(body)
    return struct.pack("<HHB3x", mtype, len(body), flags) + body


def _attribute(name: str, value) -> bytes:
    dt, data = _attr_value(value)
    nm = name.encode("utf-8") + b"\0"
    ds = _dataspace(())
    body = struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(ds))
    body += _pad8(nm) + _pad8(dt) + _pad8(ds) + data
    return _msg(0x0C, body)


def _link(name: str, addr: int) -> bytes:
    nm = name.encode("utf-8")
    if len(nm) > 255:
        raise ValueError("link names are at most 255 bytes here")
    # version 1, flags 0x10: character set present (1 = UTF-8), 1-byte length
    return _msg(0x06, bytes([1, 0x10, 1, len(nm)]) + nm + struct.pack("<Q", addr))


def _object_header(messages: list[bytes]) -> bytes:
    body = b"".join(messages)
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body)) + body


# ------------------------------------------------------------------ writer
class _Writer:
    def __init__(self):
        self.buf = bytearray(96)  # superblock + root symbol table entry

    def alloc(self, data: bytes) -> int:
        addr = len(self.buf)
        self.buf += data
        self.buf += b"\0" * (-len(self.buf) % 8)
        return addr

    def dataset(self, arr: np.ndarray, attrs: dict) -> int:
        arr = np.asarray(arr)
        # np.array(order="C") keeps a 0-d array 0-d (a scalar dataset)
        if arr.dtype.kind == "f":
            a, dt = np.array(arr, "<f8", order="C"), _dtype_float64()
        elif arr.dtype.kind in "iub":
            a, dt = np.array(arr, "<i8", order="C"), _dtype_int64()
        else:
            raise TypeError("datasets are float64 or int64 arrays")
        data_addr = self.alloc(a.tobytes()) if a.size else UNDEF
        msgs = [_msg(0x01, _dataspace(a.shape)), _msg(0x03, dt, flags=1),
                _msg(0x05, bytes([2, 2, 2, 1]) + struct.pack("<I", 0), flags=1),
                _msg(0x08, bytes([3, 1]) + struct.pack("<QQ", data_addr, a.nbytes))]
        msgs += [_attribute(k, v) for k, v in attrs.items()]
        return self.alloc(_object_header(msgs))

    def group(self, members: dict, attrs: dict) -> int:
        """members: name -> (kind, payload): ("group", (members, attrs)) or
        ("dataset", (array, attrs))."""
        links = []
        for name, (kind, payload) in members.items():
            addr = self.group(*payload) if kind == "group" else self.dataset(*payload)
            links.append(_link(name, addr))
        msgs = [_msg(0x02, bytes([0, 0]) + struct.pack("<QQ", UNDEF, UNDEF)),
                _msg(0x0A, bytes([0, 0]))] + links
        msgs += [_attribute(k, v) for k, v in attrs.items()]
        return self.alloc(_object_header(msgs))

    def finish(self, root_addr: int) -> bytes:
        sb = _SIG + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack("<HHI", 4, 16, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, len(self.buf), UNDEF)
        sb += struct.pack("<QQII16x", 0, root_addr, 0, 0)  # root entry, no cache
        assert len(sb) == 96
        self.buf[:96] = sb
        return bytes(self.buf)


def write_hdf(path: str, tree: dict) -> None:
    """Write `tree` (as returned by read_hdf: {"attrs": {...}, "members":
    {name: subtree}} for groups, {"data": array, "attrs": {...}} for
    datasets) as a new file."""
    def lower(node):
        if "data" in node:
            return "dataset", (node["data"], node.get("attrs", {}))
        return "group", ({k: lower(v) for k, v in node.get("members", {}).items()},
                         node.get("attrs", {}))
    w = _Writer()
    _, (members, attrs) = lower(tree)
    root = w.group(members, attrs)
    data = w.finish(root)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(data)
    os.replace(tmp, path)


# ------------------------------------------------------------------ reader
def _messages(d: bytes, addr: int):
    if d[addr] != 1:
        raise ValueError("only version-1 object headers are read here")
    size = struct.unpack_from("<I", d, addr + 8)[0]
    todo = [(addr + 16, addr + 16 + size)]
    while todo:
        p, end = todo.pop(0)
        while p + 8 <= end:
            mtype, msize = struct.unpack_from("<HH", d, p)
            body = d[p + 8:p + 8 + msize]
            if mtype == 0x10:  # continuation block
                a, n = struct.unpack_from("<QQ", body)
                todo.append((a, a + n))
            else:
                yield mtype, body
            p += 8 + msize


def _parse_dtype(b: bytes):
    cls, ver = b[0] & 0x0F, b[0] >> 4
    size = struct.unpack_from("<I", b, 4)[0]
    if cls == 1 and size in (4, 8) and not b[1] & 1:
        return np.dtype("<f8" if size == 8 else "<f4")
    if cls == 0 and not b[1] & 1:
        return np.dtype(("<i" if b[1] & 0x08 else "<u") + str(size))
    if cls == 3:
        return ("str", size)
    raise ValueError(f"unsupported datatype class {cls} version {ver}")


def _parse_dataspace(b: bytes):
    ver, rank, flags = b[0], b[1], b[2]
    off = 8 if ver == 1 else 4
    return tuple(struct.unpack_from("<Q", b, off + 8 * i)[0] for i in range(rank))


def _parse_attr(b: bytes):
    ver = b[0]
    if ver != 1:
        raise ValueError("only version-1 attribute messages are read here")
    nlen, dlen, slen = struct.unpack_from("<HHH", b, 2)
    p = 8
    name = b[p:p + nlen].split(b"\0")[0].decode("utf-8")
    p += nlen + (-nlen % 8)
    dt = _parse_dtype(b[p:p + dlen])
    p += dlen + (-dlen % 8)
    shape = _parse_dataspace(b[p:p + slen])
    p += slen + (-slen % 8)
    if shape:
        raise ValueError("only scalar attributes are read here")
    if isinstance(dt, tuple):
        return name, b[p:p + dt[1]].split(b"\0")[0].decode("utf-8")
    v = np.frombuffer(b, dt, 1, p)[0]
    return name, (float(v) if dt.kind == "f" else int(v))


def _node(d: bytes, addr: int) -> dict:
    links, attrs = {}, {}
    shape = dtype = None
    layout = None
    for mtype, body in _messages(d, addr):
        if mtype == 0x06:
            ver, flags = body[0], body[1]
            if ver != 1 or flags & 0x08:
                raise ValueError("only hard links are read here")
            p = 2
            if flags & 0x04:
                p += 8
            if flags & 0x10:
                p += 1
            n = 1 << (flags & 0x03)
            nlen = int.from_bytes(body[p:p + n], "little")
            p += n
            links[body[p:p + nlen].decode("utf-8")] = struct.unpack_from("<Q", body, p + nlen)[0]
        elif mtype == 0x0C:
            k, v = _parse_attr(body)
            attrs[k] = v
        elif mtype == 0x01:
            shape = _parse_dataspace(body)
        elif mtype == 0x03:
            dtype = _parse_dtype(body)
        elif mtype == 0x08:
            layout = body
    if shape is None:
        return {"attrs": attrs, "members": {k: _node(d, a) for k, a in links.items()}}
    n = int(np.prod(shape)) if shape else 1
    if layout[0] != 3:
        raise ValueError("only version-3 layouts are read here")
    if layout[1] == 1:  # contiguous
        a, size = struct.unpack_from("<QQ", layout, 2)
        raw = d[a:a + size] if a != UNDEF else b""
    elif layout[1] == 0:  # compact
        size = struct.unpack_from("<H", layout, 2)[0]
        raw = layout[4:4 + size]
    else:
        raise ValueError("chunked datasets are not read here")
    arr = np.frombuffer(raw, dtype, n).reshape(shape).copy()
    return {"data": arr, "attrs": attrs}


def read_hdf(path: str) -> dict:
    """The file as a tree: groups {"attrs", "members"}, datasets {"data",
    "attrs"}."""
    with open(path, "rb") as f:
        d = f.read()
    if d[:8] != _SIG or d[8] != 0:
        raise ValueError(f"{path}: not an HDF5 file with a version-0 superblock")
    root = struct.unpack_from("<Q", d, 64)[0]
    return _node(d, root)


# ------------------------------------------------------------- save_hdf!
def save_hdf(fname: str, group: str, dataset, attributes: dict) -> str:
    """save_hdf!(fname, group, dataset, attributes) (scripts/photon_tables/
    utils.jl:24-55): append `dataset` (a 2-D array, stored transposed as
    HDF5.jl stores a Julia matrix) as `group/dataset_<nsims+1>` with
    `attributes`, creating the file and the group (nsims = 0) as needed, and
    set the group attribute nsims. Returns the dataset name. The file is
    rewritten whole (this writer keeps no free-space map)."""
    tree = read_hdf(fname) if os.path.isfile(fname) else {"attrs": {}, "members": {}}
    g = tree["members"].setdefault(group, {"attrs": {"nsims": 0}, "members": {}})
    offset = int(g["attrs"].get("nsims", 0)) + 1
    name = f"dataset_{offset}"
    data = np.asarray(dataset)
    g["members"][name] = {"data": np.ascontiguousarray(data.T), "attrs": dict(attributes)}
    g["attrs"]["nsims"] = offset
    write_hdf(fname, tree)
    return name
