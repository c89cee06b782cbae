"""Convert the POM acceptance asset of the reference (assets/pmt_acc.hd5, read
by POMAcceptance, src/pom.jl:25-35) into the package's data file
photonpropagation.jl_amd/photonprop/data/pom_acceptance.csv.

h5py is not available here, so this reads the file's few datasets directly:
HDF5 superblock version 0, version-1 object headers (with continuation
blocks), link messages of the root group, and contiguous or compact float64
datasets - the subset of the format this small file uses.

usage: python convert_pom_acceptance.py /root/reference/assets/pmt_acc.hd5
"""
import os
import struct
import sys


def _messages(d, addr):
    """(type, body bytes) of the messages of the v1 object header at addr."""
    assert d[addr] == 1, "object header version"
    size = struct.unpack_from("<I", d, addr + 8)[0]
    todo = [(addr + 16, addr + 16 + size)]
    while todo:
        p, end = todo.pop(0)
        while p + 8 <= end:
            mtype, msize = struct.unpack_from("<HH", d, p)
            body = d[p + 8:p + 8 + msize]
            if mtype == 0x10:  # continuation
                a, n = struct.unpack_from("<QQ", body)
                todo.append((a, a + n))
            else:
                yield mtype, body
            p += 8 + msize


def _links(d, addr):
    """name -> object header address of a group's link messages."""
    out = {}
    for mtype, body in _messages(d, addr):
        if mtype != 0x06:
            continue
        ver, flags = body[0], body[1]
        assert ver == 1 and not flags & 0x08, "only hard links"
        p = 2
        if flags & 0x04:
            p += 8  # creation order
        if flags & 0x10:
            p += 1  # name character set
        nlen_size = 1 << (flags & 0x03)
        nlen = int.from_bytes(body[p:p + nlen_size], "little")
        p += nlen_size
        name = body[p:p + nlen].decode()
        out[name] = struct.unpack_from("<Q", body, p + nlen)[0]
    return out


def _dataset(d, addr):
    shape, raw, dtype = None, None, None
    for mtype, body in _messages(d, addr):
        if mtype == 0x01:  # dataspace
            ver, rank = body[0], body[1]
            o = 8 if ver == 1 else 4
            shape = tuple(struct.unpack_from("<Q", body, o + 8 * i)[0] for i in range(rank))
        elif mtype == 0x03:  # datatype
            dtype = (body[0] & 0x0F, struct.unpack_from("<I", body, 4)[0])
        elif mtype == 0x08:  # layout
            assert body[0] == 3, "layout version"
            if body[1] == 0:
                n = struct.unpack_from("<H", body, 2)[0]
                raw = body[4:4 + n]
            elif body[1] == 1:
                a, n = struct.unpack_from("<QQ", body, 2)
                raw = d[a:a + n]
            else:
                raise ValueError("chunked datasets are not supported")
    assert dtype == (1, 8), f"expected float64, got {dtype}"
    vals = list(struct.unpack("<%dd" % (len(raw) // 8), raw))
    return shape, vals


def read(path):
    d = open(path, "rb").read()
    assert d[:8] == b"\x89HDF\r\n\x1a\n" and d[8] == 0, "superblock version 0 expected"
    root = struct.unpack_from("<Q", d, 24 + 32 + 8)[0]  # root symbol-table entry
    return {name: _dataset(d, a) for name, a in _links(d, root).items()}


def main(path):
    ds = read(path)
    wl = ds["wavelengths"][1]
    a1, a2 = ds["acc_pmt_grp_1"][1], ds["acc_pmt_grp_2"][1]
    s1, s2 = ds["sigma_grp_1"][1][0], ds["sigma_grp_2"][1][0]
    repo = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    out = os.path.join(repo, "photonpropagation.jl_amd", "photonprop", "data",
                       "pom_acceptance.csv")
    with open(out, "w") as f:
        f.write(f"# POM acceptance from pmt_acc.hd5; sigma_grp_1={s1!r}; sigma_grp_2={s2!r}\n")
        f.write("wavelength,acc_pmt_grp_1,acc_pmt_grp_2\n")
        for w, x, y in zip(wl, a1, a2):
            f.write(f"{w!r},{x!r},{y!r}\n")
    print(out, len(wl), "rows; sigmas", s1, s2)


if __name__ == "__main__":
    main(sys.argv[1])
