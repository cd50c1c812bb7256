"""Voxels in .ply files (reference o_voxel/io/ply.py:27-71).

The reference goes through the plyfile package; this module reads and writes
the same files with its own codec of the PLY format (numpy structured arrays),
so it needs no extra package.  The file holds one `vertex` element: float32
properties x, y, z (the voxel coordinates) and, per attribute channel, a
property `<name>_<j>` of the attribute's type.  Writing produces what plyfile
writes for that element by default (binary in the machine's byte order, i.e.
`binary_little_endian 1.0` here); reading accepts the binary little / big
endian and ascii encodings.
"""

import io
import os
from typing import Dict

import numpy as np
import torch

__all__ = ["read_ply", "write_ply"]

# attribute dtypes accepted by write_ply (reference ply.py:14-23)
DTYPE_MAP = {
    torch.uint8: "u1",
    torch.uint16: "u2",
    torch.uint32: "u4",
    torch.int8: "i1",
    torch.int16: "i2",
    torch.int32: "i4",
    torch.float32: "f4",
    torch.float64: "f8",
}

# PLY scalar type names: the classic names (written) and the sized aliases
_PLY_TO_NP = {
    "char": "i1", "int8": "i1",
    "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2",
    "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4",
    "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4",
    "double": "f8", "float64": "f8",
}
_NP_TO_PLY = {"i1": "char", "u1": "uchar", "i2": "short", "u2": "ushort", "i4": "int", "u4": "uint",
              "f4": "float", "f8": "double"}


def _open(file, mode):
    if isinstance(file, (str, bytes, os.PathLike)):
        return open(file, mode), True
    return file, False


def _parse_header(fh):
    """(format, [(element, count, [(name, np type)])]) of a PLY header; the
    stream is left at the first byte of the body."""
    magic = fh.readline()
    if magic.strip() != b"ply":
        raise ValueError("not a PLY file (missing 'ply' magic line)")
    fmt, elements = None, []
    while True:
        line = fh.readline()
        if not line:
            raise ValueError("PLY header is not terminated by end_header")
        tok = line.decode("ascii").split()
        if not tok or tok[0] in ("comment", "obj_info"):
            continue
        if tok[0] == "end_header":
            break
        if tok[0] == "format":
            if len(tok) != 3 or tok[1] not in ("ascii", "binary_little_endian", "binary_big_endian"):
                raise ValueError(f"unsupported PLY format line: {line!r}")
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append((tok[1], int(tok[2]), []))
        elif tok[0] == "property":
            if not elements:
                raise ValueError("PLY property before any element")
            if tok[1] == "list":
                elements[-1][2].append((tok[4], None))  # list property: skipped when reading
            else:
                if tok[1] not in _PLY_TO_NP:
                    raise ValueError(f"unknown PLY property type {tok[1]!r}")
                elements[-1][2].append((tok[2], _PLY_TO_NP[tok[1]]))
        else:
            raise ValueError(f"unexpected PLY header line: {line!r}")
    if fmt is None:
        raise ValueError("PLY header has no format line")
    return fmt, elements


def _read_vertex(fh, fmt, elements):
    """Structured array of the first element (the reference reads elements[0])."""
    if not elements:
        raise ValueError("PLY file has no element")
    name, count, props = elements[0]
    if any(t is None for _, t in props):
        raise ValueError("list properties in the voxel element are not supported")
    if fmt == "ascii":
        dt = np.dtype([(p, t) for p, t in props])
        rows = [fh.readline().split() for _ in range(count)]
        out = np.empty(count, dtype=dt)
        for j, (p, t) in enumerate(props):
            out[p] = np.array([r[j] for r in rows], dtype=np.float64 if t[0] == "f" else np.int64).astype(t)
        return out
    bo = "<" if fmt == "binary_little_endian" else ">"
    dt = np.dtype([(p, bo + t) for p, t in props])
    buf = fh.read(dt.itemsize * count)
    if len(buf) != dt.itemsize * count:
        raise ValueError("PLY body is shorter than its header says")
    return np.frombuffer(buf, dtype=dt, count=count).astype(dt.newbyteorder("="))


def read_ply(file):
    """(coord int64 [N, 3], {name: tensor [N, C]}) of a PLY voxel file.

    Coordinates are the rounded x, y, z; every other property `<name>_<j>`
    becomes channel j of attribute `<name>` (reference ply.py:27-53)."""
    fh, own = _open(file, "rb")
    try:
        fmt, elements = _parse_header(fh)
        data = _read_vertex(fh, fmt, elements)
    finally:
        if own:
            fh.close()
    xyz = np.stack([data[k] for k in ("x", "y", "z")], axis=1)
    coord = torch.from_numpy(np.round(xyz).astype(np.int64))
    keys = [k for k in data.dtype.names if k not in ("x", "y", "z")]
    names = []
    for k in keys:
        n = "_".join(k.split("_")[:-1])
        if n not in names:
            names.append(n)
    attr = {}
    for n in names:
        chs = sum(1 for k in keys if k.startswith(f"{n}_"))
        attr[n] = torch.from_numpy(np.ascontiguousarray(np.stack([data[f"{n}_{j}"] for j in range(chs)], axis=1)))
    return coord, attr


def write_ply(file, coord: torch.Tensor, attr: Dict[str, torch.Tensor]):
    """One `vertex` element: x, y, z as float32 and one property per attribute
    channel, binary in the machine's byte order (reference ply.py:56-71)."""
    fields = [("x", "f4"), ("y", "f4"), ("z", "f4")]
    for k, v in attr.items():
        assert v.dtype in DTYPE_MAP, f"Unsupported data type {v.dtype} for attribute {k}"
        for j in range(v.shape[-1]):
            fields.append((f"{k}_{j}", DTYPE_MAP[v.dtype]))
    n = len(coord)
    data = np.empty(n, dtype=np.dtype([(p, "=" + t) for p, t in fields]))
    c = coord.detach().cpu().numpy().astype(np.float32).reshape(n, 3)
    data["x"], data["y"], data["z"] = c[:, 0], c[:, 1], c[:, 2]
    for k, v in attr.items():
        a = v.detach().cpu().numpy().reshape(n, -1)
        for j in range(a.shape[1]):
            data[f"{k}_{j}"] = a[:, j]
    order = "binary_little_endian" if np.little_endian else "binary_big_endian"
    head = ["ply", f"format {order} 1.0", f"element vertex {n}"]
    head += [f"property {_NP_TO_PLY[t]} {p}" for p, t in fields]
    head.append("end_header")
    fh, own = _open(file, "wb")
    try:
        fh.write(("\n".join(head) + "\n").encode("ascii"))
        fh.write(data.tobytes())
    finally:
        if own:
            fh.close()
