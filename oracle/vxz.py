"""TEST INFRASTRUCTURE -- Python restatement of the .vxz voxel file format.

Only ``tests/`` and ``bench.py``'s baseline leg use this module, as the
checker.  The product package never imports it.

Restated from
  o_voxel/io/vxz.py:120-365          (read_vxz / write_vxz: chunking, header,
                                      per-stream compression, interleaving)
  src/io/svo.cpp:17-138              (sparse voxel octree encode / decode)
  src/io/filter_parent.cpp:10-165    (attribute deltas between octree siblings)
  src/io/filter_neighbor.cpp:8-178   (attribute deltas to a lower neighbour)
sequentially, in plain Python / numpy, for small inputs.
"""

from __future__ import annotations

import json
import lzma
import struct
import zlib
from typing import Dict, List, Tuple

import numpy as np

from . import serialize as S

DEFAULT_LEVEL = {"none": 0, "deflate": 9, "lzma": 9, "zstd": 22}


# ---------------------------------------------------------------------------
# sparse voxel octree over ascending Morton codes of `depth` 3-bit digits
# ---------------------------------------------------------------------------
def svo_encode(codes, depth: int) -> np.ndarray:
    """Pre-order octree bytes (child masks) of ascending codes (svo.cpp:17-73):
    the root, then for each code the nodes it opens, shallow to deep."""
    svo = [0]
    ptr = [0] * depth
    prev = None
    for code in (int(c) & 0xFFFFFFFF for c in codes):
        digits = [(code >> (3 * (depth - 1 - j))) & 7 for j in range(depth)]
        if prev is None:
            start = 0
        else:
            start = depth - 1
            for j in range(depth - 1):
                if digits[j] != prev[j]:
                    start = j
                    break
        for j in range(start, depth):
            if j > start:
                svo.append(0)
                ptr[j] = len(svo) - 1
            svo[ptr[j]] |= 1 << digits[j]
        prev = digits
    return np.array(svo, dtype=np.uint8)


def svo_decode(svo, depth: int) -> np.ndarray:
    """Ascending codes of a pre-order octree (svo.cpp:76-137)."""
    svo = np.asarray(svo, dtype=np.uint8)
    out: List[int] = []
    pos = 0

    def visit(level: int, prefix: int):
        nonlocal pos
        node = int(svo[pos])
        pos += 1
        for i in range(8):
            if node >> i & 1:
                if level == depth - 1:
                    out.append((prefix << 3) | i)
                else:
                    visit(level + 1, (prefix << 3) | i)

    visit(0, 0)
    return np.array(out, dtype=np.int64).astype(np.uint32).view(np.int32)


def parent_encode(svo, depth: int, attr) -> np.ndarray:
    """Deltas between consecutive children of every octree node
    (filter_parent.cpp:10-87).  Row 0 holds the root's value (its first
    leaf's attributes); each internal node with c children owns c - 1 rows,
    in pre-order, holding child[i] - child[i-1] (uint8 wrap), where a node's
    value is its first leaf's."""
    svo = np.asarray(svo, dtype=np.uint8)
    attr = np.asarray(attr, dtype=np.uint8)
    N, C = attr.shape
    delta = np.zeros((N, C), np.uint8)
    state = {"svo": 0, "leaf": 0, "row": 1}

    def visit(level: int) -> np.ndarray:
        if level == depth:
            v = attr[state["leaf"]].copy()
            state["leaf"] += 1
            return v
        node = int(svo[state["svo"]])
        state["svo"] += 1
        cnt = bin(node).count("1")
        first_row = state["row"]
        state["row"] += cnt - 1
        prev = None
        value = None
        for i in range(cnt):
            v = visit(level + 1)
            if i == 0:
                value = v
            else:
                delta[first_row + i - 1] = (v.astype(np.int64) - prev) & 0xFF
            prev = v.astype(np.int64)
        return value

    delta[0] = visit(0)
    return delta


def parent_decode(svo, depth: int, delta) -> np.ndarray:
    """Inverse of parent_encode (filter_parent.cpp:90-164)."""
    svo = np.asarray(svo, dtype=np.uint8)
    delta = np.asarray(delta, dtype=np.uint8)
    N, C = delta.shape
    attr = np.zeros((N, C), np.uint8)
    state = {"svo": 0, "leaf": 0, "row": 1}

    def visit(level: int, value: np.ndarray):
        if level == depth:
            attr[state["leaf"]] = value
            state["leaf"] += 1
            return
        node = int(svo[state["svo"]])
        state["svo"] += 1
        cnt = bin(node).count("1")
        first_row = state["row"]
        state["row"] += cnt - 1
        v = value.copy()
        for i in range(cnt):
            if i > 0:
                v = ((v.astype(np.int64) + delta[first_row + i - 1]) & 0xFF).astype(np.uint8)
            visit(level + 1, v)

    visit(0, delta[0].copy())
    return attr


# lower neighbours in increasing priority (the last one present wins),
# filter_neighbor.cpp:46-67
_NEIGHBOURS = [(-1, 0, 0), (0, -1, 0), (0, 0, -1), (-1, -1, 0), (-1, 0, -1), (0, -1, -1), (-1, -1, -1)]


def _neighbour_rows(coord: np.ndarray, res: int) -> Tuple[np.ndarray, np.ndarray]:
    """(cell row of each voxel, chosen lower-neighbour cell row or -1)."""
    coord = np.asarray(coord, dtype=np.int64)
    cell = np.full(res * res * res, -1, np.int64)
    key = (coord[:, 2] * res + coord[:, 1]) * res + coord[:, 0]
    cell[key] = np.arange(len(coord))  # duplicates: the last row wins
    own = cell[key]
    nb = np.full(len(coord), -1, np.int64)
    for dx, dy, dz in _NEIGHBOURS:
        x, y, z = coord[:, 0] + dx, coord[:, 1] + dy, coord[:, 2] + dz
        ok = (x >= 0) & (y >= 0) & (z >= 0)
        r = np.where(ok, cell[np.where(ok, (z * res + y) * res + x, 0)], -1)
        nb = np.where(r >= 0, r, nb)
    return own, nb


def neighbor_encode(coord, res: int, attr) -> np.ndarray:
    """attr of each voxel minus that of its chosen lower neighbour
    (filter_neighbor.cpp:18-99): neighbours are taken before any delta is
    formed, as the reference's descending sweep does."""
    attr = np.asarray(attr, dtype=np.uint8)
    own, nb = _neighbour_rows(coord, res)
    base = attr[own].astype(np.int64)
    sub = np.where(nb[:, None] >= 0, attr[np.maximum(nb, 0)].astype(np.int64), 0)
    return ((base - sub) & 0xFF).astype(np.uint8)


def neighbor_decode(coord, res: int, delta) -> np.ndarray:
    """Inverse sweep in ascending (z, y, x) order (filter_neighbor.cpp:110-178)."""
    coord = np.asarray(coord, dtype=np.int64)
    delta = np.asarray(delta, dtype=np.uint8)
    own, nb = _neighbour_rows(coord, res)
    val = delta.astype(np.int64).copy()  # per row; rows sharing a cell read the cell's row
    done = np.zeros(len(coord), bool)
    order = np.argsort(coord[:, 0] + coord[:, 1] + coord[:, 2], kind="stable")
    s = (coord[:, 0] + coord[:, 1] + coord[:, 2])[order]
    # a neighbour's coordinate sum is strictly smaller: finish sums in order
    for lo, hi in _runs(s):
        rows = order[lo:hi]
        cells = own[rows]
        add = np.where(nb[cells, None] >= 0, val[np.maximum(nb[cells], 0)], 0)
        val[cells] = (delta[cells].astype(np.int64) + add) & 0xFF
        done[cells] = True
    return val[own].astype(np.uint8)


def _runs(s: np.ndarray):
    if len(s) == 0:
        return
    cut = np.flatnonzero(np.diff(s)) + 1
    b = np.concatenate([[0], cut, [len(s)]])
    for i in range(len(b) - 1):
        yield int(b[i]), int(b[i + 1])


# ---------------------------------------------------------------------------
# file
# ---------------------------------------------------------------------------
def compress(data: bytes, algo: str, level: int) -> bytes:
    if algo == "none":
        return data
    if algo == "deflate":
        c = zlib.compressobj(level, wbits=-15)
        return c.compress(data) + c.flush()
    if algo == "lzma":
        c = lzma.LZMACompressor(format=lzma.FORMAT_RAW, filters=[{"id": lzma.FILTER_LZMA2, "preset": level}])
        return c.compress(data) + c.flush()
    raise ValueError(f"compression {algo} not available to the oracle")


def decompress(data: bytes, algo: str, level: int) -> bytes:
    if algo == "none":
        return data
    if algo == "deflate":
        d = zlib.decompressobj(wbits=-15)
        return d.decompress(data) + d.flush()
    if algo == "lzma":
        return lzma.LZMADecompressor(format=lzma.FORMAT_RAW,
                                     filters=[{"id": lzma.FILTER_LZMA2, "preset": level}]).decompress(data)
    raise ValueError(f"compression {algo} not available to the oracle")


def write_vxz_bytes(coord, attr: Dict[str, np.ndarray], chunk_size: int = 256, filter: str = "none",
                    compression: str = "lzma", compression_level=None, attr_interleave: str = "as_is") -> bytes:
    """The bytes write_vxz produces (vxz.py:223-365)."""
    coord = np.asarray(coord, dtype=np.int64)
    level = compression_level or DEFAULT_LEVEL[compression]
    info = {
        "num_voxel": int(coord.shape[0]),
        "chunk_size": chunk_size,
        "filter": filter,
        "compression": compression,
        "compression_level": level,
        "raw_size": int(sum([coord.size * 4] + [v.size for v in attr.values()])),
        "compressed_size": 0,
        "compress_ratio": 0.0,
        "attr_interleave": attr_interleave,
        "attr": [[k, int(v.shape[1])] for k, v in attr.items()],
        "chunks": [],
    }
    depth = int(np.log2(chunk_size))
    ccoord = coord // chunk_size
    local = coord % chunk_size
    uniq, inverse = np.unique(ccoord, axis=0, return_inverse=True)
    inverse = inverse.reshape(-1)
    bin_data = b""
    for idx, cxyz in enumerate(uniq.tolist()):
        m = inverse == idx
        lc = local[m]
        codes = S.z_order_encode(lc[:, 0], lc[:, 1], lc[:, 2])
        o = np.argsort(codes, kind="stable")
        lc, codes = lc[o], codes[o]
        a = np.concatenate([np.asarray(v)[m][o] for v in attr.values()], axis=1).astype(np.uint8)
        svo = svo_encode(codes, depth)
        svo_b = compress(svo.tobytes(), compression, level)
        if filter == "parent":
            a = parent_encode(svo, depth, a)
        elif filter == "neighbor":
            a = neighbor_encode(lc, chunk_size, a)
        streams = {}
        if attr_interleave == "none":
            ch = 0
            for k, chs in info["attr"]:
                for i in range(chs):
                    streams[f"{k}_{i}"] = compress(np.ascontiguousarray(a[:, ch]).tobytes(), compression, level)
                    ch += 1
        elif attr_interleave == "as_is":
            ch = 0
            for k, chs in info["attr"]:
                streams[k] = compress(np.ascontiguousarray(a[:, ch:ch + chs]).tobytes(), compression, level)
                ch += chs
        else:
            streams["attr"] = compress(a.tobytes(), compression, level)
        chunk_info = {"idx": cxyz}
        data = b""
        chunk_info["svo"] = [len(data), len(svo_b)]
        data += svo_b
        for k, v in streams.items():
            chunk_info[k] = [len(data), len(v)]
            data += v
        chunk_info["ptr"] = [len(bin_data), len(data)]
        bin_data += data
        info["chunks"].append(chunk_info)
    info["compressed_size"] = len(bin_data)
    info["compress_ratio"] = info["raw_size"] / info["compressed_size"]
    structure = json.dumps(info).encode()
    return b"VXZ\x00" + struct.pack(">I", len(structure) + 8) + structure + bin_data


def read_vxz_bytes(data: bytes):
    """(coord int32 [N,3], {name: uint8 [N, chs]}) of a .vxz file (vxz.py:120-220)."""
    assert data[:3] == b"VXZ" and data[3] == 0
    start = struct.unpack(">I", data[4:8])[0]
    info = json.loads(data[8:start].decode())
    bin_data = data[start:]
    cs = info["chunk_size"]
    depth = int(np.log2(cs))
    coords, attrs = [], {k: [] for k, _ in info["attr"]}
    for ci in info["chunks"]:
        chunk = bin_data[ci["ptr"][0]:ci["ptr"][0] + ci["ptr"][1]]
        dec = {k: np.frombuffer(decompress(chunk[v[0]:v[0] + v[1]], info["compression"], info["compression_level"]),
                                np.uint8) for k, v in ci.items() if k not in ("ptr", "idx")}
        codes = svo_decode(dec["svo"], depth)
        lc = np.stack(S.z_order_decode(codes), 1)
        if info["attr_interleave"] == "none":
            a = np.stack([dec[f"{k}_{i}"] for k, chs in info["attr"] for i in range(chs)], 1)
        elif info["attr_interleave"] == "as_is":
            a = np.concatenate([dec[k].reshape(-1, chs) for k, chs in info["attr"]], 1)
        else:
            a = dec["attr"].reshape(-1, sum(chs for _, chs in info["attr"]))
        if info["filter"] == "parent":
            a = parent_decode(dec["svo"], depth, a)
        elif info["filter"] == "neighbor":
            a = neighbor_decode(lc, cs, a)
        coords.append(lc + np.array(ci["idx"]) * cs)
        ch = 0
        for k, chs in info["attr"]:
            attrs[k].append(a[:, ch:ch + chs])
            ch += chs
    coord = np.concatenate(coords).astype(np.int32) if coords else np.zeros((0, 3), np.int32)
    return coord, {k: np.concatenate(v) for k, v in attrs.items()}
