"""The .vxz voxel file (MI355X build).

Same functions, arguments, defaults, assertions and file bytes as the
reference module o_voxel/io/vxz.py.  Writing does the per-voxel work for all
chunks at once on the GPU (chunk split, Morton codes and ordering, octrees,
attribute filters) and compresses each stream on host threads; reading
decompresses and parses each chunk's octree on host threads and does the
per-voxel work (Morton decode, neighbour filter) for all chunks on the GPU.

Format (header, then the compressed streams):
  b"VXZ", version byte 0, big-endian u32 offset of the binary data, JSON
  {"num_voxel", "chunk_size", "filter", "compression", "compression_level",
   "raw_size", "compressed_size", "compress_ratio", "attr_interleave",
   "attr": [[name, channels], ...],
   "chunks": [{"idx": [cx, cy, cz], "svo": [off, len], <stream>: [off, len],
               ..., "ptr": [off, len]}, ...]}
Each chunk holds the chunk_size^3 block at idx * chunk_size: an octree of its
voxels' Morton codes (child masks in pre-order) and the voxels' uint8
attributes in Morton order, optionally filtered ("parent": differences of
octree siblings; "neighbor": differences to a lower neighbour).
"""

from __future__ import annotations

import json
import lzma
import os
import struct
import zlib
from concurrent.futures import ThreadPoolExecutor
from typing import Dict, Literal, Optional, Union

import numpy as np
import torch

from .. import _C
from ..serialize import decode_seq, encode_seq

__all__ = ["read_vxz", "read_vxz_info", "write_vxz"]

DEFAULT_COMPRESION_LEVEL = {"none": 0, "deflate": 9, "lzma": 9, "zstd": 22}


def _zstd():
    try:
        import zstandard
    except ImportError as e:  # not part of this environment
        raise ImportError("compression 'zstd' needs the zstandard package, which is not installed") from e
    return zstandard


def _compress(data: bytes, algo: str, level: Optional[int]) -> bytes:
    if algo == "none":
        return data
    if level is None:
        level = DEFAULT_COMPRESION_LEVEL[algo]
    if algo == "deflate":
        c = zlib.compressobj(level, wbits=-15)
        return c.compress(data) + c.flush()
    if algo == "lzma":
        c = lzma.LZMACompressor(format=lzma.FORMAT_RAW, filters=[{"id": lzma.FILTER_LZMA2, "preset": level}])
        return c.compress(data) + c.flush()
    if algo == "zstd":
        return _zstd().ZstdCompressor(level=level, write_checksum=False, write_content_size=True,
                                      threads=-1).compress(data)
    raise ValueError(f"Invalid compression algorithm: {algo}")


def _decompress(data: bytes, algo: str, level: Optional[int]) -> bytes:
    if algo == "none":
        return data
    if level is None:
        level = DEFAULT_COMPRESION_LEVEL[algo]
    if algo == "deflate":
        d = zlib.decompressobj(wbits=-15)
        return d.decompress(data) + d.flush()
    if algo == "lzma":
        return lzma.LZMADecompressor(format=lzma.FORMAT_RAW,
                                     filters=[{"id": lzma.FILTER_LZMA2, "preset": level}]).decompress(data)
    if algo == "zstd":
        z = _zstd()
        return z.ZstdDecompressor(format=z.FORMAT_ZSTD1).decompress(data)
    raise ValueError(f"Invalid compression algorithm: {algo}")


def _read_all(file) -> bytes:
    if isinstance(file, str):
        with open(file, "rb") as f:
            return f.read()
    return file.read()


def _parse_header(data: bytes):
    assert data[:3] == b"VXZ", "Invalid file type"
    assert data[3] == 0, "Invalid file version"
    start = struct.unpack(">I", data[4:8])[0]
    return json.loads(data[8:start].decode()), start


def read_vxz_info(file) -> Dict:
    """The JSON header of a .vxz file (reference vxz.py:101-119)."""
    return _parse_header(_read_all(file))[0]


def read_vxz(file, num_threads: int = -1):
    """(coord int32 [N, 3], {name: uint8 [N, channels]}) of a .vxz file
    (reference vxz.py:122-220), chunk after chunk, Morton order inside."""
    data = _read_all(file)
    num_threads = num_threads if num_threads > 0 else os.cpu_count()
    info, start = _parse_header(data)
    bin_data = memoryview(data)[start:]  # slices below are views, not copies
    cs = info["chunk_size"]
    depth = np.log2(cs)
    assert depth.is_integer(), f"Chunk size must be a power of 2, got {cs}"
    depth = int(depth)
    algo, level = info["compression"], info["compression_level"]
    names = [k for k, _ in info["attr"]]
    chs_of = [int(c) for _, c in info["attr"]]
    all_chs = sum(chs_of)

    def worker(ci):
        chunk = bin_data[ci["ptr"][0]:ci["ptr"][0] + ci["ptr"][1]]
        dec = {k: np.frombuffer(_decompress(chunk[v[0]:v[0] + v[1]], algo, level), dtype=np.uint8)
               for k, v in ci.items() if k not in ("ptr", "idx")}
        codes = _C.svo_decode_host(dec["svo"], depth)
        n = len(codes)
        if info["attr_interleave"] == "none":
            cols = [dec[f"{k}_{i}"] for k, c in zip(names, chs_of) for i in range(c)]
            a = np.stack(cols, 1) if cols else np.zeros((n, 0), np.uint8)
        elif info["attr_interleave"] == "as_is":
            parts = [dec[k].reshape(-1, c) for k, c in zip(names, chs_of)]
            a = np.concatenate(parts, 1) if parts else np.zeros((n, 0), np.uint8)
        else:
            a = dec["attr"].reshape(-1, all_chs)
        if info["filter"] == "parent" and all_chs:
            a = _C.svo_parent_decode_host(dec["svo"], depth, a)
        return codes, np.ascontiguousarray(a)

    chunks = info["chunks"]
    if num_threads == 1:
        parts = [worker(ci) for ci in chunks]
    else:
        with ThreadPoolExecutor(max_workers=num_threads) as ex:
            parts = list(ex.map(worker, chunks))
    if not parts:
        return torch.zeros((0, 3), dtype=torch.int32), {k: torch.zeros((0, c), dtype=torch.uint8)
                                                       for k, c in zip(names, chs_of)}
    counts = [len(p[0]) for p in parts]
    dev = _C._gpu_for_host_call()
    codes = torch.from_numpy(np.concatenate([p[0] for p in parts])).to(dev)
    attr = torch.from_numpy(np.concatenate([p[1] for p in parts])).to(dev)
    local = decode_seq(codes)
    chunk = torch.repeat_interleave(torch.arange(len(parts), dtype=torch.int32, device=dev),
                                    torch.tensor(counts, device=dev))
    if info["filter"] == "neighbor" and all_chs:
        attr = _C.svo_neighbor_filter(local, chunk, len(parts), cs, attr, decode=True)
    origin = torch.tensor([ci["idx"] for ci in chunks], dtype=torch.int32, device=dev) * cs
    coord = (local + origin[chunk.long()]).cpu()
    attr = attr.cpu()
    out, ch = {}, 0
    for k, c in zip(names, chs_of):
        out[k] = attr[:, ch:ch + c]
        ch += c
    return coord, out


def write_vxz(
    file,
    coord: torch.Tensor,
    attr: Dict[str, torch.Tensor],
    chunk_size: int = 256,
    filter: Literal["none", "parent", "neighbor"] = "none",
    compression: Literal["none", "deflate", "lzma", "zstd"] = "lzma",
    compression_level: Optional[int] = None,
    attr_interleave: Literal["none", "as_is", "all"] = "as_is",
    num_threads: int = -1,
):
    """Write voxels and their uint8 attributes as a .vxz file (reference
    vxz.py:223-365); the bytes are the reference writer's.  Coordinates must
    be distinct."""
    for k, v in attr.items():
        assert coord.shape[0] == v.shape[0], f"Number of coordinates and attributes do not match for key {k}"
        assert v.dtype == torch.uint8, f"Attributes must be uint8, got {v.dtype} for key {k}"
    assert attr_interleave in ["none", "as_is", "all"], f"Invalid attr_interleave value: {attr_interleave}"
    compression_level = compression_level or DEFAULT_COMPRESION_LEVEL[compression]
    num_threads = num_threads if num_threads > 0 else os.cpu_count()
    info = {
        "num_voxel": coord.shape[0],
        "chunk_size": chunk_size,
        "filter": filter,
        "compression": compression,
        "compression_level": compression_level,
        "raw_size": sum([coord.numel() * 4] + [v.numel() for v in attr.values()]),
        "compressed_size": 0,
        "compress_ratio": 0.0,
        "attr_interleave": attr_interleave,
        "attr": [[k, v.shape[1]] for k, v in attr.items()],
        "chunks": [],
    }
    depth = np.log2(chunk_size)
    assert depth.is_integer(), f"Chunk size must be a power of 2, got {chunk_size}"
    depth = int(depth)

    # per-voxel work, all chunks at once on the device
    dev = _C._gpu_for_host_call() if not coord.is_cuda else coord.device
    c = coord.to(dev)
    ccoord = c // chunk_size
    local = (c % chunk_size).int()
    uniq, inverse = torch.unique(ccoord, dim=0, return_inverse=True)
    n_chunks = uniq.shape[0]
    codes = encode_seq(local)
    key = (inverse.long() << 30) | codes.long()
    order = torch.argsort(key)
    skey = key[order]
    if skey.numel() > 1 and bool((skey[1:] == skey[:-1]).any()):
        raise ValueError("write_vxz: coordinates must be distinct")
    chunk = inverse[order].int()
    codes = codes[order].contiguous()
    local = local[order].contiguous()
    off = torch.zeros((n_chunks + 1,), dtype=torch.int64, device=dev)
    off[1:] = torch.cumsum(torch.bincount(chunk.long(), minlength=n_chunks), 0)
    parts = [v.to(dev)[order] for v in attr.values()]
    a = torch.cat(parts, dim=1).contiguous() if parts else torch.zeros((c.shape[0], 0), dtype=torch.uint8,
                                                                      device=dev)
    svo, svo_off = _C.svo_encode_chunks(codes, chunk, off, n_chunks, depth)
    if a.shape[1]:
        if filter == "parent":
            a = _C.svo_parent_encode(a)
        elif filter == "neighbor":
            a = _C.svo_neighbor_filter(local, chunk, n_chunks, chunk_size, a, decode=False)
    svo_h = svo.cpu().numpy()
    a_h = a.cpu().numpy()
    off_h = off.cpu().tolist()
    svo_off_h = svo_off.tolist()
    chunk_xyz = uniq.cpu().tolist()

    # streams, compressed on host threads
    def worker(k):
        s0, s1 = svo_off_h[k], svo_off_h[k + 1]
        ak = a_h[off_h[k]:off_h[k + 1]]
        streams = {"svo": _compress(svo_h[s0:s1].tobytes(), compression, compression_level)}
        if attr_interleave == "none":
            ch = 0
            for name, chs in info["attr"]:
                for i in range(chs):
                    streams[f"{name}_{i}"] = _compress(ak[:, ch].tobytes(), compression, compression_level)
                    ch += 1
        elif attr_interleave == "as_is":
            ch = 0
            for name, chs in info["attr"]:
                streams[name] = _compress(ak[:, ch:ch + chs].tobytes(), compression, compression_level)
                ch += chs
        else:
            streams["attr"] = _compress(ak.tobytes(), compression, compression_level)
        ci, parts, size = {"idx": chunk_xyz[k]}, [], 0
        for name, v in streams.items():
            ci[name] = [size, len(v)]
            parts.append(v)
            size += len(v)
        return ci, parts, size

    with ThreadPoolExecutor(max_workers=num_threads) as ex:
        chunks = list(ex.map(worker, range(n_chunks)))
    pieces, total = [], 0
    for ci, parts, size in chunks:
        ci["ptr"] = [total, size]
        pieces += parts
        total += size
        info["chunks"].append(ci)
    bin_data = b"".join(pieces)
    info["compressed_size"] = len(bin_data)
    info["compress_ratio"] = info["raw_size"] / info["compressed_size"]
    structure = json.dumps(info).encode()
    header = b"VXZ\x00" + struct.pack(">I", len(structure) + 8)
    if isinstance(file, str):
        with open(file, "wb") as f:
            f.write(header)
            f.write(structure)
            f.write(bin_data)
    else:
        file.write(header)
        file.write(structure)
        file.write(bin_data)
