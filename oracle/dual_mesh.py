"""TEST INFRASTRUCTURE -- numpy restatement of the dual grid -> mesh path and
of the reference's generic hashmap.

Only ``tests/`` and ``bench.py``'s baseline leg use this module, as the
checker.  The product package never imports it.

Restated from
  o_voxel/convert/flexible_dual_grid.py:331-477  (flexible_dual_grid_to_mesh)
  src/hash/hash.cuh:1-85, src/hash/hash.cu        (hashmap insert / lookup)
with sequential semantics: where the reference's GPU insertions race
(duplicate keys), the last inserted row wins here.
"""

from __future__ import annotations

from typing import Optional, Sequence, Tuple

import numpy as np

M32 = 0xFFFFFFFF
M64 = 0xFFFFFFFFFFFFFFFF


def fmix32(k: int) -> int:
    """Murmur3 32-bit finalizer (hash.cuh:2-9)."""
    k &= M32
    k ^= k >> 16
    k = (k * 0x85EBCA6B) & M32
    k ^= k >> 13
    k = (k * 0xC2B2AE35) & M32
    k ^= k >> 16
    return k


def fmix64(k: int) -> int:
    """Murmur3 64-bit finalizer (hash.cuh:13-20)."""
    k &= M64
    k ^= k >> 33
    k = (k * 0xFF51AFD7ED558CCD) & M64
    k ^= k >> 33
    k = (k * 0xC4CEB9FE1A85EC53) & M64
    k ^= k >> 33
    return k


def flat_key(b: int, x: int, y: int, z: int, W: int, H: int, D: int) -> int:
    """b*W*H*D + x*H*D + y*D + z in 64-bit unsigned arithmetic, every int
    sign-extended first (hash.cu insert_3d / lookup_3d kernels)."""
    u = lambda v: v & M64  # noqa: E731
    return (u(b) * u(W) * u(H) * u(D) + u(x) * u(H) * u(D) + u(y) * u(D) + u(z)) & M64


class RefHashmap:
    """The reference's open-addressing table (hash.cuh:23-85), sequential:
    keys[cap] / values[cap], all-ones empty key, hash % cap, linear probing."""

    def __init__(self, capacity: int, key_bits: int = 32, val_bits: int = 32, keys=None, values=None):
        self.cap = int(capacity)
        self.kb, self.vb = key_bits, val_bits
        self.empty = (1 << key_bits) - 1
        self.missing = (1 << val_bits) - 1
        self.keys = [self.empty] * self.cap if keys is None else [int(k) for k in keys]
        self.vals = [0] * self.cap if values is None else [int(v) for v in values]

    def _hash(self, key: int) -> int:
        return (fmix32(key) if self.kb == 32 else fmix64(key)) % self.cap

    def insert(self, key: int, val: int) -> bool:
        key &= self.empty
        slot = self._hash(key)
        for _ in range(self.cap):
            prev = self.keys[slot]
            if prev == self.empty:
                self.keys[slot] = key
            if prev == self.empty or prev == key:
                self.vals[slot] = val & self.missing
                return True
            slot = (slot + 1) % self.cap
        return False

    def lookup(self, key: int) -> int:
        key &= self.empty
        slot = self._hash(key)
        for _ in range(self.cap):
            k = self.keys[slot]
            if k == self.empty:
                return self.missing
            if k == key:
                return self.vals[slot]
            slot = (slot + 1) % self.cap
        return self.missing

    def lookup_3d(self, c: Sequence[int], W: int, H: int, D: int) -> int:
        b, x, y, z = (int(v) for v in c)
        if x < 0 or x >= W or y < 0 or y >= H or z < 0 or z >= D:
            return self.missing
        return self.lookup(flat_key(b, x, y, z, W, H, D))


# The four voxels around the grid edge of each axis, in quad order
# (flexible_dual_grid.py:367-374).
EDGE_NEIGHBORS = np.array(
    [
        [[0, 0, 0], [0, 0, 1], [0, 1, 1], [0, 1, 0]],
        [[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]],
        [[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0]],
    ],
    dtype=np.int64,
)


def mesh_vertices(coords: np.ndarray, dual: np.ndarray, voxel_size, aabb_min) -> np.ndarray:
    """(coords + dual) * voxel_size + aabb_min, float32, one rounding per
    operation (flexible_dual_grid.py:437)."""
    t = coords.astype(np.float32) + dual.astype(np.float32)
    t = t * np.asarray(voxel_size, dtype=np.float32)
    return (t + np.asarray(aabb_min, dtype=np.float32)).astype(np.float32)


def quads_of(coords: np.ndarray, flags: np.ndarray, grid_size) -> np.ndarray:
    """int64 [L, 4] rows of the quads whose four voxels exist, in (voxel,
    axis) order (flexible_dual_grid.py:422-431)."""
    W, H, D = (int(g) for g in grid_size)
    wide = W * H * D >= (1 << 32)
    kmask = M64 if wide else M32
    table = {}
    for i, (x, y, z) in enumerate(coords.tolist()):
        table[flat_key(0, x, y, z, W, H, D) & kmask] = i
    out = []
    fl = np.asarray(flags).astype(bool)
    for i, (x, y, z) in enumerate(coords.tolist()):
        for a in range(3):
            if not fl[i, a]:
                continue
            q = []
            for dx, dy, dz in EDGE_NEIGHBORS[a].tolist():
                xx, yy, zz = x + dx, y + dy, z + dz
                if not (0 <= xx < W and 0 <= yy < H and 0 <= zz < D):
                    break
                r = table.get(flat_key(0, xx, yy, zz, W, H, D) & kmask)
                if r is None:
                    break
                q.append(r)
            if len(q) == 4:
                out.append(q)
    return np.array(out, dtype=np.int64).reshape(-1, 4)


def _cross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.stack(
        [a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
         a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=1)


def _abs_dot(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    p = a * b
    return np.abs((p[:, 0] + p[:, 1]) + p[:, 2])


def round_to(x: np.ndarray, dtype: str) -> np.ndarray:
    """float32 values rounded (to nearest even) to "float32", "float16" or
    "bfloat16" and widened back to float32."""
    x = np.asarray(x, dtype=np.float32)
    if dtype == "float16":
        return x.astype(np.float16).astype(np.float32)
    if dtype == "bfloat16":
        b = x.view(np.uint32).astype(np.uint64)
        finite = np.isfinite(x)
        r = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
        return np.where(finite, r.view(np.float32), x)
    return x


def split_first(verts: np.ndarray, quads: np.ndarray, split_weight: Optional[np.ndarray] = None,
                weight_dtype: str = "float32") -> np.ndarray:
    """bool [L]: True where the reference picks split 1 = (0,1,2),(0,2,3)
    (flexible_dual_grid.py:439-454).  Without weights: the alignment of the
    normals its code forms from columns (0,1,2) and (1,2,3) of each 6-index
    split pattern, split 1 on a strict win.  With weights: w0*w2 > w1*w3,
    products rounded to the weights' dtype."""
    if split_weight is not None:
        w = np.asarray(split_weight, dtype=np.float32).reshape(-1)[quads]
        return round_to(w[:, 0] * w[:, 2], weight_dtype) > round_to(w[:, 1] * w[:, 3], weight_dtype)
    v = verts.astype(np.float32)[quads]  # [L, 4, 3]
    with np.errstate(invalid="ignore", over="ignore"):
        s1 = v[:, [0, 1, 2, 0, 2, 3]]
        s2 = v[:, [0, 1, 3, 3, 1, 2]]
        align = []
        for s in (s1, s2):
            n0 = _cross(s[:, 1] - s[:, 0], s[:, 2] - s[:, 0])
            n1 = _cross(s[:, 2] - s[:, 1], s[:, 3] - s[:, 1])
            align.append(_abs_dot(n0, n1))
    return align[0] > align[1]


def dual_grid_to_mesh(coords, dual, flags, split_weight, grid_size, voxel_size, aabb_min, train: bool = False,
                      weight_dtype: str = "float32") -> Tuple[np.ndarray, np.ndarray]:
    """(vertices float32, faces) of flexible_dual_grid_to_mesh
    (flexible_dual_grid.py:419-477).  Faces are int32 [2L,3], or int64 [4L,3]
    with ``train``."""
    coords = np.asarray(coords, dtype=np.int64)
    verts = mesh_vertices(coords, np.asarray(dual, dtype=np.float32), voxel_size, aabb_min)
    quads = quads_of(coords, flags, grid_size)
    N, L = len(coords), len(quads)
    if not train:
        first = split_first(verts, quads, split_weight, weight_dtype)
        tri1 = quads[:, [0, 1, 2, 0, 2, 3]]
        tri2 = quads[:, [0, 1, 3, 3, 1, 2]]
        faces = np.where(first[:, None], tri1, tri2).reshape(-1, 3).astype(np.int32)
        return verts, faces
    w = np.asarray(split_weight, dtype=np.float32).reshape(-1)[quads]
    w02 = round_to(w[:, 0] * w[:, 2], weight_dtype)[:, None]
    w13 = round_to(w[:, 1] * w[:, 3], weight_dtype)[:, None]
    qv = verts[quads]
    m02 = (qv[:, 0] + qv[:, 2]) / np.float32(2)
    m13 = (qv[:, 1] + qv[:, 3]) / np.float32(2)
    mid = ((w02 * m02 + w13 * m13) / (w02 + w13)).astype(np.float32)
    q5 = np.concatenate([quads, np.arange(N, N + L, dtype=np.int64)[:, None]], axis=1)
    faces = q5[:, [0, 1, 4, 1, 2, 4, 2, 3, 4, 3, 0, 4]].reshape(-1, 3)
    return np.concatenate([verts, mid], axis=0), faces
