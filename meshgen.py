"""Seeded synthetic meshes for the mesh_to_flexible_dual_grid workloads.

All meshes live in the aabb [-0.5, 0.5]^3 with float32 vertices and int64
faces (what a caller of ``o_voxel.convert.mesh_to_flexible_dual_grid``
passes).  Workloads:

* ``ico_cap``      icosphere r=0.4 (subdivision 4, 2,562 V / 5,120 F) plus an
                   open hemispherical cap of 1,000 F offset +0.05 z (two
                   boundary loops); grid 64.
* ``scene``        48 randomly rotated, anisotropically scaled closed
                   icospheres and tori (radii 0.02-0.2), 32 open planar or
                   quadratic patches with free boundaries, and 1% sliver
                   triangles (one edge ~1e-6 of the others); texture seams
                   split the shapes into small charts (~0.95 vertices per
                   face); ~280k F / ~263k V; grid 512.
* ``scene_large``  the ``scene`` generator at a larger face budget, then two
                   1-to-4 midpoint subdivisions with Gaussian vertex jitter
                   (sigma 1e-4), trimmed by a seeded face drop to 10.3M F.
* ``scene_rankdef`` ``scene`` plus large coplanar sheets (rank-1 quadrics),
                   90-degree creases (rank-2) and cube corners (rank-3).
* ``batch``        many ``scene``-generator meshes with log-uniform face
                   counts and per-mesh grid sizes.
"""

from __future__ import annotations

import functools
import math
from typing import List, Tuple

import numpy as np

Mesh = Tuple[np.ndarray, np.ndarray]  # (vertices f32 [V,3], faces i64 [F,3])


# ---------------------------------------------------------------------------
# primitives
# ---------------------------------------------------------------------------
def _icosahedron() -> Mesh:
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def subdivide_midpoint(v: np.ndarray, f: np.ndarray) -> Mesh:
    """1-to-4 midpoint subdivision with shared (welded) edge midpoints."""
    F = f.shape[0]
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
    e = np.sort(e, axis=1)
    key = e[:, 0] * (int(v.shape[0]) + 1) + e[:, 1]
    uniq, inv = np.unique(key, return_inverse=True)
    a = uniq // (v.shape[0] + 1)
    b = uniq % (v.shape[0] + 1)
    mids = 0.5 * (v[a] + v[b])
    base = v.shape[0]
    m = (inv + base).reshape(3, F).T  # m01, m12, m20
    v2 = np.concatenate([v, mids], axis=0)
    f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
    m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
    nf = np.concatenate([np.stack([f0, m01, m20], 1), np.stack([m01, f1, m12], 1), np.stack([m20, m12, f2], 1),
                         np.stack([m01, m12, m20], 1)], axis=0)
    return v2, nf


@functools.lru_cache(maxsize=None)
def icosphere(level: int) -> Mesh:
    """Unit icosphere; the arrays are cached and shared, callers must not modify them in place."""
    v, f = _icosahedron()
    for _ in range(level):
        v, f = subdivide_midpoint(v, f)
        v = v / np.linalg.norm(v, axis=1, keepdims=True)
    return v, f


def torus(R: float, r: float, nu: int, nv: int) -> Mesh:
    u = np.arange(nu) * (2 * np.pi / nu)
    w = np.arange(nv) * (2 * np.pi / nv)
    uu, ww = np.meshgrid(u, w, indexing="ij")
    x = (R + r * np.cos(ww)) * np.cos(uu)
    y = (R + r * np.cos(ww)) * np.sin(uu)
    z = r * np.sin(ww)
    v = np.stack([x, y, z], -1).reshape(-1, 3)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    f = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)], 0)
    return v, f.astype(np.int64)


def grid_patch(n: int, curv: Tuple[float, float, float]) -> Mesh:
    """Open n x n quad patch over [-1,1]^2, z = a x^2 + b y^2 + c x y."""
    s = np.linspace(-1.0, 1.0, n + 1)
    xx, yy = np.meshgrid(s, s, indexing="ij")
    a, b, c = curv
    zz = a * xx * xx + b * yy * yy + c * xx * yy
    v = np.stack([xx, yy, zz], -1).reshape(-1, 3)
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    p00 = i * (n + 1) + j
    p10 = (i + 1) * (n + 1) + j
    p11 = (i + 1) * (n + 1) + j + 1
    p01 = i * (n + 1) + j + 1
    f = np.concatenate([np.stack([p00, p10, p11], -1).reshape(-1, 3), np.stack([p00, p11, p01], -1).reshape(-1, 3)])
    return v, f.astype(np.int64)


def uv_cap(radius: float, n_lon: int, n_lat: int, lat_max: float) -> Mesh:
    """Open hemispherical band from the equator toward a pole hole."""
    lon = np.arange(n_lon) * (2 * np.pi / n_lon)
    lat = np.linspace(0.0, lat_max, n_lat + 1)
    ll, tt = np.meshgrid(lon, lat, indexing="ij")
    v = np.stack([radius * np.cos(tt) * np.cos(ll), radius * np.cos(tt) * np.sin(ll), radius * np.sin(tt)],
                 -1).reshape(-1, 3)
    i = np.arange(n_lon)[:, None]
    j = np.arange(n_lat)[None, :]
    a = i * (n_lat + 1) + j
    b = ((i + 1) % n_lon) * (n_lat + 1) + j
    c = ((i + 1) % n_lon) * (n_lat + 1) + j + 1
    d = i * (n_lat + 1) + j + 1
    f = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    return v, f.astype(np.int64)


def _rotation(rng: np.random.Generator) -> np.ndarray:
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _merge(parts: List[Mesh]) -> Mesh:
    vs, fs, off = [], [], 0
    for v, f in parts:
        vs.append(v)
        fs.append(f + off)
        off += v.shape[0]
    return np.concatenate(vs, 0), np.concatenate(fs, 0)


def _finish(v: np.ndarray, f: np.ndarray) -> Mesh:
    v = np.clip(v, -0.4999, 0.4999)
    return np.ascontiguousarray(v, dtype=np.float32), np.ascontiguousarray(f, dtype=np.int64)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def ico_cap() -> Mesh:
    """Icosphere-4 r=0.4 plus a 1,000-face open cap offset +0.05 z."""
    v1, f1 = icosphere(4)
    v1 = v1 * 0.4
    v2, f2 = uv_cap(0.3, 20, 25, 0.45 * np.pi)
    v2 = v2 + np.array([0.0, 0.0, 0.05])
    return _finish(*_merge([(v1, f1), (v2, f2)]))


def split_seams(v: np.ndarray, f: np.ndarray, chart_edges: float) -> Mesh:
    """Texture seams: faces are grouped into charts (cells of a lattice of
    ``chart_edges`` mean edge lengths, by face centroid) and every chart gets
    its own copies of its vertices, as a mesh exported with a fragmented UV
    atlas has.  Edges along chart borders become boundary edges (their two
    sides no longer share vertex indices).  ``chart_edges`` <= 0 keeps the
    mesh welded."""
    if chart_edges <= 0 or f.shape[0] == 0:
        return v, f
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    d = b - a
    el = float(np.sqrt((d * d).sum(axis=1)).mean())
    cent = (a + b + c) / 3.0
    cell = np.floor((cent - cent.min(axis=0)) / (chart_edges * max(el, 1e-12))).astype(np.int64)
    chart = cell[:, 0] + 4096 * (cell[:, 1] + 4096 * cell[:, 2])
    key = chart[:, None] * int(v.shape[0]) + f
    uniq, inv = np.unique(key.ravel(), return_inverse=True)
    return v[uniq % v.shape[0]], inv.reshape(f.shape).astype(np.int64)


def _scene_parts(rng: np.random.Generator, closed_budget: float, patch_budget: float, n_closed: int,
                 n_patches: int, chart_edges: float) -> List[Mesh]:
    radii = rng.uniform(0.02, 0.2, size=n_closed)
    kinds = rng.integers(0, 2, size=n_closed)  # 0 sphere, 1 torus
    areas = np.where(kinds == 0, 4 * np.pi * radii ** 2, 4 * np.pi ** 2 * radii * (0.35 * radii))
    faces_per = closed_budget * areas / areas.sum()
    parts: List[Mesh] = []
    for k in range(n_closed):
        r = radii[k]
        if kinds[k] == 0:
            level = int(np.clip(np.round(np.log(max(faces_per[k], 20.0) / 20.0) / np.log(4.0)), 1, 7))
            v, f = icosphere(level)
            v = v * r
        else:
            rr = 0.35 * r
            n_tot = max(faces_per[k] / 2.0, 24.0)
            nu = max(int(np.sqrt(n_tot * r / rr)), 6)
            nv = max(int(n_tot / nu), 4)
            v, f = torus(r, rr, nu, nv)
        v, f = split_seams(v, f, chart_edges)
        scale = rng.uniform(0.7, 1.3, size=3)
        R = _rotation(rng)
        v = (v * scale[None, :]) @ R.T
        ext = np.abs(v).max(axis=0)
        lo = -0.5 + ext + 1e-3
        hi = 0.5 - ext - 1e-3
        c = rng.uniform(np.minimum(lo, 0.0), np.maximum(hi, 0.0))
        parts.append((v + c[None, :], f))

    per_patch = patch_budget / n_patches
    for k in range(n_patches):
        n = max(int(np.sqrt(per_patch / 2.0)), 2)
        curv = (0.0, 0.0, 0.0) if k % 2 == 0 else tuple(rng.uniform(-0.5, 0.5, size=3))
        v, f = grid_patch(n, curv)
        v, f = split_seams(v, f, chart_edges)
        half = rng.uniform(0.05, 0.25)
        v = v * half
        R = _rotation(rng)
        v = v @ R.T
        ext = np.abs(v).max(axis=0)
        lo = -0.5 + ext + 1e-3
        hi = 0.5 - ext - 1e-3
        c = rng.uniform(np.minimum(lo, 0.0), np.maximum(hi, 0.0))
        parts.append((v + c[None, :], f))
    return parts


def scene(seed: int = 2, target_faces: int = 280_000, n_closed: int = 48, n_patches: int = 32,
          sliver_frac: float = 0.01, chart_edges: float = 3.0) -> Mesh:
    """Closed shapes + open patches + slivers, ~target_faces faces.

    Shapes are tessellated to a face budget; the budget is rescaled once by
    the face count of a first draw (same seed), so the result lands within a
    few percent of ``target_faces``.  Texture seams (``split_seams``, charts of
    ``chart_edges`` mean edge lengths) give about 0.95 vertices per face, as
    in a textured mesh export; 0 keeps every shape welded (about 0.5)."""
    n_sliver = int(round(sliver_frac * target_faces))
    budget = float(target_faces - n_sliver)
    # face counts do not depend on the seams (the split draws no random numbers)
    for _ in range(2):
        rng = np.random.default_rng(seed)
        got = sum(int(p[1].shape[0]) for p in _scene_parts(rng, 0.7 * budget, 0.3 * budget, n_closed, n_patches, 0.0))
        if got == 0 or abs(got - (target_faces - n_sliver)) <= 0.005 * target_faces:
            break
        budget *= (target_faces - n_sliver) / got
    rng = np.random.default_rng(seed)
    parts = _scene_parts(rng, 0.7 * budget, 0.3 * budget, n_closed, n_patches, chart_edges)

    # slivers: one edge ~1e-6 of the other two
    base = rng.uniform(-0.45, 0.45, size=(n_sliver, 3))
    d1 = rng.normal(size=(n_sliver, 3))
    d1 /= np.linalg.norm(d1, axis=1, keepdims=True)
    d2 = rng.normal(size=(n_sliver, 3))
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    L = rng.uniform(0.002, 0.02, size=(n_sliver, 1))
    p0 = base
    p1 = base + L * d1
    p2 = p1 + 1e-6 * L * d2
    sv = np.stack([p0, p1, p2], 1).reshape(-1, 3)
    sf = np.arange(3 * n_sliver, dtype=np.int64).reshape(-1, 3)
    parts.append((sv, sf))

    v, f = _merge(parts)
    return _finish(v, f)


def scene_large(seed: int = 3, target_faces: int = 10_300_000) -> Mesh:
    """scene generator -> two midpoint subdivisions + jitter -> seeded trim."""
    base_faces = int(math.ceil(target_faces / 16.0 * 1.01))
    v, f = scene(seed=seed, target_faces=base_faces)
    v = v.astype(np.float64)
    f = f.astype(np.int64)
    for _ in range(2):
        v, f = subdivide_midpoint(v, f)
    rng = np.random.default_rng(seed + 1000)
    v = v + rng.normal(scale=1e-4, size=v.shape)
    if f.shape[0] > target_faces:
        keep = np.sort(rng.choice(f.shape[0], size=target_faces, replace=False))
        f = f[keep]
    return _finish(v, f)


def scene_rankdef(seed: int = 4, target_faces: int = 280_000) -> Mesh:
    """scene plus coplanar sheets (rank-1), creases (rank-2), cube corners (rank-3)."""
    rng = np.random.default_rng(seed)
    v0, f0 = scene(seed=seed, target_faces=target_faces)
    parts: List[Mesh] = [(v0.astype(np.float64), f0)]
    for k in range(4):  # large axis-aligned coplanar sheets
        v, f = grid_patch(40, (0.0, 0.0, 0.0))
        v = v * 0.45
        perm = [(0, 1, 2), (1, 2, 0), (2, 0, 1)][k % 3]
        v = v[:, perm]
        v[:, perm.index(2)] += rng.uniform(-0.3, 0.3)
        parts.append((v, f))
    for k in range(8):  # 90-degree creases: two half-planes sharing an edge
        v1, f1 = grid_patch(16, (0.0, 0.0, 0.0))
        v1 = v1 * 0.1
        v1[:, 0] = np.abs(v1[:, 0])
        v2 = v1.copy()
        v2[:, [0, 2]] = v1[:, [2, 0]]
        off = rng.uniform(-0.35, 0.35, size=3)
        parts.append((v1 + off, f1))
        parts.append((v2 + off, f1))
    for k in range(8):  # axis-aligned cubes -> corners
        s = rng.uniform(0.03, 0.1)
        c = rng.uniform(-0.35, 0.35, size=3)
        cv = np.array([[x, y, z] for z in (-1, 1) for y in (-1, 1) for x in (-1, 1)], dtype=np.float64) * s + c
        cf = np.array([[0, 2, 1], [1, 2, 3], [4, 5, 6], [5, 7, 6], [0, 1, 4], [1, 5, 4], [2, 6, 3], [3, 6, 7],
                       [0, 4, 2], [2, 4, 6], [1, 3, 5], [3, 7, 5]], dtype=np.int64)
        parts.append((cv, cf))
    return _finish(*_merge(parts))


def batch(n_meshes: int = 256, seed: int = 5, f_lo: int = 500, f_hi: int = 500_000,
          grids=(64, 128, 256, 512)) -> List[Tuple[np.ndarray, np.ndarray, int]]:
    """Meshes with log-uniform face counts and per-mesh grid sizes."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_meshes):
        nf = int(np.exp(rng.uniform(np.log(f_lo), np.log(f_hi))))
        g = int(rng.choice(grids))
        n_closed = int(np.clip(nf // 2000, 1, 48))
        n_patch = int(np.clip(nf // 4000, 1, 32))
        v, f = scene(seed=k, target_faces=nf, n_closed=n_closed, n_patches=n_patch)
        out.append((v, f, g))
    return out


def grid_args(grid_size: int, aabb=((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))):
    """(voxel_size, grid_range, aabb_min) exactly as the reference wrapper derives them
    (o_voxel/convert/flexible_dual_grid.py:129-181) for an explicit aabb."""
    import torch

    a = torch.tensor(np.array(aabb), dtype=torch.float32)
    g = torch.tensor(np.array([grid_size] * 3), dtype=torch.int32)
    vs = ((a[1] - a[0]) / g).to(dtype=torch.float32)
    return [float(x) for x in vs.tolist()], [0, 0, 0] + [int(x) for x in g.tolist()], a[0].numpy()


WORKLOADS = {
    "ico_cap": dict(grid=64),
    "scene": dict(grid=512),
    "scene_large": dict(grid=512),
    "scene_rankdef": dict(grid=1024),
}
