"""GPU parity of the fold's slow paths against the CPU oracle: bricks with
more records than the LDS tile of the main fold kernel (folded by voxel
ranges) and single voxels with more records than any tile (folded in
key-ordered batches).  Same bar as the other parity tests: occupancy and
flags exact, duals bit-identical."""

import numpy as np
import pytest
import torch

from test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def _grid(G):
    return [1.0 / G] * 3, [0, 0, 0, G, G, G]


def _fan(n, center, radius, rng):
    """n triangles sharing one vertex at `center`, their other vertices on a
    small circle in a tilted plane: every triangle crosses the voxels around
    the centre."""
    ang = np.sort(rng.uniform(0, 2 * np.pi, n + 1))
    u = np.array([1.0, 0.3, 0.2])
    u /= np.linalg.norm(u)
    w = np.cross(u, [0.1, 1.0, 0.4])
    w /= np.linalg.norm(w)
    ring = center + radius * (np.cos(ang)[:, None] * u + np.sin(ang)[:, None] * w)
    v = np.concatenate([center[None], ring]).astype(np.float32)
    f = np.stack([np.zeros(n, np.int64), np.arange(1, n + 1), np.arange(2, n + 2)], 1)
    return v, f


def _merge(parts):
    vo, fo, off = [], [], 0
    for v, f in parts:
        vo.append(np.asarray(v, np.float32))
        fo.append(np.asarray(f, np.int64) + off)
        off += len(v)
    return np.concatenate(vo), np.concatenate(fo)


def test_giant_voxel(oracle):
    """A tilted sheet makes voxel (13, 17, 9) active; ~9000 triangles inside
    that voxel then give it more face records than any LDS tile holds."""
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    rng = np.random.default_rng(1)
    G = 32
    vs, gr = _grid(G)
    sheet_v = np.array([[2, 2, 9.3], [30, 3, 9.7], [29, 30, 9.6], [3, 29, 9.4]], np.float64) / G
    sheet = (sheet_v, np.array([[0, 1, 2], [0, 2, 3]]))
    center = np.array([13.5, 17.5, 9.5]) / G
    v, f = _merge([sheet, _fan(9000, center, 0.3 / G, rng)])
    compare(oracle, v, f, vs, gr, label="giant voxel")
    from o_voxel import _C

    assert _C.last_stats()["giant_voxels"] >= 1


def test_large_brick_ranges(oracle):
    """20000 triangles of about a voxel inside one 8^3 brick: more records
    than the large-brick tile, folded by voxel ranges."""
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    rng = np.random.default_rng(2)
    G = 64
    vs, gr = _grid(G)
    n = 20000
    c = np.array([16.0, 24.0, 40.0]) + rng.uniform(1.0, 7.0, (n, 1, 3))
    v = (c + rng.uniform(-1.0, 1.0, (n, 3, 3))).reshape(-1, 3) / G
    f = np.arange(3 * n).reshape(n, 3)
    compare(oracle, v.astype(np.float32), f, vs, gr, label="large brick")
    from o_voxel import _C

    st = _C.last_stats()
    assert st["large_bricks"] >= 1 and st["large_brick_records"] > 4096
