"""GPU: the launch-class thresholds of the scanline and face-test kernels
and the brick-size threshold of the record grouping change launch shapes
only.  Every mesh forced into each extreme class gives the same outputs, bit
for bit, as the default classes; the autotuner keeps one of its candidates on
the context; invalid thresholds are rejected."""

import numpy as np
import pytest
import torch

import meshgen

pytestmark = pytest.mark.gpu

AABB = [[-0.5, -0.5, -0.5], [0.5, 0.5, 0.5]]
DEFAULT_MARK = [0.5, 2.0, 8.0]
DEFAULT_FACE = [4.0, 16.0]
DEFAULT_BRICK = 3.0


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda", 0)


def _meshes(dev):
    out = meshgen.batch(n_meshes=8, seed=11, f_lo=200, f_hi=30000, grids=(32, 64, 128, 256))
    return ([(torch.from_numpy(v).to(dev), torch.from_numpy(f.astype(np.int64)).to(dev)) for v, f, _ in out],
            [g for _, _, g in out])


def _run(ms, grids):
    import o_voxel

    out = o_voxel.convert.mesh_to_flexible_dual_grid_batch(ms, grid_sizes=grids, aabb=AABB, face_weight=1.0,
                                                           boundary_weight=0.2, regularization_weight=1e-2)
    torch.cuda.synchronize()
    return [[t.cpu().numpy() for t in o] for o in out]


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        for u, w in zip(x, y):
            assert u.shape == w.shape
            assert np.array_equal(u.view(np.uint8), w.view(np.uint8))


# every mesh in one scanline class / one face-test class
EXTREMES = [
    ("all 64-item merged scanline tiles", [1e30, 1e30, 1e30], None),
    ("all 128-item merged scanline tiles", [0.0, 1e30, 1e30], None),
    ("all 128-item plain scanline tiles", [0.0, 0.0, 1e30], None),
    ("all 256-item scanline tiles", [0.0, 0.0, 0.0], None),
    ("all 32-face tiles", None, [0.0, 0.0]),
    ("all 64-face tiles", None, [0.0, 1e30]),
    ("all one-lane-per-face tiles", None, [1e30, 1e30]),
]


def test_tile_classes_do_not_change_results():
    from o_voxel import _C

    dev = _dev()
    ms, grids = _meshes(dev)
    _C.set_tile_classes(dev, DEFAULT_MARK, DEFAULT_FACE)
    want = _run(ms, grids)
    try:
        for name, mark, face in EXTREMES:
            print(f"[case] {name}", flush=True)
            _C.set_tile_classes(dev, mark, face)
            _same(_run(ms, grids), want)
            _C.set_tile_classes(dev, DEFAULT_MARK, DEFAULT_FACE)
        for fpv in (0.0, 1e30):
            print(f"[case] every mesh in {'4x4x4' if fpv == 0.0 else '8x8x8'} bricks", flush=True)
            _C.set_brick_class(dev, fpv)
            _same(_run(ms, grids), want)
    finally:
        _C.set_tile_classes(dev, DEFAULT_MARK, DEFAULT_FACE)
        _C.set_brick_class(dev, DEFAULT_BRICK)


def test_autotune_keeps_a_candidate():
    import o_voxel
    from o_voxel import _C

    dev = _dev()
    ms, grids = _meshes(dev)
    _C.set_tile_classes(dev, DEFAULT_MARK, DEFAULT_FACE)
    want = _run(ms, grids)
    try:
        print("[case] autotune over scales 0.5, 1, 2", flush=True)
        res = o_voxel.convert.autotune_tile_classes(ms, scales=(0.5, 1.0, 2.0), repeats=1, grid_sizes=grids,
                                                    aabb=AABB, face_weight=1.0, boundary_weight=0.2,
                                                    regularization_weight=1e-2)
        marks = [[x * s for x in DEFAULT_MARK] for s in (0.5, 1.0, 2.0)]
        faces = [[x * s for x in DEFAULT_FACE] for s in (0.5, 1.0, 2.0)]
        assert res["mark_fpc"] in marks and res["face_vpf"] in faces
        assert res["brick_fpv"] in [DEFAULT_BRICK * s for s in (0.5, 1.0, 2.0)]
        assert len(res["trials"]) == 9
        assert all(np.isfinite(t.get("mark_ms", t.get("face_ms", t.get("fold_ms")))) for t in res["trials"])
        assert _C.get_tile_classes(dev) == (res["mark_fpc"], res["face_vpf"])
        assert _C.get_brick_class(dev) == res["brick_fpv"]
        _same(_run(ms, grids), want)
    finally:
        _C.set_tile_classes(dev, DEFAULT_MARK, DEFAULT_FACE)
        _C.set_brick_class(dev, DEFAULT_BRICK)


@pytest.mark.parametrize("mark,face", [([2.0, 1.0, 8.0], None), ([-1.0, 2.0, 8.0], None), (None, [4.0, 1.0]),
                                       (None, [float("nan"), 4.0]), ([0.5, 2.0, float("inf")], None)])
def test_invalid_tile_classes(mark, face):
    from o_voxel import _C

    dev = _dev()
    with pytest.raises(RuntimeError):
        _C.set_tile_classes(dev, mark, face)
    assert _C.get_tile_classes(dev) == (DEFAULT_MARK, DEFAULT_FACE)


@pytest.mark.parametrize("fpv", [-1.0, float("nan"), float("inf")])
def test_invalid_brick_class(fpv):
    from o_voxel import _C

    dev = _dev()
    with pytest.raises(RuntimeError):
        _C.set_brick_class(dev, fpv)
    assert _C.get_brick_class(dev) == DEFAULT_BRICK
