"""GPU: graph mode (ovx_fdg_set_graph).  The device work of a conversion
after its voxel set is captured as a HIP graph and replayed for conversions of
the same plan; every replay must equal the eager conversion bit for bit, with
the program's own eager work and synchronizations between the replays, for a
batch and for a single mesh, and a conversion whose record buffers overflow
inside the graph is re-run eagerly with exact sizes."""

import numpy as np
import pytest
import torch

import meshgen

pytestmark = pytest.mark.gpu

AABB = [[-0.5, -0.5, -0.5], [0.5, 0.5, 0.5]]
KW = dict(face_weight=1.0, boundary_weight=0.2, regularization_weight=1e-2)


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda", 0)


def _bytes(outs):
    return [[t.cpu().view(torch.uint8) if t.dtype != torch.bool else t.cpu() for t in o] for o in outs]


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        for u, w in zip(x, y):
            assert u.shape == w.shape and torch.equal(u, w)


def _eager_work(dev, k):
    """Unrelated eager work of the program between replays, then a sync."""
    x = torch.arange(1 << 20, device=dev, dtype=torch.float32) * (k + 1)
    y = (x * 0.5 + 1.0).sum()
    torch.cuda.synchronize()
    return float(y)


@pytest.mark.parametrize("precision", ["reference", "compensated"])
def test_graph_replays_equal_eager_batch(precision):
    import o_voxel
    from o_voxel import _C

    dev = _dev()
    meshes = meshgen.batch(n_meshes=16, seed=9, f_lo=300, f_hi=40000, grids=(32, 64, 128, 256))
    ms = [(torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev)) for v, f, _ in meshes]
    grids = [g for _, _, g in meshes]

    def conv():
        return o_voxel.convert.mesh_to_flexible_dual_grid_batch(ms, grid_sizes=grids, aabb=AABB, precision=precision,
                                                               **KW)

    print(f"[case] batch of 16 in graph mode, {precision}", flush=True)
    _C.set_graph(dev, False)
    want = _bytes(conv())
    s0 = _C.graph_stats(dev)
    _C.set_graph(dev, True)
    try:
        for k in range(4):
            out = conv()
            _same(_bytes(out), want)
            del out
            _eager_work(dev, k)
        st = _C.graph_stats(dev)
        assert st["captures"] - s0["captures"] >= 1
        assert st["replays"] - s0["replays"] >= 1, st
    finally:
        _C.set_graph(dev, False)


def test_graph_single_mesh_and_other_plans():
    """Single conversions alternating between two meshes (two plans, two
    graphs kept) replay the right graph each time."""
    import o_voxel
    from o_voxel import _C

    dev = _dev()
    inputs = []
    for seed, nf, g in ((51, 20000, 128), (52, 6000, 64)):
        v, f = meshgen.scene(seed=seed, target_faces=nf, n_closed=6, n_patches=6)
        inputs.append((torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev), g))

    def conv(i):
        v, f, g = inputs[i]
        return o_voxel.convert.mesh_to_flexible_dual_grid(v, f, grid_size=g, aabb=AABB, **KW)

    print("[case] two meshes alternating in graph mode", flush=True)
    _C.set_graph(dev, False)
    want = [_bytes([conv(i)])[0] for i in range(2)]
    _C.set_graph(dev, True)
    try:
        for k in range(6):
            out = conv(k % 2)
            _same(_bytes([out]), [want[k % 2]])
            del out
            _eager_work(dev, k)
    finally:
        _C.set_graph(dev, False)


def test_graph_overflow_reruns_eagerly():
    """Record buffers far too small: the graph's overflow flag makes the fold
    and solve return at once, and finish re-runs the record stages eagerly
    with exact sizes; results equal the eager ones."""
    import o_voxel
    from o_voxel import _C

    dev = _dev()
    v, f = meshgen.scene(seed=53, target_faces=15000, n_closed=6, n_patches=6)
    vt, ft = torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev)

    def conv():
        return o_voxel.convert.mesh_to_flexible_dual_grid(vt, ft, grid_size=128, aabb=AABB, **KW)

    _C.set_graph(dev, False)
    want = _bytes([conv()])
    _C.set_graph(dev, True)
    try:
        print("[case] graph mode with record buffers of 4096 items", flush=True)
        _C.set_capacity(dev, events=0, hits=4096)
        _same(_bytes([conv()]), want)
        _same(_bytes([conv()]), want)
    finally:
        _C.set_graph(dev, False)
