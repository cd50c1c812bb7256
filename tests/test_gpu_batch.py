"""GPU: batched conversion (many meshes of mixed size and grid in one pass)
equals converting each mesh alone, bit for bit; and the batch of one is the
plain path.  The oracle checks one batch member end to end."""

import numpy as np
import pytest
import torch

import meshgen

pytestmark = pytest.mark.gpu

AABB = [[-0.5, -0.5, -0.5], [0.5, 0.5, 0.5]]


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda", 0)


def _meshes(n, seed, f_hi=20000):
    out = meshgen.batch(n_meshes=n, seed=seed, f_lo=200, f_hi=f_hi, grids=(32, 64, 128, 256))
    return [(v, f.astype(np.int64), g) for v, f, g in out]


def _single(v, f, g, dev, **kw):
    import o_voxel

    vt = torch.from_numpy(v).to(dev)
    ft = torch.from_numpy(f).to(dev)
    return [t.cpu().numpy() for t in o_voxel.convert.mesh_to_flexible_dual_grid(
        vt, ft, grid_size=g, aabb=AABB, face_weight=1.0, boundary_weight=0.2, regularization_weight=1e-2, **kw)]


def _batch(meshes, dev, **kw):
    import o_voxel

    ms = [(torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev)) for v, f, _ in meshes]
    out = o_voxel.convert.mesh_to_flexible_dual_grid_batch(
        ms, grid_sizes=[g for _, _, g in meshes], aabb=AABB, face_weight=1.0, boundary_weight=0.2,
        regularization_weight=1e-2, **kw)
    torch.cuda.synchronize()
    return [[t.cpu().numpy() for t in o] for o in out]


@pytest.mark.parametrize("precision", ["reference", "compensated"])
def test_batch_equals_single_calls(precision):
    dev = _dev()
    meshes = _meshes(12, seed=7)
    # an empty mesh and a mesh entirely outside its grid inside the batch
    meshes.insert(3, (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int64), 64))
    v, f = meshgen.icosphere(2)
    meshes.insert(7, ((v * 0.1 + 2.0).astype(np.float32), f.astype(np.int64), 64))
    print(f"[case] batch of {len(meshes)} meshes, precision={precision}, "
          f"faces={[len(f) for _, f, _ in meshes]}", flush=True)
    got = _batch(meshes, dev, precision=precision)
    assert len(got) == len(meshes)
    for k, (v, f, g) in enumerate(meshes):
        want = _single(v, f, g, dev, precision=precision)
        for a, b, name in zip(got[k], want, ("coords", "dual", "intersected")):
            assert a.shape == b.shape and np.array_equal(a, b), f"mesh {k} ({len(f)} faces, grid {g}): {name} differs"


def test_batch_member_matches_oracle(oracle):
    dev = _dev()
    meshes = _meshes(6, seed=11, f_hi=8000)
    got = _batch(meshes, dev)
    for k in (0, len(meshes) - 1):
        v, f, g = meshes[k]
        vs, gr, a0 = meshgen.grid_args(g)
        ref = oracle.flexible_dual_grid((v - a0).astype(np.float32), f.astype(np.int32), vs, gr, 1.0, 0.2, 1e-2)
        perm = oracle.canonical_order(ref.coords, gr)
        c, d, i = got[k]
        assert np.array_equal(c, ref.coords[perm]) and np.array_equal(i, ref.flags[perm])
        assert np.array_equal(d, ref.dual[perm])


def test_batch_of_one_is_plain_path():
    dev = _dev()
    (v, f, g), = _meshes(1, seed=3)
    got = _batch([(v, f, g)], dev)[0]
    want = _single(v, f, g, dev)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_batch_prepare_matches_torch():
    """The batch wrapper's packing kernel gives exactly torch's concatenation
    of `vertices - origin` and of the faces rebased by their mesh's first
    vertex row, over meshes with int32 and int64 faces, an empty mesh, a mesh
    without faces and meshes spanning several kernel blocks; an index outside
    its own mesh becomes -1."""
    from o_voxel import _C

    dev = _dev()
    rng = np.random.default_rng(17)
    sizes = [(5000, 9000), (0, 0), (3, 1), (1777, 0), (20011, 31337), (1, 1), (4096, 1366)]
    vs, fs = [], []
    for k, (nv, nf) in enumerate(sizes):
        vs.append(torch.from_numpy(rng.normal(0.0, 2.0, (nv, 3)).astype(np.float32)).to(dev))
        f = rng.integers(0, max(nv, 1), (nf, 3))
        fs.append(torch.from_numpy(f.astype(np.int64 if k % 2 else np.int32)).to(dev))
    org = torch.from_numpy(rng.normal(0.0, 1.0, (len(sizes), 3)).astype(np.float32)).to(dev)
    print("[case] batch prepare: 7 meshes", flush=True)
    v_all, f_all, nv, nf = _C._batch_prepare(vs, fs, org)
    torch.cuda.synchronize()
    want_v = torch.cat([v - org[k].reshape(1, 3) for k, v in enumerate(vs)], 0)
    base = np.cumsum([0] + nv[:-1])
    want_f = torch.cat([(f.to(torch.int64) + int(base[k])).to(torch.int32) for k, f in enumerate(fs)], 0)
    assert nv == [s[0] for s in sizes] and nf == [s[1] for s in sizes]
    assert torch.equal(v_all, want_v) and torch.equal(f_all, want_f)
    print("[case] batch prepare: indices outside their mesh", flush=True)
    fs[4] = fs[4].clone()
    fs[4][7, 2] = 20011          # first vertex of the next mesh: outside this one
    fs[4][9, 0] = -1
    fs[5] = fs[5].clone()        # int64 faces
    fs[5][0, 1] = 1 << 40
    _, f_bad, _, _ = _C._batch_prepare(vs, fs, org)
    r4 = int(sum(nf[:4]))
    r5 = int(sum(nf[:5]))
    assert f_bad[r4 + 7, 2].item() == -1 and f_bad[r4 + 9, 0].item() == -1 and f_bad[r5, 1].item() == -1
    print("[case] batch prepare: no origins", flush=True)
    v0, _, _, _ = _C._batch_prepare(vs, fs, None)
    assert torch.equal(v0, torch.cat(vs, 0))


def test_batch_invalid_face_index_raises():
    """An index that names a vertex of another mesh of the batch is an input error."""
    dev = _dev()
    meshes = _meshes(3, seed=5, f_hi=2000)
    v, f, g = meshes[1]
    f = f.copy()
    f[0, 0] = len(v)  # one past this mesh's vertices: a vertex of the next mesh in the batch
    meshes[1] = (v, f, g)
    with pytest.raises(RuntimeError, match="vertex index"):
        _batch(meshes, dev)
    _batch(_meshes(2, seed=6, f_hi=2000), dev)  # the context keeps working


def test_batch_with_fan_vertex_equals_single_calls():
    """A mesh whose centre vertex has 80 spokes (more half-edges than the
    per-vertex slots of the boundary-edge stage) makes the batch take the
    general per-vertex lists; every mesh's outputs still equal its single call."""
    dev = _dev()
    meshes = _meshes(5, seed=21)
    n = 80
    ang = np.linspace(0.0, 1.8 * np.pi, n + 1)
    ring = np.stack([0.3 * np.cos(ang), 0.3 * np.sin(ang), 0.05 * np.sin(3 * ang)], 1)
    v = np.concatenate([np.zeros((1, 3)), ring]).astype(np.float32)
    f = np.stack([np.zeros(n, np.int64), np.arange(1, n + 1), np.arange(2, n + 2)], 1)
    meshes.insert(2, (v, f, 64))
    print(f"[case] batch of {len(meshes)} meshes with an {n}-spoke fan", flush=True)
    got = _batch(meshes, dev)
    for k, (v, f, g) in enumerate(meshes):
        want = _single(v, f, g, dev)
        for a, b in zip(got[k], want):
            assert a.shape == b.shape
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), k
