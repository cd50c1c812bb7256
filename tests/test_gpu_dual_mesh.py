"""GPU parity of flexible_dual_grid_to_mesh and of the generic hashmap
(o_voxel._C.hashmap_*) against the numpy oracle (oracle/dual_mesh.py).

Bar: bit-exact vertices, faces and quads (integer and index work, float32
vertex arithmetic with one rounding per operation); train-mode mid vertices
within 1e-6 (torch's elementwise ops on the same values); hashmap lookups
exact, tables readable by the oracle's reference-format probe.
"""

import sys

import numpy as np
import pytest
import torch

import meshgen
from oracle import dual_mesh as DM

pytestmark = pytest.mark.gpu

DEV = "cuda"
M32 = DM.M32


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ov():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import o_voxel

    return o_voxel


# ---------------------------------------------------------------------------
# generic hashmap
# ---------------------------------------------------------------------------
TDT = {32: torch.uint32, 64: torch.uint64}


def _u(a: np.ndarray, bits: int) -> torch.Tensor:
    """numpy unsigned ints -> device uint tensor (through int64 bit patterns)."""
    a = np.asarray(a, dtype=np.uint64 if bits == 64 else np.uint32)
    return torch.from_numpy(a.view(np.int64 if bits == 64 else np.int32).copy()).view(TDT[bits]).to(DEV)


def _np(t: torch.Tensor, bits: int) -> np.ndarray:
    return t.cpu().view(torch.int64 if bits == 64 else torch.int32).numpy().view(
        np.uint64 if bits == 64 else np.uint32)


@pytest.mark.parametrize("kb,vb", [(32, 32), (32, 64), (64, 32), (64, 64)])
def test_hashmap_insert_lookup(ov, kb, vb):
    _say(f"hashmap kb={kb} vb={vb}")
    _C = ov._C
    rng = np.random.default_rng(kb * 3 + vb)
    M, cap = 20000, 40009
    hi = (1 << 40) if kb == 64 else (1 << 31)
    keys = np.unique(rng.integers(0, hi, 2 * M))[:M]
    rng.shuffle(keys)
    vals = rng.integers(0, 1 << (vb - 2), M)
    tk = _u(np.full(cap, (1 << kb) - 1, dtype=np.uint64), kb)
    tv = torch.zeros((cap,), dtype=TDT[vb], device=DEV)
    _C.hashmap_insert_cuda(tk, tv, _u(keys, kb), _u(vals, vb))
    absent = np.arange(hi, hi + 1000)
    q = np.concatenate([keys, absent])
    out = _np(_C.hashmap_lookup_cuda(tk, tv, _u(q, kb)), vb)
    np.testing.assert_array_equal(out[:M], vals.astype(out.dtype))
    assert (out[M:] == (1 << vb) - 1).all()
    # the table is the reference's format: the oracle's sequential probe reads it
    ref = DM.RefHashmap(cap, kb, vb, keys=_np(tk, kb).tolist(), values=_np(tv, vb).tolist())
    for i in rng.choice(M, 300, replace=False):
        assert ref.lookup(int(keys[i])) == int(vals[i])
    for k in absent[:50]:
        assert ref.lookup(int(k)) == (1 << vb) - 1
    stored = _np(tk, kb)
    assert np.sort(stored[stored != (1 << kb) - 1]).tolist() == np.sort(keys).tolist()
    # re-inserting a key overwrites its value in place
    _C.hashmap_insert_cuda(tk, tv, _u(keys[:10], kb), _u(np.arange(10), vb))
    out = _np(_C.hashmap_lookup_cuda(tk, tv, _u(keys[:10], kb)), vb)
    assert out.tolist() == list(range(10))
    assert (_np(tk, kb) != (1 << kb) - 1).sum() == M


def test_hashmap_3d(ov):
    _say("hashmap 3d")
    _C = ov._C
    rng = np.random.default_rng(7)
    W, H, D = 37, 41, 43
    n = 5000
    flat = rng.choice(W * H * D, n, replace=False)
    c = np.stack([np.zeros(n, np.int64), flat // (H * D), (flat // D) % H, flat % D], 1).astype(np.int32)
    coords = torch.from_numpy(c).to(DEV)
    cap = 2 * n
    for kb in (32, 64):
        tk = _u(np.full(cap, (1 << kb) - 1, dtype=np.uint64), kb)
        tv = torch.zeros((cap,), dtype=torch.uint32, device=DEV)
        _C.hashmap_insert_3d_idx_as_val_cuda(tk, tv, coords, W, H, D)
        probe = np.concatenate([c, np.array([[0, -1, 0, 0], [0, W, 0, 0], [0, 0, H, 0], [0, 0, 0, D],
                                             [0, 1, 2, 3]], np.int32)])
        out = _np(_C.hashmap_lookup_3d_cuda(tk, tv, torch.from_numpy(probe).to(DEV), W, H, D), 32)
        assert out[:n].tolist() == list(range(n))
        assert (out[n:n + 4] == M32).all()
        want = np.nonzero(flat == 1 * H * D + 2 * D + 3)[0]
        assert out[n + 4] == (int(want[0]) if len(want) else M32)
        # explicit values
        tk = _u(np.full(cap, (1 << kb) - 1, dtype=np.uint64), kb)
        tv = torch.zeros((cap,), dtype=torch.uint64, device=DEV)
        vals = rng.integers(0, 1 << 62, n)
        _C.hashmap_insert_3d_cuda(tk, tv, coords, _u(vals, 64), W, H, D)
        out = _np(_C.hashmap_lookup_3d_cuda(tk, tv, coords, W, H, D), 64)
        np.testing.assert_array_equal(out, vals.astype(np.uint64))
        ref = DM.RefHashmap(cap, kb, 64, keys=_np(tk, kb).tolist(), values=_np(tv, 64).tolist())
        for i in range(0, n, 97):
            assert ref.lookup_3d(c[i], W, H, D) == int(vals[i])


def test_hashmap_type_errors(ov):
    _C = ov._C
    tk = _u(np.full(16, M32, dtype=np.uint64), 32)
    tv = torch.zeros((16,), dtype=torch.uint32, device=DEV)
    with pytest.raises(RuntimeError, match="Keys must be uint32"):
        _C.hashmap_insert_cuda(tk, tv, torch.zeros(3, dtype=torch.int64, device=DEV), tv[:3])
    with pytest.raises(RuntimeError, match="Values must be uint32"):
        _C.hashmap_insert_cuda(tk, tv, tk[:3], torch.zeros(3, dtype=torch.int64, device=DEV))
    with pytest.raises(RuntimeError, match="Coords must be int32"):
        _C.hashmap_insert_3d_idx_as_val_cuda(tk, tv, torch.zeros((3, 4), dtype=torch.int64, device=DEV), 4, 4, 4)
    with pytest.raises(RuntimeError, match="Unsupported data type"):
        _C.hashmap_lookup_cuda(torch.zeros(16, dtype=torch.int32, device=DEV), tv, tk[:3])


def test_hashmap_full_table_terminates(ov):
    """More distinct keys than slots: the extra keys are dropped, lookups of
    them report missing, and both kernels return."""
    _C = ov._C
    tk = _u(np.full(8, M32, dtype=np.uint64), 32)
    tv = torch.zeros((8,), dtype=torch.uint32, device=DEV)
    keys = np.arange(100, 112)
    _C.hashmap_insert_cuda(tk, tv, _u(keys, 32), _u(keys, 32))
    out = _np(_C.hashmap_lookup_cuda(tk, tv, _u(np.arange(100, 120), 32)), 32)
    torch.cuda.synchronize()
    found = out[:12]
    assert (found != M32).sum() == 8
    assert ((found == keys) | (found == M32)).all()
    assert (out[12:] == M32).all()


# ---------------------------------------------------------------------------
# flexible_dual_grid_to_mesh
# ---------------------------------------------------------------------------
def _fdg_input(ov, G, seed=0, faces=60000):
    v, f = meshgen.scene(seed=seed, target_faces=faces, n_closed=12, n_patches=8)
    coords, dual, inter = ov.convert.mesh_to_flexible_dual_grid(
        torch.from_numpy(v).to(DEV), torch.from_numpy(f).to(DEV), grid_size=G, aabb=[[-0.5] * 3, [0.5] * 3],
        face_weight=1.0, boundary_weight=0.2, regularization_weight=1e-2)
    dual_vox = dual * G - coords  # voxel units, as the reference's example packs them
    return coords, dual_vox, inter


def _check_mesh(ov, coords, dual, inter, G, split_weight=None, wdt="float32", perm=None):
    verts, faces = ov.convert.flexible_dual_grid_to_mesh(coords, dual, inter, split_weight,
                                                         aabb=[[-0.5] * 3, [0.5] * 3], grid_size=G)
    sw = None if split_weight is None else split_weight.float().cpu().numpy()
    rv, rf = DM.dual_grid_to_mesh(coords.cpu().numpy(), dual.float().cpu().numpy(), inter.cpu().numpy(), sw,
                                  (G, G, G), [np.float32(1.0) / np.float32(G)] * 3, [-0.5] * 3, weight_dtype=wdt)
    assert verts.dtype == torch.float32 and faces.dtype == torch.int32
    np.testing.assert_array_equal(verts.cpu().numpy(), rv)
    np.testing.assert_array_equal(faces.cpu().numpy(), rf)
    return verts, faces


@pytest.mark.parametrize("G", [64, 128])
def test_to_mesh_matches_oracle(ov, G):
    _say(f"to_mesh G={G}")
    coords, dual, inter = _fdg_input(ov, G, seed=G)
    verts, faces = _check_mesh(ov, coords, dual, inter, G)
    assert faces.shape[0] == 2 * int(inter.sum())  # every flag of a conversion finds its voxels


def test_to_mesh_split_weights(ov):
    _say("to_mesh split weights")
    G = 64
    coords, dual, inter = _fdg_input(ov, G, seed=11)
    N = coords.shape[0]
    g = torch.Generator(device="cpu").manual_seed(5)
    w = torch.rand((N, 1), generator=g)
    # quantised weights make ties and near-ties common
    w = (torch.floor(w * 64) / 64 + 1 / 128).to(DEV)
    _check_mesh(ov, coords, dual, inter, G, split_weight=w)
    _check_mesh(ov, coords, dual, inter, G, split_weight=(w * 200).half(), wdt="float16")
    _check_mesh(ov, coords, dual, inter, G, split_weight=(w * 300).bfloat16(), wdt="bfloat16")
    _check_mesh(ov, coords, dual, inter, G, split_weight=w.reshape(N))


def test_to_mesh_unsorted_sparse_and_half(ov):
    _say("to_mesh permuted / dropped / half")
    G = 64
    coords, dual, inter = _fdg_input(ov, G, seed=13)
    N = coords.shape[0]
    g = torch.Generator(device="cpu").manual_seed(9)
    perm = torch.randperm(N, generator=g).to(DEV)
    _check_mesh(ov, coords[perm], dual[perm], inter[perm], G)
    keep = torch.rand(N, generator=g).to(DEV) > 0.05  # missing voxels drop quads
    _check_mesh(ov, coords[keep], dual[keep], inter[keep], G)
    _check_mesh(ov, coords, dual.half(), inter, G)
    # voxel_size instead of grid_size gives the same grid
    v1, f1 = ov.convert.flexible_dual_grid_to_mesh(coords, dual, inter, None, aabb=[[-0.5] * 3, [0.5] * 3],
                                                   voxel_size=1.0 / G)
    v2, f2 = ov.convert.flexible_dual_grid_to_mesh(coords, dual, inter, None, aabb=[[-0.5] * 3, [0.5] * 3],
                                                   grid_size=G)
    assert torch.equal(v1, v2) and torch.equal(f1, f2)


def test_to_mesh_train_mode(ov):
    _say("to_mesh train")
    G = 64
    coords, dual, inter = _fdg_input(ov, G, seed=17, faces=20000)
    N = coords.shape[0]
    g = torch.Generator(device="cpu").manual_seed(2)
    w = (torch.rand((N, 1), generator=g) + 0.1).to(DEV)
    d = dual.clone().requires_grad_(True)
    ww = w.clone().requires_grad_(True)
    verts, faces = ov.convert.flexible_dual_grid_to_mesh(coords, d, inter, ww, aabb=[[-0.5] * 3, [0.5] * 3],
                                                         grid_size=G, train=True)
    rv, rf = DM.dual_grid_to_mesh(coords.cpu().numpy(), dual.cpu().numpy(), inter.cpu().numpy(),
                                  w.cpu().numpy(), (G, G, G), [np.float32(1.0 / G)] * 3, [-0.5] * 3, train=True)
    assert faces.dtype == torch.int64
    np.testing.assert_array_equal(faces.cpu().numpy(), rf)
    np.testing.assert_array_equal(verts[:N].detach().cpu().numpy(), rv[:N])
    np.testing.assert_allclose(verts[N:].detach().cpu().numpy(), rv[N:], rtol=0, atol=1e-6)
    (verts ** 2).sum().backward()
    assert d.grad is not None and torch.isfinite(d.grad).all()
    assert ww.grad is not None and torch.isfinite(ww.grad).all()


def test_to_mesh_matches_reference_composition(ov):
    """The quads of the fused path equal those of the reference's own recipe
    run through this build's generic hashmap (insert rows, look up the four
    edge neighbours of every flag, keep rows with no miss)."""
    _say("to_mesh vs hashmap composition")
    _C = ov._C
    G = 64
    coords, dual, inter = _fdg_input(ov, G, seed=19)
    N = coords.shape[0]
    tk = torch.full((2 * N,), -1, dtype=torch.int64, device=DEV).to(torch.int32).view(torch.uint32)
    tv = torch.empty((2 * N,), dtype=torch.uint32, device=DEV)
    _C.hashmap_insert_3d_idx_as_val_cuda(tk, tv, torch.cat([torch.zeros_like(coords[:, :1]), coords], 1), G, G, G)
    off = torch.from_numpy(DM.EDGE_NEIGHBORS.astype(np.int32)).to(DEV).unsqueeze(0)
    cand = (coords.reshape(N, 1, 1, 3) + off)[inter]
    M = cand.shape[0]
    key = torch.cat([torch.zeros((M * 4, 1), dtype=torch.int32, device=DEV), cand.reshape(-1, 3)], 1)
    idx = _C.hashmap_lookup_3d_cuda(tk, tv, key, G, G, G).reshape(M, 4).int()
    quads_ref = idx[(idx != -1).all(1)]
    _, _, quads = _C.flexible_dual_grid_to_mesh_cuda(coords, dual, inter, None, [G] * 3, [1.0 / G] * 3,
                                                     [-0.5] * 3, faces=False, quads=True)
    assert torch.equal(quads, quads_ref)


def test_to_mesh_empty_and_host(ov):
    v, f = ov.convert.flexible_dual_grid_to_mesh(
        torch.zeros((0, 3), dtype=torch.int32, device=DEV), torch.zeros((0, 3), device=DEV),
        torch.zeros((0, 3), dtype=torch.bool, device=DEV), None, aabb=[[-0.5] * 3, [0.5] * 3], grid_size=16)
    assert v.shape == (0, 3) and f.shape == (0, 3)
    c = torch.tensor([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], dtype=torch.int32)
    inter = torch.zeros((4, 3), dtype=torch.bool)
    inter[0, 2] = True
    d = torch.full((4, 3), 0.5)
    v, f = ov.convert.flexible_dual_grid_to_mesh(c, d, inter, None, aabb=[[0] * 3, [1] * 3], grid_size=4)
    assert not v.is_cuda and f.tolist() == [[0, 2, 3], [0, 3, 1]]


def _oracle_mesh(coords, dual, inter, G):
    return DM.dual_grid_to_mesh(coords.cpu().numpy(), dual.cpu().numpy(), inter.cpu().numpy(), None, G,
                                [np.float32(1.0) / np.float32(g) for g in G], [-0.5] * 3)


@pytest.mark.parametrize("G", [(64, 64, 64), (1100, 1100, 1100), (2048, 2048, 2048), (40, 3000, 9000)])
def test_to_mesh_table_modes(ov, G):
    """Direct-mapped table (grids up to 2^30 cells), packed 8-byte hash
    (below 2^32 cells) and 16-byte hash (2^32 cells and more), with
    duplicate rows and coordinates outside the grid whose keys alias cells
    inside it."""
    _say(f"to_mesh table mode grid={G}")
    coords, dual, inter = _fdg_input(ov, 64, seed=23, faces=20000)
    N = coords.shape[0]
    g = torch.Generator(device="cpu").manual_seed(4)
    dup = torch.randint(0, N, (200,), generator=g).to(DEV)
    W, H, D = G
    odd = torch.tensor([[0, 1, -1], [0, 0, D], [-1, 0, 0], [W, 0, 0], [2, -3, 5]], dtype=torch.int32, device=DEV)
    c = torch.cat([coords, coords[dup], odd])
    d = torch.cat([dual, dual[dup].flip(1), torch.full((5, 3), 0.25, device=DEV)])
    f = torch.cat([inter, inter[dup], torch.ones((5, 3), dtype=torch.bool, device=DEV)])
    verts, faces = ov.convert.flexible_dual_grid_to_mesh(c, d, f, None, aabb=[[-0.5] * 3, [0.5] * 3], grid_size=list(G))
    rv, rf = _oracle_mesh(c, d, f, G)
    np.testing.assert_array_equal(verts.cpu().numpy(), rv)
    np.testing.assert_array_equal(faces.cpu().numpy(), rf)


def test_to_mesh_begin_without_finish(ov):
    """A begin whose finish never ran leaves the next call correct (the
    direct table's cells of the abandoned call are cleared first)."""
    import ctypes

    _C = ov._C
    L = _C._load()
    coords, dual, inter = _fdg_input(ov, 64, seed=29, faces=20000)
    dev = coords.device
    ctx = _C._ctx(dev)
    gs = (ctypes.c_int32 * 3)(64, 64, 64)
    vs = (ctypes.c_float * 3)(*([1.0 / 64] * 3))
    a0 = (ctypes.c_float * 3)(*([-0.5] * 3))
    verts = torch.empty_like(dual)
    flags = inter.contiguous().view(torch.uint8)
    nq = ctypes.c_int64(0)
    with torch.cuda.device(dev), ctx.lock:
        st = L.ovx_dual_mesh_begin(ctx.h, coords.data_ptr(), dual.data_ptr(), flags.data_ptr(), coords.shape[0], gs,
                                   vs, a0, verts.data_ptr(), _C._stream_handle(dev), ctypes.byref(nq))
    assert st == 0 and nq.value > 0
    # a different voxel set on the same grid
    keep = torch.arange(coords.shape[0], device=DEV) % 3 != 0
    _check_mesh(ov, coords[keep], dual[keep], inter[keep], 64)
    _check_mesh(ov, coords, dual, inter, 64)
