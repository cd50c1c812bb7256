"""CPU pins of the dual grid -> mesh oracle and of the reference-format
hashmap oracle (oracle/dual_mesh.py).

The Murmur3 finalizers are pinned by the algorithm's published test vectors
(MurmurHash3_x86_32 of the empty input, and the SMHasher verification codes
of MurmurHash3_x86_32 / _x64_128, which run every constant of the hash
through 256 keys).  The edge-neighbour table is pinned by geometry (the four
voxels share one grid edge, consecutive ones share a face, the flagged voxel
comes first, the winding is fixed) and by the round trip mesh -> flexible
dual grid -> mesh, where every intersected edge must find its four voxels.
The split rules are pinned by their closed forms.
"""

import numpy as np
import pytest

import meshgen
from oracle import dual_mesh as DM

M32 = DM.M32
M64 = DM.M64


# ---------------------------------------------------------------------------
# Murmur3 (published algorithm, written out here independently of the oracle
# except for the finalizers under test)
# ---------------------------------------------------------------------------
def _rotl(x, r, bits):
    m = (1 << bits) - 1
    return ((x << r) | (x >> (bits - r))) & m


def murmur3_x86_32(data: bytes, seed: int) -> bytes:
    c1, c2 = 0xCC9E2D51, 0x1B873593
    h, n = seed & M32, len(data)
    for i in range(n // 4):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = _rotl((k * c1) & M32, 15, 32) * c2 & M32
        h = (_rotl(h ^ k, 13, 32) * 5 + 0xE6546B64) & M32
    t, k = data[4 * (n // 4):], 0
    for i in range(len(t) - 1, -1, -1):
        k ^= t[i] << (8 * i)
    if t:
        h ^= _rotl((k * c1) & M32, 15, 32) * c2 & M32
    return DM.fmix32(h ^ n).to_bytes(4, "little")


def murmur3_x64_128(data: bytes, seed: int) -> bytes:
    c1, c2 = 0x87C37B91114253D5, 0x4CF5AD432745937F
    h1 = h2 = seed & M64
    n = len(data)
    for i in range(n // 16):
        k1 = int.from_bytes(data[16 * i:16 * i + 8], "little")
        k2 = int.from_bytes(data[16 * i + 8:16 * i + 16], "little")
        h1 ^= _rotl((k1 * c1) & M64, 31, 64) * c2 & M64
        h1 = ((_rotl(h1, 27, 64) + h2) * 5 + 0x52DCE729) & M64
        h2 ^= _rotl((k2 * c2) & M64, 33, 64) * c1 & M64
        h2 = ((_rotl(h2, 31, 64) + h1) * 5 + 0x38495AB5) & M64
    t = data[16 * (n // 16):]
    k1 = k2 = 0
    for i in range(len(t) - 1, 7, -1):
        k2 ^= t[i] << (8 * (i - 8))
    if len(t) > 8:
        h2 ^= _rotl((k2 * c2) & M64, 33, 64) * c1 & M64
    for i in range(min(len(t), 8) - 1, -1, -1):
        k1 ^= t[i] << (8 * i)
    if t:
        h1 ^= _rotl((k1 * c1) & M64, 31, 64) * c2 & M64
    h1 ^= n
    h2 ^= n
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    h1, h2 = DM.fmix64(h1), DM.fmix64(h2)
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    return h1.to_bytes(8, "little") + h2.to_bytes(8, "little")


def smhasher_verification(fn) -> int:
    key = bytes(range(256))
    hashes = b"".join(fn(key[:i], 256 - i) for i in range(256))
    return int.from_bytes(fn(hashes, 0)[:4], "little")


def test_fmix32_published_vectors():
    # MurmurHash3_x86_32("", seed) = fmix32(seed)
    assert DM.fmix32(0) == 0
    assert DM.fmix32(1) == 0x514E28B7
    assert DM.fmix32(0xFFFFFFFF) == 0x81F16F39
    assert smhasher_verification(murmur3_x86_32) == 0xB0F57EE3


def test_fmix64_smhasher_verification():
    assert smhasher_verification(murmur3_x64_128) == 0x6384BA69


def test_flat_key_wraps_like_size_t():
    assert DM.flat_key(0, 1, 2, 3, 10, 20, 30) == 1 * 600 + 2 * 30 + 3
    assert DM.flat_key(1, 0, 0, 0, 10, 20, 30) == 6000
    # negative coordinates sign-extend and wrap modulo 2^64
    assert DM.flat_key(0, 0, 0, -1, 10, 20, 30) == M64
    assert DM.flat_key(0, 0, 1, -1, 10, 20, 30) == 29  # aliases (0, 0, 29)


# ---------------------------------------------------------------------------
# reference-format table
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kb,vb", [(32, 32), (32, 64), (64, 32), (64, 64)])
def test_ref_hashmap_insert_lookup(kb, vb):
    rng = np.random.default_rng(kb + vb)
    keys = rng.choice(1 << 20, 300, replace=False).tolist()
    t = DM.RefHashmap(701, kb, vb)
    for i, k in enumerate(keys):
        assert t.insert(k, 7 * i + 1)
    for i, k in enumerate(keys):
        assert t.lookup(k) == 7 * i + 1
    missing = set(range(1 << 20, (1 << 20) + 50))
    for k in missing:
        assert t.lookup(k) == (1 << vb) - 1
    # every stored key is reachable from its hash without crossing an empty slot
    for slot, k in enumerate(t.keys):
        if k == t.empty:
            continue
        s = t._hash(k)
        while s != slot:
            assert t.keys[s] != t.empty
            s = (s + 1) % t.cap
    # overwrite keeps one slot
    t.insert(keys[0], 5)
    assert t.lookup(keys[0]) == 5
    assert sum(k != t.empty for k in t.keys) == len(keys)


def test_ref_hashmap_full_and_3d_bounds():
    t = DM.RefHashmap(4)
    assert all(t.insert(k, k) for k in range(4))
    assert not t.insert(99, 1)
    assert t.lookup(98) == M32
    t = DM.RefHashmap(64)
    t.insert(DM.flat_key(0, 1, 2, 3, 4, 5, 6), 11)
    assert t.lookup_3d((0, 1, 2, 3), 4, 5, 6) == 11
    for c in ((0, -1, 2, 3), (0, 4, 2, 3), (0, 1, 5, 3), (0, 1, 2, 6)):
        assert t.lookup_3d(c, 4, 5, 6) == M32


# ---------------------------------------------------------------------------
# edge-neighbour table: geometry
# ---------------------------------------------------------------------------
def test_edge_neighbors_surround_one_edge():
    for a in range(3):
        off = DM.EDGE_NEIGHBORS[a]
        assert off[0].tolist() == [0, 0, 0]  # the flagged voxel comes first
        assert (off[:, a] == 0).all()  # all four at the same position along the edge
        perp = [b for b in range(3) if b != a]
        # the four voxels are the 2x2 block around the edge at the flagged
        # voxel's (+, +) corner
        assert sorted(map(tuple, off[:, perp].tolist())) == [(0, 0), (0, 1), (1, 0), (1, 1)]
        # consecutive voxels share a face (a 4-cycle)
        for k in range(4):
            assert np.abs(off[k] - off[(k + 1) % 4]).sum() == 1
        # winding: the quad normal (right-hand rule over the voxel centres)
        # points along -axis for every axis
        n = np.cross(off[1] - off[0], off[2] - off[1])
        e = np.zeros(3, np.int64)
        e[a] = -1
        assert n.tolist() == e.tolist()


def test_mesh_vertices_closed_form():
    rng = np.random.default_rng(0)
    G = 64
    c = rng.integers(0, G, (500, 3))
    d = rng.integers(0, 257, (500, 3)) / 256.0
    v = DM.mesh_vertices(c, d.astype(np.float32), [1.0 / G] * 3, [-0.5] * 3)
    np.testing.assert_array_equal(v, ((c + d) / G - 0.5).astype(np.float32))


def test_round_trip_every_flag_finds_its_voxels(oracle):
    """mesh -> flexible dual grid -> mesh: every intersected edge of a closed
    mesh has its four voxels occupied, so every flag yields one quad, and the
    reconstructed vertices stay within a voxel of the surface."""
    v, f = meshgen.icosphere(3)
    v = (v * 0.37 + np.array([0.013, -0.021, 0.007])).astype(np.float32)
    G = 32
    vs = [1.0 / G] * 3
    r = oracle.flexible_dual_grid(v + 0.5, f.astype(np.int32), vs, [0, 0, 0, G, G, G], 1.0, 0.2, 1e-2)
    dual_vox = r.dual * G - r.coords  # voxel units, as the reference example packs them
    quads = DM.quads_of(r.coords, r.flags, (G, G, G))
    assert len(quads) == int(r.flags.sum()) > 0
    verts, faces = DM.dual_grid_to_mesh(r.coords, dual_vox, r.flags, None, (G, G, G), vs, [-0.5] * 3)
    assert faces.shape == (2 * len(quads), 3)
    rad = np.linalg.norm(verts - np.array([0.013, -0.021, 0.007]), axis=1)
    assert np.abs(rad - 0.37).max() < 1.0 / G
    # every quad edge joins face-adjacent voxels
    c = r.coords.astype(np.int64)
    for k in range(4):
        assert (np.abs(c[quads[:, k]] - c[quads[:, (k + 1) % 4]]).sum(1) == 1).all()


def test_quads_skip_missing_and_out_of_grid_voxels():
    # a 2x2x2 block with the corner voxel flagged on all axes -> 3 quads; drop
    # one voxel -> the quads that need it disappear; a flagged voxel on the
    # grid's upper face has no neighbours
    c = np.array([[x, y, z] for z in range(2) for y in range(2) for x in range(2)])
    flags = np.zeros((8, 3), bool)
    flags[0] = True
    q = DM.quads_of(c, flags, (4, 4, 4))
    assert len(q) == 3 and (q[:, 0] == 0).all()
    keep = [i for i in range(8) if tuple(c[i]) != (0, 1, 1)]  # the x-axis quad's diagonal voxel
    q2 = DM.quads_of(c[keep], flags[keep], (4, 4, 4))
    assert len(q2) == 2
    q3 = DM.quads_of(c, flags, (1, 4, 4))  # x = 1 lies outside a grid 1 wide
    assert len(q3) == 1  # only the x-axis quad (all four voxels at x = 0)


def test_split_rules_closed_form():
    rng = np.random.default_rng(1)
    L = 200
    verts = rng.normal(size=(4 * L, 3)).astype(np.float32)
    quads = np.arange(4 * L).reshape(L, 4)
    # without weights the second alignment is |n . cross(v3 - v1, v3 - v1)| = 0,
    # so split 1 wins exactly when triangle (v0, v1, v2) is not degenerate
    assert DM.split_first(verts, quads).all()
    deg = verts.copy()
    deg[quads[:, 1]] = deg[quads[:, 0]]
    assert not DM.split_first(deg, quads).any()
    # with weights: w0*w2 > w1*w3, ties to split 2
    w = np.array([2, 1, 2, 1, 1, 2, 1, 2, 1, 1, 1, 1], np.float32)
    q = np.arange(12).reshape(3, 4)
    assert DM.split_first(None, q, w).tolist() == [True, False, False]
    # products round in the weights' dtype: 2049*1 vs 2048*1 differ in float32,
    # tie in float16 (11-bit significand) -> split 2 there
    w = np.array([2049, 1, 1, 2048], np.float32)
    q = np.arange(4).reshape(1, 4)
    assert DM.split_first(None, q, w, "float32").tolist() == [True]
    assert DM.split_first(None, q, w, "float16").tolist() == [False]
    w = np.array([257, 1, 1, 256], np.float32)  # bfloat16: 8-bit significand
    assert DM.split_first(None, q, w, "bfloat16").tolist() == [False]


def test_faces_follow_split_patterns():
    c = np.array([[x, y, z] for z in range(2) for y in range(2) for x in range(2)])
    flags = np.zeros((8, 3), bool)
    flags[0] = True
    dual = np.full((8, 3), 0.5, np.float32)
    verts, faces = DM.dual_grid_to_mesh(c, dual, flags, None, (4, 4, 4), [0.25] * 3, [0.0] * 3)
    q = DM.quads_of(c, flags, (4, 4, 4))
    assert faces.dtype == np.int32 and faces.shape == (6, 3)
    for j, qq in enumerate(q):
        # split 1 (non-degenerate quads): the two triangles share the 0-2 diagonal
        assert faces[2 * j].tolist() == [qq[0], qq[1], qq[2]]
        assert faces[2 * j + 1].tolist() == [qq[0], qq[2], qq[3]]


def test_train_mode_closed_form():
    c = np.array([[x, y, z] for z in range(2) for y in range(2) for x in range(2)])
    flags = np.zeros((8, 3), bool)
    flags[0, 2] = True
    rng = np.random.default_rng(3)
    dual = rng.random((8, 3)).astype(np.float32)
    N = 8
    w = np.ones((N, 1), np.float32)
    verts, faces = DM.dual_grid_to_mesh(c, dual, flags, w, (4, 4, 4), [0.25] * 3, [0.0] * 3, train=True)
    q = DM.quads_of(c, flags, (4, 4, 4))[0]
    assert verts.shape == (N + 1, 3) and faces.shape == (4, 3) and faces.dtype == np.int64
    # equal weights: the mid vertex is the quad's centroid
    np.testing.assert_allclose(verts[N], verts[q].mean(0), rtol=0, atol=1e-6)
    # fan: every triangle is one quad edge (in cyclic order) plus the mid vertex
    for k in range(4):
        assert faces[k].tolist() == [q[k], q[(k + 1) % 4], N]
    # zero weight on the 1-3 diagonal: the mid vertex is the 0-2 midpoint
    w2 = np.ones((N, 1), np.float32)
    w2[q[1]] = 0
    verts2, _ = DM.dual_grid_to_mesh(c, dual, flags, w2, (4, 4, 4), [0.25] * 3, [0.0] * 3, train=True)
    np.testing.assert_allclose(verts2[N], (verts2[q[0]] + verts2[q[2]]) / 2, rtol=0, atol=1e-7)
