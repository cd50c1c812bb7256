"""GPU parity: the HIP path against the CPU oracle on the same inputs.

Bar (per the workload definition): occupancy and intersected flags
bit-exact; dual vertices RMS <= 2e-4 voxel and max <= 1e-3 voxel against the
oracle float path after canonical voxel ordering.  The kernels replay the
oracle's operation order, so the duals are expected to be bit-identical; the
tests also report the bit-identical fraction.
"""

import numpy as np
import pytest
import torch

import meshgen

pytestmark = pytest.mark.gpu

DUAL_RMS_TOL = 2e-4  # voxel units
DUAL_MAX_TOL = 1e-3  # voxel units


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda", 0)


def _C():
    from o_voxel import _C

    return _C


def run_gpu(v, f, vs, gr, fw, bw, rw):
    dev = _dev()
    vt = torch.from_numpy(np.ascontiguousarray(v, np.float32)).to(dev)
    ft = torch.from_numpy(np.ascontiguousarray(f, np.int32)).to(dev)
    c, d, i = _C().mesh_to_flexible_dual_grid_cuda(vt, ft, vs, gr, fw, bw, rw)
    torch.cuda.synchronize()
    return c.cpu().numpy(), d.cpu().numpy(), i.cpu().numpy()


def compare(oracle, v, f, vs, gr, fw=1.0, bw=0.2, rw=1e-2, label=""):
    print(f"[case] {label} F={len(f)} grid={gr}", flush=True)
    ref = oracle.flexible_dual_grid(v, f, vs, gr, fw, bw, rw)
    c, d, i = run_gpu(v, f, vs, gr, fw, bw, rw)
    perm = oracle.canonical_order(ref.coords, gr)
    assert c.shape == ref.coords.shape, f"{label}: N {c.shape[0]} vs oracle {ref.coords.shape[0]}"
    assert np.array_equal(c, ref.coords[perm]), f"{label}: occupancy mismatch"
    assert np.array_equal(i, ref.flags[perm]), f"{label}: intersected flags mismatch"
    if len(c) == 0:
        return 0.0, 0.0, 1.0
    vox = np.asarray(vs, np.float64)[None, :]
    diff = (d.astype(np.float64) - ref.dual[perm].astype(np.float64)) / vox
    rms = float(np.sqrt((diff ** 2).sum(1).mean()))
    mx = float(np.abs(diff).max())
    exact = float(np.mean(np.all(d == ref.dual[perm], axis=1)))
    print(f"[case] {label}: N={len(c)} dual rms={rms:.3e} max={mx:.3e} voxel, bit-identical {exact:.6f}", flush=True)
    assert rms <= DUAL_RMS_TOL and mx <= DUAL_MAX_TOL, (label, rms, mx)
    return rms, mx, exact


def test_ico_cap_grid64(oracle):
    v, f = meshgen.ico_cap()
    vs, gr, a0 = meshgen.grid_args(64)
    rms, mx, exact = compare(oracle, (v - a0).astype(np.float32), f, vs, gr, label="ico_cap/64")
    assert exact == 1.0


@pytest.mark.parametrize("grid", [16, 37, 128])
def test_ico_cap_other_grids(oracle, grid):
    v, f = meshgen.ico_cap()
    vs, gr, a0 = meshgen.grid_args(grid)
    compare(oracle, (v - a0).astype(np.float32), f, vs, gr, label=f"ico_cap/{grid}")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_small_scene(oracle, seed):
    v, f = meshgen.scene(seed=seed, target_faces=20000)
    vs, gr, a0 = meshgen.grid_args(128)
    compare(oracle, (v - a0).astype(np.float32), f, vs, gr, label=f"scene20k/s{seed}")


def test_weights_variants(oracle):
    v, f = meshgen.scene(seed=7, target_faces=8000)
    vs, gr, a0 = meshgen.grid_args(96)
    vl = (v - a0).astype(np.float32)
    for fw, bw, rw in [(0.0, 0.0, 0.0), (1.0, 0.0, 0.1), (0.0, 1.0, 0.1), (2.5, 0.7, 0.0), (1.0, 1.0, 1.0)]:
        compare(oracle, vl, f, vs, gr, fw, bw, rw, label=f"weights fw={fw} bw={bw} rw={rw}")


def test_anisotropic_grid_and_offset(oracle):
    v, f = meshgen.scene(seed=11, target_faces=6000)
    vl = (v + 0.5).astype(np.float32)
    vs = [1.0 / 70, 1.0 / 45, 1.0 / 90]
    gr = [0, 0, 0, 70, 45, 90]
    compare(oracle, vl, f, vs, gr, label="anisotropic")
    gr2 = [5, -3, 10, 60, 40, 80]
    compare(oracle, vl, f, vs, gr2, label="offset grid_min")


def test_empty_and_outside(oracle):
    vs, gr, a0 = meshgen.grid_args(32)
    v = np.zeros((0, 3), np.float32)
    f = np.zeros((0, 3), np.int32)
    c, d, i = run_gpu(v, f, vs, gr, 1.0, 0.2, 1e-2)
    assert c.shape == (0, 3) and d.shape == (0, 3) and i.shape == (0, 3)
    # all faces outside the grid
    v = np.array([[5, 5, 5], [6, 5, 5], [5, 6, 5]], np.float32)
    f = np.array([[0, 1, 2]], np.int32)
    c, d, i = run_gpu(v, f, vs, gr, 1.0, 0.2, 1e-2)
    assert c.shape == (0, 3)


def test_degenerate_triangles(oracle):
    rng = np.random.default_rng(5)
    pts = rng.uniform(0.1, 0.9, size=(60, 3)).astype(np.float32)
    faces = []
    for k in range(20):
        a, b, c = 3 * k, 3 * k + 1, 3 * k + 2
        faces.append([a, b, c])
    # zero-area (collinear / repeated vertex) triangles and a repeated index
    pts[3:6] = np.array([[0.2, 0.2, 0.2], [0.4, 0.4, 0.4], [0.6, 0.6, 0.6]], np.float32)
    faces.append([0, 0, 1])
    faces.append([7, 7, 7])
    f = np.array(faces, np.int32)
    vs = [1.0 / 40] * 3
    gr = [0, 0, 0, 40, 40, 40]
    compare(oracle, pts, f, vs, gr, label="degenerate")


def test_intersect_occ(oracle):
    v, f = meshgen.scene(seed=3, target_faces=10000)
    vs, gr, a0 = meshgen.grid_args(100)
    tri = (v - a0).astype(np.float32)[f]
    ref = oracle.intersect_occ(tri, vs, gr)
    dev = _dev()
    c = _C().intersect_occ_cuda(torch.from_numpy(tri).to(dev), vs, gr).cpu().numpy()
    perm = oracle.canonical_order(ref, gr)
    assert np.array_equal(c, ref[perm])


def test_public_api_matches_c_boundary(oracle):
    import o_voxel

    v, f = meshgen.ico_cap()
    dev = _dev()
    aabb = [[-0.5] * 3, [0.5] * 3]
    c, d, i = o_voxel.convert.mesh_to_flexible_dual_grid(torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev),
                                                         grid_size=64, aabb=aabb, face_weight=1.0,
                                                         boundary_weight=0.2, regularization_weight=1e-2)
    vs, gr, a0 = meshgen.grid_args(64)
    ref = oracle.flexible_dual_grid((v - a0).astype(np.float32), f.astype(np.int32), vs, gr, 1.0, 0.2, 1e-2)
    perm = oracle.canonical_order(ref.coords, gr)
    assert np.array_equal(c.cpu().numpy(), ref.coords[perm])
    assert np.array_equal(d.cpu().numpy(), ref.dual[perm])
    assert np.array_equal(i.cpu().numpy(), ref.flags[perm])


def test_prepare_mesh_matches_torch():
    """The wrapper's input preparation (one library kernel) gives exactly
    torch's `vertices - aabb[0]` and `faces.to(int32)`; an int64 index outside
    [0, 2^31) becomes -1, which the public API reports as out of range
    (a plain int32 cast would wrap 2^32 to the valid index 0)."""
    dev = _dev()
    rng = np.random.default_rng(11)
    v = torch.from_numpy(rng.normal(0.0, 3.0, (100003, 3)).astype(np.float32)).to(dev)
    o = torch.from_numpy(rng.normal(0.0, 1.0, 3).astype(np.float32)).to(dev)
    f = torch.from_numpy(rng.integers(0, 100003, (77777, 3))).to(dev)
    print("[case] prepare: int64 faces", flush=True)
    vp, fp = _C()._prepare_mesh(v, f, o)
    torch.cuda.synchronize()
    assert torch.equal(vp, v - o.reshape(1, 3))
    assert fp.dtype == torch.int32 and torch.equal(fp, f.to(torch.int32))
    print("[case] prepare: int32 faces", flush=True)
    f32 = f.to(torch.int32)
    vp2, fp2 = _C()._prepare_mesh(v, f32, o)
    assert fp2 is f32 and torch.equal(vp2, vp)
    print("[case] prepare: views not 16-byte aligned (element-wise path)", flush=True)
    vp3, fp3 = _C()._prepare_mesh(v[1:], f[1:], o)
    assert torch.equal(vp3, v[1:] - o.reshape(1, 3)) and torch.equal(fp3, f[1:].to(torch.int32))
    print("[case] prepare: int64 indices outside [0, 2^31)", flush=True)
    fb = f.clone()
    fb[5, 1] = 1 << 33
    fb[9, 2] = -7
    fb[11, 0] = 1 << 31
    _, fpb = _C()._prepare_mesh(v, fb, o)
    assert fpb[5, 1].item() == -1 and fpb[9, 2].item() == -1 and fpb[11, 0].item() == -1
    assert torch.equal(fpb[12:], fp[12:])
    import o_voxel

    vv, ff = meshgen.ico_cap()
    ff = ff.astype(np.int64).copy()
    ff[3, 0] = 1 << 32
    print("[case] public API: int64 index 2^32", flush=True)
    with pytest.raises(RuntimeError, match="vertex index"):
        o_voxel.convert.mesh_to_flexible_dual_grid(torch.from_numpy(vv).to(dev), torch.from_numpy(ff).to(dev),
                                                   grid_size=64, aabb=[[-0.5] * 3, [0.5] * 3])


def test_determinism(oracle):
    v, f = meshgen.scene(seed=4, target_faces=15000)
    vs, gr, a0 = meshgen.grid_args(128)
    vl = (v - a0).astype(np.float32)
    a = run_gpu(vl, f, vs, gr, 1.0, 0.2, 1e-2)
    b = run_gpu(vl, f, vs, gr, 1.0, 0.2, 1e-2)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_buffer_overflow_fallbacks(oracle):
    """Tiny crossing / hit buffers force the re-recording of crossings and the
    traversal-based emit; the results must not change."""
    v, f = meshgen.scene(seed=6, target_faces=12000)
    vs, gr, a0 = meshgen.grid_args(128)
    vl = (v - a0).astype(np.float32)
    base = run_gpu(vl, f, vs, gr, 1.0, 0.2, 1e-2)
    dev = _dev()
    _C().set_capacity(dev, events=1000, hits=1000)
    small = run_gpu(vl, f, vs, gr, 1.0, 0.2, 1e-2)   # both buffers overflow
    again = run_gpu(vl, f, vs, gr, 1.0, 0.2, 1e-2)   # grown buffers
    for x, y, z in zip(base, small, again):
        assert np.array_equal(x, y) and np.array_equal(x, z)
    compare(oracle, vl, f, vs, gr, label="after overflow")


def test_invalid_inputs_raise(oracle):
    """A face index outside [0, V) or a non-finite vertex is reported as an
    input error (the kernels that run before the report stay inside the
    vertex array), and the context keeps working afterwards."""
    v, f = meshgen.ico_cap()
    vs, gr, a0 = meshgen.grid_args(64)
    vl = (v - a0).astype(np.float32)
    for bad in (len(vl), len(vl) + 1000, -1):
        fb = f.astype(np.int32).copy()
        fb[7, 1] = bad
        print(f"[case] face index {bad}", flush=True)
        with pytest.raises(RuntimeError, match="vertex index"):
            run_gpu(vl, fb, vs, gr, 1.0, 0.2, 1e-2)
    vb = vl.copy()
    vb[3, 0] = np.nan
    print("[case] non-finite vertex", flush=True)
    with pytest.raises(RuntimeError, match="non-finite"):
        run_gpu(vb, f, vs, gr, 1.0, 0.2, 1e-2)
    compare(oracle, vl, f, vs, gr, label="after invalid inputs")


def test_caller_stream_and_back_to_back(oracle):
    """The conversion forks its boundary-edge and record stages onto a second
    stream and joins them back: on a caller's non-default stream, with
    conversions queued back to back and no host wait between them, the
    results equal the default-stream ones bit for bit and match the oracle."""
    import o_voxel

    dev = _dev()
    v, f = meshgen.scene(seed=4, target_faces=30000, n_closed=6, n_patches=6)
    vs, gr, a0 = meshgen.grid_args(128)
    vl = (v - a0).astype(np.float32)  # grid-local, as the public API shifts them
    want = run_gpu(vl, f, vs, gr, 1.0, 0.2, 1e-2)
    compare(oracle, vl, f, vs, gr, label="scene 30k @128 (default stream)")
    vt = torch.from_numpy(v).to(dev)
    ft = torch.from_numpy(f.astype(np.int64)).to(dev)
    aabb = [[-0.5] * 3, [0.5] * 3]
    st = torch.cuda.Stream(dev)  # default priority
    print("[case] 3 conversions queued on a caller stream", flush=True)
    with torch.cuda.stream(st):
        outs = [o_voxel.convert.mesh_to_flexible_dual_grid(vt, ft, grid_size=128, aabb=aabb, face_weight=1.0,
                                                           boundary_weight=0.2, regularization_weight=1e-2)
                for _ in range(3)]
    st.synchronize()
    for out in outs:
        got = [t.cpu().numpy() for t in out]
        for a, b in zip(got, want):
            assert a.shape == b.shape
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_alternating_streams_share_the_context():
    """Conversions alternating between two caller streams with no host wait
    between them (the library's workspace is one per device): each call's
    work is ordered after the previous call's, so every result equals the
    single-stream one bit for bit."""
    import o_voxel

    dev = _dev()
    aabb = [[-0.5] * 3, [0.5] * 3]
    meshes = []
    for seed, nf, g in ((41, 30000, 128), (42, 8000, 64), (43, 60000, 256)):
        v, f = meshgen.scene(seed=seed, target_faces=nf, n_closed=6, n_patches=6)
        meshes.append((torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev), g))

    def conv(m):
        return o_voxel.convert.mesh_to_flexible_dual_grid(m[0], m[1], grid_size=m[2], aabb=aabb, face_weight=1.0,
                                                          boundary_weight=0.2, regularization_weight=1e-2)

    want = []
    for m in meshes:
        want.append([t.cpu() for t in conv(m)])
    streams = [torch.cuda.Stream(dev), torch.cuda.Stream(dev)]  # default priority
    print("[case] 12 conversions alternating between two streams", flush=True)
    outs = []
    for k in range(12):
        with torch.cuda.stream(streams[k % 2]):
            outs.append((k % 3, conv(meshes[k % 3])))
    for s in streams:
        s.synchronize()
    for k, out in outs:
        for a, b in zip(out, want[k]):
            a = a.cpu()
            assert a.shape == b.shape and torch.equal(a.view(torch.uint8), b.view(torch.uint8))


@pytest.mark.parametrize("gmin", [0, -4096])
def test_full_extent_axis(oracle, gmin):
    """A zig-zag strip across a grid axis of the maximum extent (8192
    voxels): scanline rows run along that axis for one scan axis and columns
    for another, so row / column offsets from the grid minimum reach 8191;
    also with a negative grid minimum."""
    n = 300
    x = (gmin + np.linspace(0.0, 8192.0, n + 1)[:-1] + 0.37) / 8192.0
    x = np.concatenate([x, [(gmin + 8191.6) / 8192.0]])
    y = np.where(np.arange(n + 1) % 2 == 0, 0.13, 0.87)
    z = 0.21 + 0.55 * y + 0.05 * np.sin(np.arange(n + 1))
    v = np.stack([x, y, z], axis=1).astype(np.float32)
    f = np.array([[i, i + 1, i + 2] for i in range(n - 1)], np.int32)
    vs = [1.0 / 8192, 1.0 / 8, 1.0 / 8]
    gr = [gmin, 0, 0, gmin + 8192, 8, 8]
    rms, mx, exact = compare(oracle, v, f, vs, gr, label=f"strip over 8192 voxels, gmin {gmin}")
    assert exact == 1.0


def _fan_mesh(k_closed, seed):
    """Fans whose centre vertex is the smaller vertex of all its spokes: a
    closed fan of k triangles gives it 2k half-edges, an open one of k
    triangles 2k as well (two of them boundary edges).  Centres sit on both
    sides of the 512-vertex bins' borders, among unreferenced vertices."""
    rng = np.random.default_rng(seed)
    V = 1500
    pts = rng.uniform(0.05, 0.95, size=(V, 3))
    faces = []
    for n, c in enumerate((0, 500, 511, 1000, 1023, 1400)):
        closed = n % 2 == 0
        k = k_closed if (closed and n == 2) else 8
        ring = list(range(c + 1, c + 1 + (k if closed else k + 1)))
        centre = rng.uniform(0.2, 0.8, size=3)
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        ang = np.linspace(0.0, 2.0 * np.pi, len(ring), endpoint=False) if closed else np.linspace(0.0, 1.6 * np.pi, len(ring))
        r = rng.uniform(0.05, 0.12)
        local = np.stack([r * np.cos(ang), r * np.sin(ang), 0.3 * r * np.sin(3.0 * ang)], 1)
        pts[c] = centre
        pts[ring] = centre + local @ q.T
        for i in range(len(ring) if closed else len(ring) - 1):
            faces.append([c, ring[i], ring[(i + 1) % len(ring)]])
    return pts.astype(np.float32), np.array(faces, np.int32)


@pytest.mark.parametrize("k_closed", [8, 9])
def test_boundary_edges_at_bin_borders_and_slot_limit(oracle, k_closed):
    """A centre vertex with 16 half-edges fits the per-vertex slots of the
    binned boundary-edge stage; with 18 the conversion falls back to per-vertex
    lists of any length.  Both equal the oracle."""
    v, f = _fan_mesh(k_closed, seed=31 + k_closed)
    vs = [1.0 / 64] * 3
    gr = [0, 0, 0, 64, 64, 64]
    compare(oracle, v, f, vs, gr, label=f"fans, {2 * k_closed} half-edges at one vertex")
