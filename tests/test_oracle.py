"""Pins of the CPU oracle (no GPU).

The reference ships no tests or golden vectors for this path, and its CPU
implementation cannot be built here (it needs Eigen).  Each oracle building
block is therefore checked against something independent of its own code:
closed forms, brute-force geometric predicates evaluated a different way
(point-in-triangle for the scanline, 13-axis separating-axis test for the face
overlap, slab clipping for the boundary DDA), LAPACK (scipy) for the
column-pivoted QR, and exhaustive search for the constrained solve.
Inputs that fall within a small margin of a decision boundary are excluded
from the set comparisons, since there the result legitimately depends on the
exact float rounding the oracle reproduces.
"""

import numpy as np
import pytest
import scipy.linalg

import meshgen


# ---------------------------------------------------------------------------
# scanline events vs point-in-triangle brute force
# ---------------------------------------------------------------------------
def brute_events(tri, G, vs):
    """Events (ax2, x_idx, y_idx, z_idx) by testing every interior grid line
    against the projected triangle (float64 barycentrics).  Returns the set
    of sure events and the set of ambiguous (near-boundary) ones."""
    sure, amb = set(), set()
    t = tri.astype(np.float64)
    for ax2 in range(3):
        ax0, ax1 = (ax2 + 1) % 3, (ax2 + 2) % 3
        P = t[:, [ax0, ax1]]
        Z = t[:, ax2]
        (x0, y0), (x1, y1), (x2, y2) = P
        det = (y1 - y2) * (x0 - x2) + (x2 - x1) * (y0 - y2)
        if abs(det) < 1e-14:
            continue
        ks = np.arange(1, G)  # interior grid lines k * vs, k = 1 .. G-1
        X, Y = np.meshgrid(ks * vs, ks * vs, indexing="ij")
        l0 = ((y1 - y2) * (X - x2) + (x2 - x1) * (Y - y2)) / det
        l1 = ((y2 - y0) * (X - x2) + (x0 - x2) * (Y - y2)) / det
        l2 = 1.0 - l0 - l1
        lmin = np.minimum(np.minimum(l0, l1), l2)
        z = l0 * Z[0] + l1 * Z[1] + l2 * Z[2]
        zq = z / vs
        zi = np.trunc(zq).astype(np.int64)  # the reference's int() truncates toward zero
        inside = lmin > 1e-9
        near = np.abs(lmin) <= 1e-9
        zfrac = np.abs(zq - np.round(zq))
        for (a, b) in zip(*np.nonzero(inside | near)):
            ev = (ax2, int(ks[a] - 1), int(ks[b] - 1), int(zi[a, b]))
            ok_z = 0 <= zi[a, b] < G
            if near[a, b] or zfrac[a, b] < 1e-9:
                amb.add(ev)
                # the neighbouring z layer is also plausible at a z boundary
                amb.add((ax2, ev[1], ev[2], ev[3] - 1))
            elif ok_z:
                sure.add(ev)
    return sure, amb


@pytest.mark.parametrize("seed", range(12))
def test_scan_events_vs_point_in_triangle(oracle, seed):
    rng = np.random.default_rng(seed)
    G = int(rng.integers(8, 40))
    vs = 1.0 / G
    for _ in range(20):
        scale = rng.choice([0.05, 0.2, 0.6])
        c = rng.uniform(0.1, 0.9, size=3)
        tri = (c + rng.uniform(-scale, scale, size=(3, 3))).astype(np.float32)
        ev = oracle.scan_events(tri, [vs] * 3, [0, 0, 0, G, G, G])
        got = set(map(tuple, ev.tolist()))
        sure, amb = brute_events(tri, G, np.float64(np.float32(vs)))
        assert sure <= got, (seed, sorted(sure - got)[:5])
        assert got <= (sure | amb), (seed, sorted(got - sure - amb)[:5])


def test_scan_square_closed_form(oracle):
    """An axis-aligned square at a non-grid z: events are exactly the interior
    grid lines along z, one voxel layer, z-flags on those lines."""
    G = 32
    vs = 1.0 / G
    x0, x1, y0, y1, z0 = 0.2131, 0.6789, 0.3047, 0.7211, 0.4443
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    res = oracle.flexible_dual_grid(v, f, [vs] * 3, [0, 0, 0, G, G, G], 1.0, 0.0, 0.1)
    zi = int(np.floor(np.float32(z0) / np.float32(vs)))
    ii = [i for i in range(G - 1) if x0 < (i + 1) * vs < x1]
    jj = [j for j in range(G - 1) if y0 < (j + 1) * vs < y1]
    want_vox = {(i + dx, j + dy, zi) for i in ii for j in jj for dx in (0, 1) for dy in (0, 1)}
    got_vox = set(map(tuple, res.coords.tolist()))
    assert got_vox == want_vox
    flags = {tuple(c): tuple(fl) for c, fl in zip(res.coords.tolist(), res.flags.tolist())}
    for (x, y, z), fl in flags.items():
        assert fl[0] is False and fl[1] is False
        assert fl[2] == (x in ii and y in jj)
    assert res.stats["events"] == len(ii) * len(jj)
    # every dual lies in the plane z = z0 (single plane quadric + regularisation towards the mean)
    assert np.allclose(res.dual[:, 2], np.float32(z0), atol=2e-7)


def test_occupancy_invariants(oracle):
    v, f = meshgen.scene(seed=21, target_faces=3000)
    G = 48
    vs, gr, a0 = meshgen.grid_args(G)
    vl = (v - a0).astype(np.float32)
    base = oracle.flexible_dual_grid(vl, f, vs, gr, 1.0, 0.2, 1e-2)
    key = lambda c: set(map(tuple, c.tolist()))
    # face permutation: same voxel set, same flags
    perm = np.random.default_rng(0).permutation(len(f))
    r2 = oracle.flexible_dual_grid(vl, f[perm], vs, gr, 1.0, 0.2, 1e-2)
    assert key(r2.coords) == key(base.coords)
    fl = {tuple(c): tuple(x) for c, x in zip(base.coords.tolist(), base.flags.tolist())}
    fl2 = {tuple(c): tuple(x) for c, x in zip(r2.coords.tolist(), r2.flags.tolist())}
    assert fl == fl2
    # winding flip
    r3 = oracle.flexible_dual_grid(vl, f[:, [0, 2, 1]], vs, gr, 1.0, 0.2, 1e-2)
    assert key(r3.coords) == key(base.coords)
    # an intersected flag along axis a implies the 4 voxels around that grid edge exist
    occ = key(base.coords)
    for (x, y, z), flg in fl.items():
        for a in range(3):
            if flg[a]:
                a0_, a1_ = (a + 1) % 3, (a + 2) % 3
                for dx in (0, 1):
                    for dy in (0, 1):
                        c = [x, y, z]
                        c[a0_] += dx
                        c[a1_] += dy
                        assert tuple(c) in occ


def test_intersect_occ_matches_fdg_occupancy(oracle):
    v, f = meshgen.scene(seed=22, target_faces=4000)
    vs, gr, a0 = meshgen.grid_args(40)
    vl = (v - a0).astype(np.float32)
    occ = oracle.intersect_occ(vl[f], vs, gr)
    res = oracle.flexible_dual_grid(vl, f, vs, gr, 1.0, 0.2, 1e-2)
    assert np.array_equal(occ, res.coords)  # same first-insertion order too


# ---------------------------------------------------------------------------
# face overlap test vs 13-axis separating-axis test
# ---------------------------------------------------------------------------
def sat_overlap(tri, lo, hi):
    """Closed triangle vs closed box, float64 SAT.  Returns (overlap, margin)."""
    c = 0.5 * (lo + hi)
    h = 0.5 * (hi - lo)
    t = tri - c
    e = [t[1] - t[0], t[2] - t[1], t[0] - t[2]]
    axes = [np.eye(3)[k] for k in range(3)]
    n = np.cross(e[0], e[1])
    axes.append(n)
    for ei in e:
        for k in range(3):
            axes.append(np.cross(ei, np.eye(3)[k]))
    margin = np.inf
    for a in axes:
        na = np.linalg.norm(a)
        if na < 1e-12:
            continue
        a = a / na
        p = t @ a
        r = np.abs(a) @ h
        gap = max(p.min() - r, -r - p.max())  # > 0 separated
        margin = min(margin, abs(gap))
        if gap > 0:
            return False, abs(gap)
    return True, margin


@pytest.mark.parametrize("seed", range(6))
def test_face_hits_vs_sat(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    G = 24
    vs = np.float32(1.0 / G)
    for _ in range(15):
        c = rng.uniform(0.2, 0.8, size=3)
        tri = (c + rng.uniform(-0.12, 0.12, size=(3, 3))).astype(np.float32)
        hits = set(map(tuple, oracle.face_hits(tri, [vs] * 3, [0, 0, 0, G, G, G]).tolist()))
        lo_v = np.floor(tri.min(0) / vs).astype(int) - 1
        hi_v = np.floor(tri.max(0) / vs).astype(int) + 1
        t64 = tri.astype(np.float64)
        for x in range(max(lo_v[0], 0), min(hi_v[0] + 1, G)):
            for y in range(max(lo_v[1], 0), min(hi_v[1] + 1, G)):
                for z in range(max(lo_v[2], 0), min(hi_v[2] + 1, G)):
                    lo = np.array([x, y, z], np.float64) * np.float64(vs)
                    ov, margin = sat_overlap(t64, lo, lo + np.float64(vs))
                    if margin < 1e-6:
                        continue
                    assert ((x, y, z) in hits) == ov, (seed, (x, y, z), ov, margin)


# ---------------------------------------------------------------------------
# boundary DDA vs slab clipping; boundary quadric closed form
# ---------------------------------------------------------------------------
def seg_box(p0, p1, lo, hi):
    """Closed segment vs closed box via slab clipping; returns (hit, margin)."""
    d = p1 - p0
    t0, t1 = 0.0, 1.0
    margin = np.inf
    for a in range(3):
        if abs(d[a]) < 1e-15:
            if p0[a] < lo[a] or p0[a] > hi[a]:
                return False, min(abs(p0[a] - lo[a]), abs(p0[a] - hi[a]))
            margin = min(margin, p0[a] - lo[a], hi[a] - p0[a])
            continue
        ta, tb = (lo[a] - p0[a]) / d[a], (hi[a] - p0[a]) / d[a]
        if ta > tb:
            ta, tb = tb, ta
        t0, t1 = max(t0, ta), min(t1, tb)
    L = np.linalg.norm(d)
    margin = min(margin, abs(t1 - t0) * L)
    return t0 <= t1, margin


@pytest.mark.parametrize("seed", range(6))
def test_boundary_dda_vs_slab(oracle, seed):
    rng = np.random.default_rng(200 + seed)
    G = 20
    vs = np.float32(1.0 / G)
    for _ in range(25):
        p0 = rng.uniform(0.05, 0.95, size=3)
        p1 = p0 + rng.normal(scale=0.2, size=3)
        seg = np.stack([p0, p1]).astype(np.float32)
        vox, Q = oracle.boundary_voxels(seg, [vs] * 3, [0, 0, 0, G, G, G])
        got = set(map(tuple, vox.tolist()))
        assert len(got) == len(vox), "DDA visited a voxel twice"
        s64 = seg.astype(np.float64)
        for x in range(G):
            for y in range(G):
                for z in range(G):
                    lo = np.array([x, y, z], np.float64) * np.float64(vs)
                    hit, margin = seg_box(s64[0], s64[1], lo, lo + np.float64(vs))
                    if margin < 1e-6:
                        continue
                    assert ((x, y, z) in got) == hit, (seed, (x, y, z), hit)
        # quadric: A = I - d d^T (eigenvalues 0, 1, 1); zero on the line
        d = s64[1] - s64[0]
        d /= np.linalg.norm(d)
        A = Q[:3, :3].astype(np.float64)
        assert np.allclose(A, np.eye(3) - np.outer(d, d), atol=1e-6)
        assert np.allclose(Q, Q.T)
        for tpar in (0.0, 0.3, 1.0):
            p = np.append(s64[0] + tpar * (s64[1] - s64[0]), 1.0)
            assert abs(p @ Q.astype(np.float64) @ p) < 1e-5


def test_boundary_edges_follow_edge_multiset(oracle):
    """A single open triangle has 3 boundary edges, a closed tetrahedron none,
    and a quad split in two has 4 (the shared diagonal is interior)."""
    G = 16
    vs = [1.0 / G] * 3
    gr = [0, 0, 0, G, G, G]
    tri_v = np.array([[0.1, 0.1, 0.3], [0.8, 0.2, 0.35], [0.3, 0.9, 0.4]], np.float32)
    r = oracle.flexible_dual_grid(tri_v, np.array([[0, 1, 2]], np.int32), vs, gr, 1.0, 1.0, 0.1)
    assert r.stats["boundary_edges"] == 3
    tet_v = np.array([[0.2, 0.2, 0.2], [0.8, 0.25, 0.3], [0.4, 0.8, 0.25], [0.45, 0.4, 0.85]], np.float32)
    tet_f = np.array([[0, 2, 1], [0, 1, 3], [1, 2, 3], [2, 0, 3]], np.int32)
    r = oracle.flexible_dual_grid(tet_v, tet_f, vs, gr, 1.0, 1.0, 0.1)
    assert r.stats["boundary_edges"] == 0
    quad_v = np.array([[0.1, 0.1, 0.5], [0.9, 0.1, 0.52], [0.9, 0.9, 0.55], [0.1, 0.9, 0.5]], np.float32)
    r = oracle.flexible_dual_grid(quad_v, np.array([[0, 1, 2], [0, 2, 3]], np.int32), vs, gr, 1.0, 1.0, 0.1)
    assert r.stats["boundary_edges"] == 4


# ---------------------------------------------------------------------------
# quadric accumulation: single plane closed form, PSD
# ---------------------------------------------------------------------------
def test_triangle_plane_closed_form(oracle):
    rng = np.random.default_rng(3)
    for _ in range(50):
        tri = rng.uniform(-1, 1, size=(3, 3)).astype(np.float32)
        pl = oracle.triangle_plane(*tri).astype(np.float64)
        t = tri.astype(np.float64)
        n = np.cross(t[1] - t[0], t[2] - t[1])
        n /= np.linalg.norm(n)
        assert np.allclose(pl[:3], n, atol=1e-6)
        assert abs(pl[3] + n @ t[0]) < 1e-6
        for p in t:  # all three vertices lie on the plane
            assert abs(pl[:3] @ p + pl[3]) < 1e-5
    # degenerate: zero normal stays zero (normalized() of a zero vector)
    z = oracle.triangle_plane([0, 0, 0], [1, 1, 1], [2, 2, 2])
    assert np.all(z[:3] == 0.0)


def test_single_plane_quadric_sum(oracle):
    """A planar patch: every voxel's FP64 quadric is k * p p^T with k the
    number of contributions it received, and its dual lies on the plane."""
    G = 24
    vs = [1.0 / G] * 3
    n = np.array([0.3, -0.5, 0.81])
    n /= np.linalg.norm(n)
    u = np.cross(n, [1, 0, 0])
    u /= np.linalg.norm(u)
    w = np.cross(n, u)
    c = np.array([0.5, 0.5, 0.5])
    quad = np.array([c + 0.4 * (a * u + b * w) for a, b in [(-1, -1), (1, -1), (1, 1), (-1, 1)]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    r = oracle.flexible_dual_grid(quad, f, vs, [0, 0, 0, G, G, G], 1.0, 0.0, 0.0, fp64=True, keep_qef=True)
    # the two float32 triangles are coplanar only to float32 rounding (~1e-7)
    pl = np.append(n, -n @ c)
    P = np.outer(pl, pl)
    for q in r.qef:
        k = q[0, 0] / P[0, 0]
        assert abs(k - round(k)) < 1e-5
        assert np.allclose(q, round(k) * P, atol=1e-5 * max(1, k))
    d = r.dual64
    assert np.allclose(d @ n - n @ c, 0.0, atol=1e-6)


def test_quadrics_psd(oracle):
    v, f = meshgen.scene(seed=9, target_faces=2000)
    vs, gr, a0 = meshgen.grid_args(32)
    r = oracle.flexible_dual_grid((v - a0).astype(np.float32), f, vs, gr, 1.0, 0.5, 0.0, fp64=True, keep_qef=True)
    ev = np.linalg.eigvalsh(r.qef)
    assert (ev.min(axis=1) > -1e-9 * np.abs(ev).max(axis=1)).all()
    # every intersection point of a voxel lies on one of its edges (a grid line
    # crossing inside the voxel's z range), so their mean lies in the closed voxel box
    m = r.mean / r.cnt[:, None]
    lo = r.coords.astype(np.float64) * np.float64(np.float32(vs[0]))
    assert (m >= lo - 1e-12).all() and (m <= lo + np.float64(np.float32(vs[0])) + 1e-12).all()
    assert (r.cnt >= 1).all() and (r.cnt == np.round(r.cnt)).all()


# ---------------------------------------------------------------------------
# column-pivoted QR vs LAPACK (scipy geqp3), closed forms
# ---------------------------------------------------------------------------
def lapack_basic_solution(A, b, rank):
    Q, R, P = scipy.linalg.qr(A, pivoting=True)
    y = Q.T @ b
    x = np.zeros(A.shape[1])
    x[P[:rank]] = scipy.linalg.solve_triangular(R[:rank, :rank], y[:rank])
    return x, P


@pytest.mark.parametrize("n", [2, 3])
def test_cpqr_full_rank(oracle, n):
    rng = np.random.default_rng(n)
    for _ in range(200):
        A = rng.normal(size=(n, n))
        if np.linalg.cond(A) > 1e3:
            continue
        b = rng.normal(size=n)
        x = oracle.cpqr_solve(A.astype(np.float32), b.astype(np.float32)).astype(np.float64)
        xr = np.linalg.solve(A.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64))
        assert np.allclose(x, xr, rtol=2e-4, atol=2e-4 * np.abs(xr).max())
        if n == 3:
            x64 = oracle.cpqr_solve(A, b)
            assert np.allclose(x64, np.linalg.solve(A, b), rtol=1e-10, atol=1e-10)


def test_cpqr_rank_deficient_matches_lapack_pivoting(oracle):
    """Exactly rank-deficient systems (zero columns survive the Householder
    steps exactly, so the nonzero-pivot count is the true rank): the basic
    solution equals LAPACK geqp3's basic solution with the same pivots, and
    the dropped unknowns are zero."""
    rng = np.random.default_rng(7)
    for rank in (1, 2):
        for _ in range(100):
            A = rng.normal(size=(3, 3))
            zero_cols = rng.choice(3, size=3 - rank, replace=False)
            A[:, zero_cols] = 0.0
            b = rng.normal(size=3)
            x = oracle.cpqr_solve(A, b)  # FP64 variant
            xr, P = lapack_basic_solution(A, b, rank)
            assert np.all(x[zero_cols] == 0.0)
            assert set(P[:rank].tolist()).isdisjoint(zero_cols.tolist())
            assert np.allclose(x, xr, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(xr).max()))
            # float variant: same structure, float tolerance
            xf = oracle.cpqr_solve(A.astype(np.float32), b.astype(np.float32)).astype(np.float64)
            assert np.all(xf[zero_cols] == 0.0)
            if np.linalg.cond(A[:, P[:rank]]) < 1e3:
                assert np.allclose(xf, xr, rtol=1e-3, atol=1e-3 * max(1.0, np.abs(xr).max()))


def test_cpqr_zero_matrix(oracle):
    """All-zero A: the nonzero-pivot test (max_norm*eps)^2/rows*(rows-k) is
    strict, 0 < 0 fails, so all 3 pivots count and the back substitution
    divides by zero: x = (nan, nan, +inf) for b = (1, 1, 1)."""
    x = oracle.cpqr_solve(np.zeros((3, 3), np.float32), np.ones(3, np.float32))
    assert np.isnan(x[0]) and np.isnan(x[1]) and np.isposinf(x[2])
    # zero right-hand side entries are skipped, so b = 0 gives x = 0
    x = oracle.cpqr_solve(np.zeros((3, 3), np.float32), np.zeros(3, np.float32))
    assert np.all(x == 0)


def test_degenerate_only_mesh_falls_back_to_corners(oracle):
    """Only zero-area triangles and no regularisation: every voxel quadric has
    a zero 3x3 block, the unconstrained solve is not finite, and the box
    search picks a voxel corner (finite output)."""
    v = np.array([[0.21, 0.33, 0.47], [0.61, 0.73, 0.87], [0.41, 0.53, 0.67]], np.float32)
    f = np.array([[0, 1, 2]], np.int32)
    G = 16
    r = oracle.flexible_dual_grid(v, f, [1.0 / G] * 3, [0, 0, 0, G, G, G], 1.0, 0.0, 0.0)
    assert np.isfinite(r.dual).all()
    corners = np.stack([r.coords, r.coords + 1]).astype(np.float32) * np.float32(1.0 / G)
    for k in range(len(r.coords)):
        for a in range(3):
            assert r.dual[k, a] in (corners[0, k, a], corners[1, k, a])


# ---------------------------------------------------------------------------
# constrained solve vs exhaustive search (FP64 path)
# ---------------------------------------------------------------------------
def test_constrained_solve_vs_search(oracle):
    """One voxel, one plane through it plus regularisation towards a point
    outside: the FP64 dual must be the box-constrained minimiser of the
    quadric (checked by dense search) and lie inside the voxel."""
    rng = np.random.default_rng(11)
    G = 8
    vs = 1.0 / G
    checked = constrained = 0
    for trial in range(40):
        c = rng.uniform(0.3, 0.7, size=3)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        u = np.cross(n, rng.normal(size=3))
        u /= np.linalg.norm(u)
        w = np.cross(n, u)
        tri = np.array([c + 0.05 * u, c - 0.04 * u + 0.03 * w, c - 0.01 * u - 0.05 * w], np.float32)
        r = oracle.flexible_dual_grid(tri, np.array([[0, 1, 2]], np.int32), [vs] * 3, [0, 0, 0, G, G, G],
                                      1.0, 0.0, 0.05, fp64=True, keep_qef=True)
        for k in range(len(r.coords)):
            q = r.qef[k].copy()
            cnt = r.cnt[k]
            p = r.mean[k] / cnt
            Rg = np.zeros((4, 4))
            Rg[:3, :3] = np.eye(3)
            Rg[:3, 3] = -p
            Rg[3, :3] = -p
            Rg[3, 3] = p @ p
            q = q + 0.05 * cnt * Rg
            lo = r.coords[k] * vs
            hi = lo + vs
            d = r.dual64[k]
            assert np.all(d >= lo - 1e-12) and np.all(d <= hi + 1e-12)
            # dense search over the box
            s = np.linspace(0, 1, 41)
            X, Y, Z = np.meshgrid(s, s, s, indexing="ij")
            pts = lo + vs * np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
            ph = np.concatenate([pts, np.ones((len(pts), 1))], 1)
            e = np.einsum("ni,ij,nj->n", ph, q, ph)
            dh = np.append(d, 1.0)
            assert dh @ q @ dh <= e.min() + 1e-12
            checked += 1
            xu = np.linalg.solve(q[:3, :3], -q[:3, 3])
            constrained += int(np.any(xu < lo) or np.any(xu > hi))
    assert checked > 20 and constrained >= 5, (checked, constrained)


def test_duals_inside_voxels(oracle):
    v, f = meshgen.ico_cap()
    vs, gr, a0 = meshgen.grid_args(64)
    r = oracle.flexible_dual_grid((v - a0).astype(np.float32), f, vs, gr, 1.0, 0.2, 1e-2)
    lo = r.coords.astype(np.float32) * np.float32(vs[0])
    hi = (r.coords + 1).astype(np.float32) * np.float32(vs[0])
    assert np.all(r.dual >= lo) and np.all(r.dual <= hi)


def test_fp64_and_float_paths_agree_on_membership(oracle):
    v, f = meshgen.ico_cap()
    vs, gr, a0 = meshgen.grid_args(64)
    vl = (v - a0).astype(np.float32)
    a = oracle.flexible_dual_grid(vl, f, vs, gr, 1.0, 0.2, 1e-2)
    b = oracle.flexible_dual_grid(vl, f, vs, gr, 1.0, 0.2, 1e-2, fp64=True)
    assert np.array_equal(a.coords, b.coords) and np.array_equal(a.flags, b.flags)
    assert a.stats == b.stats
    dv = (a.dual.astype(np.float64) - b.dual64) / vs[0]
    # the float path's rounding error relative to FP64, in voxels (reported, loose bound)
    assert np.sqrt((dv ** 2).sum(1).mean()) < 5e-2
