"""TEST INFRASTRUCTURE -- numpy/ctypes front end of the CPU parity oracle.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this module, and only as the checker (or the timed CPU baseline).
The product package never imports it.

The oracle is a CPU restatement of the reference CPU path
``src/convert/flexible_dual_grid.cpp:660-992`` (float path) plus an FP64
variant that keeps the float path's voxel/term membership and evaluates the
QEF terms and the solve in double.  Each building block is exposed so tests
can pin it against closed forms, scipy and brute force.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_fdg.so")
_lib: Optional[ctypes.CDLL] = None

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -O3 -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(
        os.path.join(_HERE, "fdg_oracle.cpp")
    ):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s", "liboracle_fdg.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_fdg_run.restype = ctypes.c_int64
        L.oracle_fdg_run.argtypes = [_f32p, ctypes.c_int64, _i32p, ctypes.c_int64, _f32p, _i64p, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_intersect_occ.restype = ctypes.c_int64
        L.oracle_intersect_occ.argtypes = [_f32p, ctypes.c_int64, _f32p, _i64p, ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_fdg_copy.restype = None
        L.oracle_fdg_copy.argtypes = [ctypes.c_void_p, _i32p, _f32p, _f64p, _u8p]
        L.oracle_fdg_copy_qef.restype = ctypes.c_int
        L.oracle_fdg_copy_qef.argtypes = [ctypes.c_void_p, _f64p, _f64p, _f64p]
        L.oracle_fdg_stats.restype = None
        L.oracle_fdg_stats.argtypes = [ctypes.c_void_p, _i64p]
        L.oracle_fdg_free.restype = None
        L.oracle_fdg_free.argtypes = [ctypes.c_void_p]
        L.oracle_cpqr_solve3_f32.argtypes = [_f32p, _f32p, _f32p]
        L.oracle_cpqr_solve2_f32.argtypes = [_f32p, _f32p, _f32p]
        L.oracle_cpqr_solve3_f64.argtypes = [_f64p, _f64p, _f64p]
        L.oracle_triangle_plane_f32.argtypes = [_f32p, _f32p, _f32p, _f32p]
        L.oracle_scan_events.restype = ctypes.c_int64
        L.oracle_scan_events.argtypes = [_f32p, _f32p, _i64p, _i32p, ctypes.c_int64]
        L.oracle_face_hits.restype = ctypes.c_int64
        L.oracle_face_hits.argtypes = [_f32p, _f32p, _i64p, _i32p, ctypes.c_int64]
        L.oracle_boundary_voxels.restype = ctypes.c_int64
        L.oracle_boundary_voxels.argtypes = [_f32p, _f32p, _i64p, _i32p, ctypes.c_int64, _f32p]
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def _grid_args(voxel_size, grid_range):
    vs = np.ascontiguousarray(np.asarray(voxel_size, dtype=np.float32).reshape(3))
    gr = np.ascontiguousarray(np.asarray(grid_range, dtype=np.int64).reshape(6))
    return vs, gr


class Result:
    """Outputs of one oracle conversion (numpy arrays, first-insertion order)."""

    def __init__(self, coords, dual, flags, dual64=None, qef=None, mean=None, cnt=None, stats=None):
        self.coords = coords
        self.dual = dual
        self.flags = flags
        self.dual64 = dual64
        self.qef = qef
        self.mean = mean
        self.cnt = cnt
        self.stats = stats


def flexible_dual_grid(vertices: np.ndarray, faces: np.ndarray, voxel_size, grid_range, face_weight: float = 1.0,
                       boundary_weight: float = 1.0, regularization_weight: float = 0.1, fp64: bool = False,
                       keep_qef: bool = False) -> Result:
    """Oracle for ``_C.mesh_to_flexible_dual_grid_cpu`` (FDG.cpp:660-992).

    ``vertices`` are already grid-local (shifted by aabb[0]) float32 [V,3];
    ``faces`` int32 [F,3].  Returns a :class:`Result` in first-insertion order.
    """
    L = lib()
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
    vs, gr = _grid_args(voxel_size, grid_range)
    h = ctypes.c_void_p()
    n = L.oracle_fdg_run(_p(v, _f32p), v.shape[0], _p(f, _i32p), f.shape[0], _p(vs, _f32p), _p(gr, _i64p),
                         float(face_weight), float(boundary_weight), float(regularization_weight), int(bool(fp64)),
                         int(bool(keep_qef)), ctypes.byref(h))
    if n < 0:
        raise ValueError("oracle_fdg_run: invalid arguments (face index out of range?)")
    try:
        coords = np.empty((n, 3), np.int32)
        dual = np.empty((n, 3), np.float32)
        flags = np.empty((n, 3), np.uint8)
        dual64 = np.empty((n, 3), np.float64) if fp64 else None
        L.oracle_fdg_copy(h, _p(coords, _i32p), _p(dual, _f32p), _p(dual64, _f64p) if fp64 else None,
                          _p(flags, _u8p))
        qef = mean = cnt = None
        if keep_qef:
            qef = np.empty((n, 4, 4), np.float64)
            mean = np.empty((n, 3), np.float64)
            cnt = np.empty((n,), np.float64)
            L.oracle_fdg_copy_qef(h, _p(qef, _f64p), _p(mean, _f64p), _p(cnt, _f64p))
        st = np.zeros(4, np.int64)
        L.oracle_fdg_stats(h, _p(st, _i64p))
        stats = dict(events=int(st[0]), face_hits=int(st[1]), boundary_edges=int(st[2]), boundary_hits=int(st[3]))
    finally:
        L.oracle_fdg_free(h)
    return Result(coords, dual, flags.astype(bool), dual64, qef, mean, cnt, stats)


def intersect_occ(triangles: np.ndarray, voxel_size, grid_range) -> np.ndarray:
    """Oracle for ``_C.intersect_occ_cpu`` (FDG.cpp:80-156, 605-644)."""
    L = lib()
    t = np.ascontiguousarray(triangles, dtype=np.float32).reshape(-1, 9)
    vs, gr = _grid_args(voxel_size, grid_range)
    h = ctypes.c_void_p()
    n = L.oracle_intersect_occ(_p(t, _f32p), t.shape[0], _p(vs, _f32p), _p(gr, _i64p), ctypes.byref(h))
    try:
        coords = np.empty((n, 3), np.int32)
        L.oracle_fdg_copy(h, _p(coords, _i32p), None, None, None)
    finally:
        L.oracle_fdg_free(h)
    return coords


def cpqr_solve(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Eigen-style ColPivHouseholderQR solve (float32 2x2/3x3, float64 3x3)."""
    L = lib()
    A = np.ascontiguousarray(A)
    b = np.ascontiguousarray(b)
    n = A.shape[0]
    if A.dtype == np.float64:
        assert n == 3
        x = np.zeros(3, np.float64)
        L.oracle_cpqr_solve3_f64(_p(A, _f64p), _p(b, _f64p), _p(x, _f64p))
        return x
    A = A.astype(np.float32)
    b = b.astype(np.float32)
    x = np.zeros(n, np.float32)
    if n == 3:
        L.oracle_cpqr_solve3_f32(_p(A, _f32p), _p(b, _f32p), _p(x, _f32p))
    elif n == 2:
        L.oracle_cpqr_solve2_f32(_p(A, _f32p), _p(b, _f32p), _p(x, _f32p))
    else:
        raise ValueError(n)
    return x


def triangle_plane(v0, v1, v2) -> np.ndarray:
    L = lib()
    a = [np.ascontiguousarray(np.asarray(v, np.float32).reshape(3)) for v in (v0, v1, v2)]
    out = np.zeros(4, np.float32)
    L.oracle_triangle_plane_f32(_p(a[0], _f32p), _p(a[1], _f32p), _p(a[2], _f32p), _p(out, _f32p))
    return out


def scan_events(tri: np.ndarray, voxel_size, grid_range) -> np.ndarray:
    """Scanline events of one triangle as rows (ax2, x_idx, y_idx, z_idx)."""
    L = lib()
    t = np.ascontiguousarray(tri, np.float32).reshape(9)
    vs, gr = _grid_args(voxel_size, grid_range)
    n = L.oracle_scan_events(_p(t, _f32p), _p(vs, _f32p), _p(gr, _i64p), None, 0)
    out = np.zeros((max(n, 1), 4), np.int32)
    L.oracle_scan_events(_p(t, _f32p), _p(vs, _f32p), _p(gr, _i64p), _p(out, _i32p), n)
    return out[:n]


def face_hits(tri: np.ndarray, voxel_size, grid_range) -> np.ndarray:
    """Voxels passing the face-stage overlap test for one triangle."""
    L = lib()
    t = np.ascontiguousarray(tri, np.float32).reshape(9)
    vs, gr = _grid_args(voxel_size, grid_range)
    n = L.oracle_face_hits(_p(t, _f32p), _p(vs, _f32p), _p(gr, _i64p), None, 0)
    out = np.zeros((max(n, 1), 3), np.int32)
    L.oracle_face_hits(_p(t, _f32p), _p(vs, _f32p), _p(gr, _i64p), _p(out, _i32p), n)
    return out[:n]


def boundary_voxels(seg: np.ndarray, voxel_size, grid_range) -> Tuple[np.ndarray, Optional[np.ndarray]]:
    """In-grid DDA voxels of one boundary segment and its float quadric."""
    L = lib()
    s = np.ascontiguousarray(seg, np.float32).reshape(6)
    vs, gr = _grid_args(voxel_size, grid_range)
    Q = np.zeros(16, np.float32)
    n = L.oracle_boundary_voxels(_p(s, _f32p), _p(vs, _f32p), _p(gr, _i64p), None, 0, _p(Q, _f32p))
    if n < 0:
        return np.zeros((0, 3), np.int32), None
    out = np.zeros((max(n, 1), 3), np.int32)
    L.oracle_boundary_voxels(_p(s, _f32p), _p(vs, _f32p), _p(gr, _i64p), _p(out, _i32p), n, None)
    return out[:n], Q.reshape(4, 4)


def canonical_order(coords: np.ndarray, grid_range) -> np.ndarray:
    """Permutation sorting voxels by x + Gx*(y + Gy*z) (the canonical order)."""
    gr = np.asarray(grid_range, np.int64)
    c = coords.astype(np.int64) - gr[:3]
    G = gr[3:] - gr[:3]
    key = c[:, 0] + G[0] * (c[:, 1] + G[1] * c[:, 2])
    return np.argsort(key, kind="stable")
