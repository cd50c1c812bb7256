"""``o_voxel._C`` for MI355X: the extension boundary of the flexible dual grid.

The reference binds these names from a torch C++ extension
(src/ext.cpp:18-21, stubs o_voxel/_C.pyi:10-38).  Here they are thin Python
functions over the C ABI of ``libovx_fdg.so`` (include/ovx_fdg.h): tensors
are checked, outputs are allocated by torch's caching allocator on the input
device, and the raw device pointers plus torch's current HIP stream go to the
library.  The library call releases the GIL (ctypes does), like the
reference's ``py::gil_scoped_release``.

There is no CPU implementation in this package: the ``*_cpu`` entry points
run the same HIP kernels on the current GPU (host -> device copy, convert,
device -> host copy) and raise if no GPU or no built library is present.
"""

from __future__ import annotations

import ctypes
import os
import threading
from typing import List, Tuple

import numpy as np
import torch

__all__ = [
    "mesh_to_flexible_dual_grid_cuda",
    "mesh_to_flexible_dual_grid_cpu",
    "intersect_occ_cuda",
    "intersect_occ_cpu",
    "library_path",
    "last_stats",
    "mesh_to_flexible_dual_grid_batch_cuda",
    "mesh_to_flexible_dual_grid_batch_packed_cuda",
    "stage_times",
    "set_timing",
    "kernel_times",
    "set_capacity",
    "set_brick_class",
    "get_brick_class",
    "set_graph",
    "graph_stats",
    "workspace_bytes",
    "flexible_dual_grid_to_mesh_cuda",
    "hashmap_insert_cuda",
    "hashmap_lookup_cuda",
    "hashmap_insert_3d_cuda",
    "hashmap_lookup_3d_cuda",
    "hashmap_insert_3d_idx_as_val_cuda",
    "z_order_encode_cuda",
    "z_order_decode_cuda",
    "hilbert_encode_cuda",
    "hilbert_decode_cuda",
    "z_order_encode_cpu",
    "z_order_decode_cpu",
    "hilbert_encode_cpu",
    "hilbert_decode_cpu",
    "encode_sparse_voxel_octree_cpu",
    "decode_sparse_voxel_octree_cpu",
    "encode_sparse_voxel_octree_attr_parent_cpu",
    "decode_sparse_voxel_octree_attr_parent_cpu",
    "encode_sparse_voxel_octree_attr_neighbor_cpu",
    "decode_sparse_voxel_octree_attr_neighbor_cpu",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("OVX_FDG_LIB", os.path.join(_HERE, "_lib", "libovx_fdg.so"))

_lib = None
_lib_lock = threading.Lock()   # guards loading the shared library
_ctx_lock = threading.Lock()   # guards creating per-device contexts
_ctxs = {}


class _Grid(ctypes.Structure):
    _fields_ = [("voxel_size", ctypes.c_float * 3), ("grid_range", ctypes.c_int64 * 6)]


def library_path() -> str:
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"o_voxel: the HIP library {_LIB_PATH} is not built; run `make -C o-voxel-gpu_amd/csrc` "
                "(or __graft_entry__.build()) first"
            )
        L = ctypes.CDLL(_LIB_PATH)
        vp, i64, f32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_float
        L.ovx_version.restype = ctypes.c_char_p
        L.ovx_last_error.restype = ctypes.c_char_p
        L.ovx_fdg_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.ovx_fdg_destroy.argtypes = [vp]
        L.ovx_fdg_workspace_bytes.argtypes = [vp]
        L.ovx_fdg_workspace_bytes.restype = i64
        L.ovx_fdg_trim.argtypes = [vp]
        L.ovx_fdg_prepare_mesh.argtypes = [vp, vp, i64, vp, i64, vp, vp, vp]
        L.ovx_fdg_batch_prepare.argtypes = [vp, ctypes.c_int, i64, i64, vp, vp, vp]
        L.ovx_fdg_begin.argtypes = [vp, vp, i64, vp, i64, ctypes.POINTER(_Grid), f32, f32, f32, vp,
                                    ctypes.POINTER(i64)]
        L.ovx_fdg_finish.argtypes = [vp, vp, vp, vp, vp]
        L.ovx_fdg_convert_noalloc.argtypes = [vp, vp, i64, vp, i64, ctypes.POINTER(_Grid), f32, f32, f32, vp, vp,
                                              vp, i64, vp, ctypes.POINTER(i64)]
        L.ovx_fdg_convert_noalloc.restype = ctypes.c_int
        L.ovx_fdg_batch_begin.argtypes = [vp, vp, i64, vp, i64, ctypes.c_int, ctypes.POINTER(i64),
                                          ctypes.POINTER(i64), ctypes.POINTER(_Grid), f32, f32, f32, vp,
                                          ctypes.POINTER(i64)]
        L.ovx_fdg_batch_begin.restype = ctypes.c_int
        L.ovx_occ_begin.argtypes = [vp, vp, i64, ctypes.POINTER(_Grid), vp, ctypes.POINTER(i64)]
        L.ovx_occ_finish.argtypes = [vp, vp, vp]
        L.ovx_fdg_stats.argtypes = [vp, ctypes.POINTER(i64), ctypes.c_int]
        L.ovx_fdg_set_timing.argtypes = [vp, ctypes.c_int]
        L.ovx_fdg_stage_times.argtypes = [vp, ctypes.POINTER(f32)]
        L.ovx_fdg_set_capacity.argtypes = [vp, i64, i64]
        L.ovx_fdg_set_precision.argtypes = [vp, ctypes.c_int, f32]
        L.ovx_fdg_set_precision.restype = ctypes.c_int
        f64p = ctypes.POINTER(ctypes.c_double)
        L.ovx_fdg_set_tile_classes.argtypes = [vp, f64p, f64p]
        L.ovx_fdg_get_tile_classes.argtypes = [vp, f64p, f64p]
        L.ovx_fdg_set_brick_class.argtypes = [vp, ctypes.c_double]
        L.ovx_fdg_set_brick_class.restype = ctypes.c_int
        L.ovx_fdg_get_brick_class.argtypes = [vp, f64p]
        L.ovx_fdg_get_brick_class.restype = ctypes.c_int
        L.ovx_fdg_set_graph.argtypes = [vp, ctypes.c_int]
        L.ovx_fdg_set_graph.restype = ctypes.c_int
        L.ovx_fdg_graph_stats.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.ovx_fdg_graph_stats.restype = ctypes.c_int
        L.ovx_fdg_set_capacity.restype = ctypes.c_int
        L.ovx_fdg_kernel_count.restype = ctypes.c_int
        L.ovx_fdg_kernel_name.argtypes = [ctypes.c_int]
        L.ovx_fdg_kernel_name.restype = ctypes.c_char_p
        L.ovx_fdg_kernel_times.argtypes = [vp, ctypes.POINTER(f32), ctypes.c_int]
        i32p, f32p = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(f32)
        L.ovx_dual_mesh_begin.argtypes = [vp, vp, vp, vp, i64, i32p, f32p, f32p, vp, vp, ctypes.POINTER(i64)]
        L.ovx_dual_mesh_finish.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp]
        ci = ctypes.c_int
        L.ovx_hashmap_insert.argtypes = [vp, ci, vp, ci, i64, vp, vp, i64, vp]
        L.ovx_hashmap_lookup.argtypes = [vp, ci, vp, ci, i64, vp, vp, i64, vp]
        L.ovx_hashmap_insert_3d.argtypes = [vp, ci, vp, ci, i64, vp, vp, i64, ci, ci, ci, vp]
        L.ovx_hashmap_lookup_3d.argtypes = [vp, ci, vp, ci, i64, vp, vp, i64, ci, ci, ci, vp]
        L.ovx_hashmap_insert_3d_idx_as_val.argtypes = [vp, ci, vp, ci, i64, vp, i64, ci, ci, ci, vp]
        i64p, vpp = ctypes.POINTER(i64), ctypes.POINTER(vp)
        L.ovx_seq_encode.argtypes = [vpp, i64p, ci, i64, ci, vp, vp]
        L.ovx_seq_decode.argtypes = [vp, i64, ci, vpp, i64p, vp]
        L.ovx_svo_encode_begin.argtypes = [vp, vp, vp, vp, i64, ci, ci, vp, i64p]
        L.ovx_svo_encode_finish.argtypes = [vp, vp, vp, vp]
        L.ovx_svo_parent_encode.argtypes = [vp, vp, ci, vp, vp]
        L.ovx_svo_neighbor_filter.argtypes = [vp, vp, vp, i64, ci, ci, vp, ci, ci, vp, vp]
        L.ovx_svo_decode_host.argtypes = [vp, i64, ci, vp, i64, i64p]
        L.ovx_svo_parent_decode_host.argtypes = [vp, i64, ci, vp, i64, ci, vp]
        for name in ("ovx_seq_encode", "ovx_seq_decode", "ovx_svo_encode_begin", "ovx_svo_encode_finish",
                     "ovx_svo_parent_encode", "ovx_svo_neighbor_filter", "ovx_svo_decode_host",
                     "ovx_svo_parent_decode_host"):
            getattr(L, name).restype = ctypes.c_int
        for name in ("ovx_dual_mesh_begin", "ovx_dual_mesh_finish", "ovx_hashmap_insert", "ovx_hashmap_lookup",
                     "ovx_hashmap_insert_3d", "ovx_hashmap_lookup_3d", "ovx_hashmap_insert_3d_idx_as_val"):
            getattr(L, name).restype = ctypes.c_int
        for name in ("ovx_fdg_create", "ovx_fdg_destroy", "ovx_fdg_trim", "ovx_fdg_begin", "ovx_fdg_finish",
                     "ovx_occ_begin", "ovx_occ_finish", "ovx_fdg_stats", "ovx_fdg_set_timing",
                     "ovx_fdg_stage_times", "ovx_fdg_kernel_times"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _check(status: int, what: str):
    if status != 0:
        msg = _load().ovx_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (status {status}): {msg}")


class _Ctx:
    """Per-device library context (owns the device workspace)."""

    def __init__(self, device_index: int):
        L = _load()
        h = ctypes.c_void_p()
        _check(L.ovx_fdg_create(device_index, ctypes.byref(h)), "ovx_fdg_create")
        self.h = h
        self.lock = threading.Lock()
        self.svo_keep = None  # inputs of the last octree encode (the context refers to them)

    def __del__(self):
        try:
            if _lib is not None and self.h:
                _lib.ovx_fdg_destroy(self.h)
        except Exception:
            pass


def _ctx(device: torch.device) -> _Ctx:
    idx = device.index if device.index is not None else torch.cuda.current_device()
    c = _ctxs.get(idx)
    if c is None:
        _load()  # takes _lib_lock itself; must not run while _ctx_lock is held by a loader
        with _ctx_lock:
            c = _ctxs.get(idx)
            if c is None:
                c = _Ctx(idx)
                _ctxs[idx] = c
    return c


def _grid(voxel_size: List[float], grid_range: List[int]) -> _Grid:
    if len(voxel_size) != 3 or len(grid_range) != 6:
        raise RuntimeError("voxel_size must have 3 entries and grid_range 6")
    g = _Grid()
    for i in range(3):
        g.voxel_size[i] = float(voxel_size[i])
    for i in range(6):
        g.grid_range[i] = int(grid_range[i])
    return g


def _require_device_tensor(t: torch.Tensor, name: str, dtype: torch.dtype, ndim: int):
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a device tensor")
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must be {dtype}, got {t.dtype}")
    if t.dim() != ndim:
        raise RuntimeError(f"{name} must have {ndim} dimensions, got shape {tuple(t.shape)}")


def _stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


_PRECISIONS = {"reference": 0, "compensated": 1}


def _prepare_mesh(vertices: torch.Tensor, faces: torch.Tensor, origin: torch.Tensor):
    """(vertices - origin as float32, faces as int32) of device tensors, in
    one library kernel on the current stream: the preparation the public
    wrapper does before the extension call (reference
    o_voxel/convert/flexible_dual_grid.py:89-90,181), with no framework
    kernel.  vertices float32 [V,3] and faces int32 / int64 [F,3], contiguous;
    origin float32 [3] on the same device.  int32 faces are returned as they
    are; an int64 index outside [0, 2^31) becomes -1 (rejected by the
    conversion as out of range)."""
    _require_device_tensor(vertices, "vertices", torch.float32, 2)
    if faces.dtype not in (torch.int32, torch.int64):
        raise RuntimeError(f"faces must be int32 or int64, got {faces.dtype}")
    if not faces.is_cuda or faces.device != vertices.device or origin.device != vertices.device:
        raise RuntimeError("vertices, faces and origin must be on one HIP device")
    if origin.dtype != torch.float32 or origin.numel() != 3 or not origin.is_contiguous():
        raise RuntimeError("origin must be 3 contiguous float32 values")
    if not vertices.is_contiguous() or not faces.is_contiguous():
        raise RuntimeError("vertices and faces must be contiguous")
    dev = vertices.device
    L = _load()
    v_out = torch.empty_like(vertices)
    f64 = faces.dtype == torch.int64
    f_out = torch.empty(faces.shape, dtype=torch.int32, device=dev) if f64 else faces
    with torch.cuda.device(dev):
        _check(L.ovx_fdg_prepare_mesh(vertices.data_ptr(), origin.data_ptr(), vertices.size(0),
                                      faces.data_ptr() if f64 else None, faces.size(0) if f64 else 0,
                                      v_out.data_ptr(), f_out.data_ptr() if f64 else None, _stream_handle(dev)),
               "ovx_fdg_prepare_mesh")
    return v_out, f_out


def _batch_prepare(vertices: List[torch.Tensor], faces: List[torch.Tensor], origins=None):
    """Packed inputs of a batched conversion from per-mesh device tensors, in
    one library kernel on the current stream: (vertices - origins[m], stacked,
    float32 [sum V_m, 3]; faces + first vertex row of their mesh, stacked,
    int32 [sum F_m, 3]).  vertices[m] float32 [V_m, 3], faces[m] int32 or
    int64 [F_m, 3], contiguous; origins float32 [M, 3] on the device, or None
    for no shift.  A face index outside [0, V_m) becomes -1 (rejected by the
    conversion as out of range)."""
    M = len(vertices)
    if len(faces) != M:
        raise RuntimeError("vertices and faces must have one entry per mesh")
    if M == 0:
        raise RuntimeError("a batch needs at least one mesh")
    dev = vertices[0].device
    for k in range(M):
        _require_device_tensor(vertices[k], f"vertices[{k}]", torch.float32, 2)
        f = faces[k]
        if not isinstance(f, torch.Tensor) or f.dtype not in (torch.int32, torch.int64) or f.dim() != 2:
            raise RuntimeError(f"faces[{k}] must be a 2-D int32 or int64 tensor")
        if vertices[k].size(1) != 3 or f.size(1) != 3:
            raise RuntimeError("vertices and faces must have 3 columns")
        if vertices[k].device != dev or f.device != dev:
            raise RuntimeError("all meshes must be on one HIP device")
        if not vertices[k].is_contiguous() or not f.is_contiguous():
            raise RuntimeError("vertices and faces must be contiguous")
    nv = [int(v.size(0)) for v in vertices]
    nf = [int(f.size(0)) for f in faces]
    V, F = sum(nv), sum(nf)
    if V > 2 ** 31 - 1:
        raise RuntimeError(f"{V} vertices in the batch: packed face indices are 32-bit")
    if origins is None:
        origins = torch.zeros((M, 3), dtype=torch.float32).pin_memory().to(dev, non_blocking=True)
    if (origins.dtype != torch.float32 or tuple(origins.shape) != (M, 3) or not origins.is_contiguous()
            or origins.device != dev):
        raise RuntimeError("origins must be a contiguous float32 [M, 3] tensor on the meshes' device")
    table = np.zeros((M, 8), dtype=np.int64)
    table[:, 0] = [v.data_ptr() for v in vertices]
    table[:, 1] = [f.data_ptr() for f in faces]
    table[:, 2] = origins.data_ptr() + 12 * np.arange(M, dtype=np.int64)
    table[:, 3] = nv
    table[:, 4] = nf
    table[:, 5] = np.cumsum([0] + nv[:-1])
    table[:, 6] = np.cumsum([0] + nf[:-1])
    table[:, 7] = [1 if f.dtype == torch.int64 else 0 for f in faces]  # faces_int64, then a zero int32
    tab = torch.from_numpy(table).pin_memory().to(dev, non_blocking=True)
    v_all = torch.empty((V, 3), dtype=torch.float32, device=dev)
    f_all = torch.empty((F, 3), dtype=torch.int32, device=dev)
    L = _load()
    with torch.cuda.device(dev):
        _check(L.ovx_fdg_batch_prepare(tab.data_ptr(), M, V, F, v_all.data_ptr(), f_all.data_ptr(), _stream_handle(dev)),
               "ovx_fdg_batch_prepare")
    return v_all, f_all, nv, nf


def mesh_to_flexible_dual_grid_cuda(
    vertices: torch.Tensor,
    faces: torch.Tensor,
    voxel_size: List[float],
    grid_range: List[int],
    face_weight: float,
    boundary_weight: float,
    regularization_weight: float,
    precision: str = "reference",
    cond_max: float = 1.0e6,
) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Device flexible dual grid (reference: src/ext.cpp:19 ->
    src/convert/mesh_to_flexible_dual_grid_gpu/mesh_to_flexible_dual_grid.cu:797-895).

    vertices f32 [V,3] grid-local, faces i32 [F,3], both contiguous on one
    device.  Returns (coords i32 [N,3], dual f32 [N,3], intersected bool [N,3])
    with voxels in ascending x + Gx*(y + Gy*z) order.

    ``precision`` (not in the reference): "reference" reproduces the CPU float
    path bit for bit; "compensated" accumulates each voxel's quadric in the
    voxel's frame with compensated FP32 sums and solves in FP32 with an FP64
    fallback above the condition estimate ``cond_max``.
    """
    mode = _PRECISIONS.get(precision)
    if mode is None:
        raise RuntimeError(f"precision must be one of {sorted(_PRECISIONS)}, got {precision!r}")
    _require_device_tensor(vertices, "vertices", torch.float32, 2)
    _require_device_tensor(faces, "faces", torch.int32, 2)
    if vertices.size(1) != 3 or faces.size(1) != 3:
        raise RuntimeError("vertices and faces must have 3 columns")
    if vertices.device != faces.device:
        raise RuntimeError("vertices and faces must be on the same device")
    vertices = vertices.contiguous()
    faces = faces.contiguous()
    dev = vertices.device
    ctx = _ctx(dev)
    L = _load()
    g = _grid(voxel_size, grid_range)
    with torch.cuda.device(dev), ctx.lock:
        stream = _stream_handle(dev)
        n = ctypes.c_int64(0)
        _check(L.ovx_fdg_set_precision(ctx.h, mode, float(cond_max)), "mesh_to_flexible_dual_grid")
        _check(L.ovx_fdg_begin(ctx.h, vertices.data_ptr(), vertices.size(0), faces.data_ptr(), faces.size(0),
                               ctypes.byref(g), float(face_weight), float(boundary_weight),
                               float(regularization_weight), stream, ctypes.byref(n)), "mesh_to_flexible_dual_grid")
        N = int(n.value)
        coords = torch.empty((N, 3), dtype=torch.int32, device=dev)
        dual = torch.empty((N, 3), dtype=torch.float32, device=dev)
        inter = torch.empty((N, 3), dtype=torch.bool, device=dev)
        _check(L.ovx_fdg_finish(ctx.h, coords.data_ptr(), dual.data_ptr(), inter.data_ptr(), stream),
               "mesh_to_flexible_dual_grid")
    return coords, dual, inter


def mesh_to_flexible_dual_grid_batch_packed_cuda(
    vertices: torch.Tensor,
    faces: torch.Tensor,
    vertex_counts: List[int],
    face_counts: List[int],
    voxel_sizes: List[List[float]],
    grid_ranges: List[List[int]],
    face_weight: float,
    boundary_weight: float,
    regularization_weight: float,
    precision: str = "reference",
    cond_max: float = 1.0e6,
) -> List[Tuple[torch.Tensor, torch.Tensor, torch.Tensor]]:
    """Batched device flexible dual grid on packed inputs: the meshes'
    grid-local f32 vertices one after another [sum V_m, 3] and their i32
    faces [sum F_m, 3] with batch-global vertex indices, plus per-mesh counts
    and grids.  No reference counterpart (the reference converts one mesh per
    call, src/ext.cpp:19); each mesh's outputs equal those of
    mesh_to_flexible_dual_grid_cuda on it alone.  Returns one (coords, dual,
    intersected) tuple per mesh (views of shared output tensors)."""
    mode = _PRECISIONS.get(precision)
    if mode is None:
        raise RuntimeError(f"precision must be one of {sorted(_PRECISIONS)}, got {precision!r}")
    M = len(vertex_counts)
    if M == 0:
        return []
    if len(face_counts) != M or len(voxel_sizes) != M or len(grid_ranges) != M:
        raise RuntimeError("vertex_counts, face_counts, voxel_sizes and grid_ranges must have one entry per mesh")
    _require_device_tensor(vertices, "vertices", torch.float32, 2)
    _require_device_tensor(faces, "faces", torch.int32, 2)
    if vertices.size(1) != 3 or faces.size(1) != 3:
        raise RuntimeError("vertices and faces must have 3 columns")
    if vertices.device != faces.device:
        raise RuntimeError("vertices and faces must be on the same device")
    vo = (ctypes.c_int64 * (M + 1))()
    fo = (ctypes.c_int64 * (M + 1))()
    for k in range(M):
        vo[k + 1] = vo[k] + int(vertex_counts[k])
        fo[k + 1] = fo[k] + int(face_counts[k])
    if vo[M] != vertices.size(0) or fo[M] != faces.size(0):
        raise RuntimeError("vertex / face counts do not add up to the packed arrays")
    vertices = vertices.contiguous()
    faces = faces.contiguous()
    dev = vertices.device
    grids = (_Grid * M)()
    for k in range(M):
        grids[k] = _grid(voxel_sizes[k], grid_ranges[k])
    ctx = _ctx(dev)
    L = _load()
    counts = (ctypes.c_int64 * M)()
    with torch.cuda.device(dev), ctx.lock:
        stream = _stream_handle(dev)
        _check(L.ovx_fdg_set_precision(ctx.h, mode, float(cond_max)), "mesh_to_flexible_dual_grid_batch")
        _check(L.ovx_fdg_batch_begin(ctx.h, vertices.data_ptr(), int(vo[M]), faces.data_ptr(), int(fo[M]), M, fo, vo,
                                     grids, float(face_weight), float(boundary_weight),
                                     float(regularization_weight), stream, counts), "mesh_to_flexible_dual_grid_batch")
        n = [int(counts[k]) for k in range(M)]
        N = sum(n)
        coords = torch.empty((N, 3), dtype=torch.int32, device=dev)
        dual = torch.empty((N, 3), dtype=torch.float32, device=dev)
        inter = torch.empty((N, 3), dtype=torch.bool, device=dev)
        _check(L.ovx_fdg_finish(ctx.h, coords.data_ptr(), dual.data_ptr(), inter.data_ptr(), stream),
               "mesh_to_flexible_dual_grid_batch")
    return list(zip(coords.split(n), dual.split(n), inter.split(n)))


def mesh_to_flexible_dual_grid_batch_cuda(
    vertices: List[torch.Tensor],
    faces: List[torch.Tensor],
    voxel_sizes: List[List[float]],
    grid_ranges: List[List[int]],
    face_weight: float,
    boundary_weight: float,
    regularization_weight: float,
    precision: str = "reference",
    cond_max: float = 1.0e6,
) -> List[Tuple[torch.Tensor, torch.Tensor, torch.Tensor]]:
    """Batched device flexible dual grid on per-mesh tensors (grid-local f32
    [V_m,3] vertices, i32 [F_m,3] faces with per-mesh indices): packs them
    (one concatenation each, one index rebase) and runs
    mesh_to_flexible_dual_grid_batch_packed_cuda."""
    M = len(vertices)
    if M == 0:
        return []
    if len(faces) != M:
        raise RuntimeError("vertices and faces must have one entry per mesh")
    for k in range(M):
        _require_device_tensor(vertices[k], f"vertices[{k}]", torch.float32, 2)
        _require_device_tensor(faces[k], f"faces[{k}]", torch.int32, 2)
    v_all, f_all, nv, nf = _batch_prepare([v.contiguous() for v in vertices], [f.contiguous() for f in faces])
    return mesh_to_flexible_dual_grid_batch_packed_cuda(v_all, f_all, nv, nf, voxel_sizes, grid_ranges, face_weight,
                                                        boundary_weight, regularization_weight, precision=precision,
                                                        cond_max=cond_max)


def intersect_occ_cuda(triangles: torch.Tensor, voxel_size: List[float], grid_range: List[int]) -> torch.Tensor:
    """Device occupancy of a triangle soup (reference: src/ext.cpp:21 ->
    src/convert/mesh_to_flexible_dual_grid_gpu/intersect_qef.cu:896-920).

    triangles f32 [T,3,3] grid-local.  Returns coords i32 [N,3] in ascending
    x + Gx*(y + Gy*z) order.
    """
    _require_device_tensor(triangles, "triangles", torch.float32, 3)
    if triangles.size(1) != 3 or triangles.size(2) != 3:
        raise RuntimeError("triangles must have shape [T,3,3]")
    triangles = triangles.contiguous()
    dev = triangles.device
    ctx = _ctx(dev)
    L = _load()
    g = _grid(voxel_size, grid_range)
    with torch.cuda.device(dev), ctx.lock:
        stream = _stream_handle(dev)
        n = ctypes.c_int64(0)
        _check(L.ovx_occ_begin(ctx.h, triangles.data_ptr(), triangles.size(0), ctypes.byref(g), stream,
                               ctypes.byref(n)), "intersect_occ")
        coords = torch.empty((int(n.value), 3), dtype=torch.int32, device=dev)
        _check(L.ovx_occ_finish(ctx.h, coords.data_ptr(), stream), "intersect_occ")
    return coords


_SPLIT_TYPES = {torch.float32: 0, torch.float16: 1, torch.bfloat16: 2}


def flexible_dual_grid_to_mesh_cuda(
    coords: torch.Tensor,
    dual_vertices: torch.Tensor,
    intersected_flag: torch.Tensor,
    split_weight,
    grid_size: List[int],
    voxel_size: List[float],
    aabb_min: List[float],
    faces: bool = True,
    quads: bool = False,
):
    """Device dual grid -> mesh (the body of the reference's
    flexible_dual_grid_to_mesh, o_voxel/convert/flexible_dual_grid.py:422-454,
    as one library call instead of its hashmap calls and torch glue).

    coords i32 [N,3], dual_vertices f32/f16/bf16 [N,3] in voxel units,
    intersected_flag bool/uint8 [N,3], split_weight None or f32/f16/bf16 with
    N values ([N,1] as the reference uses it, or [N]); grid_size (W, H, D),
    voxel_size and aabb_min are 3-lists of the float32 values the reference
    computes.  Returns (vertices f32 [N,3] = (coords + dual) * voxel_size +
    aabb_min, faces i32 [2L,3] or None, quads i32 [L,4] or None), quads in
    (voxel, axis) order.
    """
    _require_device_tensor(coords, "coords", torch.int32, 2)
    if not isinstance(dual_vertices, torch.Tensor) or dual_vertices.device != coords.device:
        raise RuntimeError("dual_vertices must be a tensor on the device of coords")
    if dual_vertices.dtype not in _SPLIT_TYPES:
        raise RuntimeError(f"dual_vertices must be float32, float16 or bfloat16, got {dual_vertices.dtype}")
    if not isinstance(intersected_flag, torch.Tensor) or intersected_flag.device != coords.device:
        raise RuntimeError("intersected_flag must be a tensor on the device of coords")
    if intersected_flag.dtype not in (torch.bool, torch.uint8):
        raise RuntimeError(f"intersected_flag must be bool or uint8, got {intersected_flag.dtype}")
    N = coords.size(0)
    for name, t in (("coords", coords), ("dual_vertices", dual_vertices), ("intersected_flag", intersected_flag)):
        if t.dim() != 2 or t.size(0) != N or t.size(1) != 3:
            raise RuntimeError(f"{name} must have shape [N, 3] with N = {N}, got {tuple(t.shape)}")
    wtype = 0
    if split_weight is not None:
        if not isinstance(split_weight, torch.Tensor) or split_weight.device != coords.device:
            raise RuntimeError("split_weight must be a tensor on the device of coords")
        if split_weight.dtype not in _SPLIT_TYPES:
            raise RuntimeError(f"split_weight must be float32, float16 or bfloat16, got {split_weight.dtype}")
        if split_weight.numel() != N:
            raise RuntimeError(f"split_weight must hold one value per voxel ({N}), got shape {tuple(split_weight.shape)}")
        wtype = _SPLIT_TYPES[split_weight.dtype]
        # float16 / bfloat16 -> float32 is exact; the kernel rounds products back
        split_weight = split_weight.detach().reshape(N).to(torch.float32).contiguous()
    dev = coords.device
    coords = coords.contiguous()
    dual = dual_vertices.detach().to(torch.float32).contiguous()
    flags = intersected_flag.contiguous().view(torch.uint8)
    gs = (ctypes.c_int32 * 3)(*[int(x) for x in grid_size])
    vs = (ctypes.c_float * 3)(*[float(x) for x in voxel_size])
    a0 = (ctypes.c_float * 3)(*[float(x) for x in aabb_min])
    ctx = _ctx(dev)
    L = _load()
    with torch.cuda.device(dev), ctx.lock:
        stream = _stream_handle(dev)
        verts = torch.empty((N, 3), dtype=torch.float32, device=dev)
        nq = ctypes.c_int64(0)
        _check(L.ovx_dual_mesh_begin(ctx.h, coords.data_ptr(), dual.data_ptr(), flags.data_ptr(), N, gs, vs, a0,
                                     verts.data_ptr(), stream, ctypes.byref(nq)), "flexible_dual_grid_to_mesh")
        nq = int(nq.value)
        f = torch.empty((2 * nq, 3), dtype=torch.int32, device=dev) if faces else None
        q = torch.empty((nq, 4), dtype=torch.int32, device=dev) if quads else None
        _check(L.ovx_dual_mesh_finish(ctx.h, split_weight.data_ptr() if split_weight is not None else None, wtype,
                                      f.data_ptr() if f is not None else None,
                                      q.data_ptr() if q is not None else None, stream),
               "flexible_dual_grid_to_mesh")
    return verts, f, q


# ---------------------------------------------------------------------------
# generic hashmap (reference: src/hash/api.h:27-110, src/hash/hash.cu, bound
# src/ext.cpp:12-16).  Tables are caller-owned uint32 / uint64 tensors; the
# keys tensor starts filled with all ones (empty).
# ---------------------------------------------------------------------------
_HASH_BYTES = {torch.uint32: 4, torch.uint64: 8}


def _hash_table(hashmap_keys: torch.Tensor, hashmap_values: torch.Tensor):
    for name, t in (("hashmap_keys", hashmap_keys), ("hashmap_values", hashmap_values)):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise RuntimeError(f"{name} must be a device tensor")
        if not t.is_contiguous():
            raise RuntimeError(f"{name} must be contiguous")
    kb, vb = _HASH_BYTES.get(hashmap_keys.dtype), _HASH_BYTES.get(hashmap_values.dtype)
    if kb is None or vb is None:
        raise RuntimeError("Unsupported data type")
    if hashmap_keys.device != hashmap_values.device:
        raise RuntimeError("hashmap_keys and hashmap_values must be on the same device")
    if hashmap_values.numel() < hashmap_keys.numel():
        raise RuntimeError("hashmap_values must have as many entries as hashmap_keys")
    return kb, vb


def _hash_input(t: torch.Tensor, name: str, dtype: torch.dtype, msg: str, dev: torch.device, cols: int = 0):
    if not isinstance(t, torch.Tensor) or t.dtype != dtype:
        raise RuntimeError(msg)
    if t.device != dev:
        raise RuntimeError(f"{name} must be on the device of the hashmap")
    if cols and (t.dim() != 2 or t.size(1) != cols):
        raise RuntimeError(f"{name} must have shape [M, {cols}], got {tuple(t.shape)}")
    return t.contiguous()


def _check_rows(keys, values, keys_name: str, cols: int = 0):
    """Shape checks made before any device work: 1-D keys (or [M, cols]
    coordinates) and at least one value per key row, so that the kernels
    never read past the end of a short values tensor."""
    if not isinstance(keys, torch.Tensor):
        raise RuntimeError(f"{keys_name} must be a torch.Tensor")
    if cols:
        if keys.dim() != 2 or keys.size(1) != cols:
            raise RuntimeError(f"{keys_name} must have shape [M, {cols}], got {tuple(keys.shape)}")
    elif keys.dim() != 1:
        raise RuntimeError(f"{keys_name} must be a 1-D tensor, got shape {tuple(keys.shape)}")
    if values is not None:
        if not isinstance(values, torch.Tensor):
            raise RuntimeError("values must be a torch.Tensor")
        if values.numel() < keys.size(0):
            raise RuntimeError(f"values has {values.numel()} entries for {keys.size(0)} keys")


def _type_name(nbytes: int) -> str:
    return "uint32" if nbytes == 4 else "uint64"


def hashmap_insert_cuda(hashmap_keys: torch.Tensor, hashmap_values: torch.Tensor, keys: torch.Tensor,
                        values: torch.Tensor) -> None:
    """Insert keys[M] with values[M] (reference src/hash/hash.cu hashmap_insert_cuda)."""
    _check_rows(keys, values, "keys")
    kb, vb = _hash_table(hashmap_keys, hashmap_values)
    dev = hashmap_keys.device
    keys = _hash_input(keys, "keys", hashmap_keys.dtype, f"Keys must be {_type_name(kb)}", dev)
    values = _hash_input(values, "values", hashmap_values.dtype, f"Values must be {_type_name(vb)}", dev)
    M = keys.size(0)
    with torch.cuda.device(dev):
        _check(_load().ovx_hashmap_insert(hashmap_keys.data_ptr(), kb, hashmap_values.data_ptr(), vb,
                                          hashmap_keys.numel(), keys.data_ptr(), values.data_ptr(), M,
                                          _stream_handle(dev)), "hashmap_insert")


def hashmap_lookup_cuda(hashmap_keys: torch.Tensor, hashmap_values: torch.Tensor, keys: torch.Tensor) -> torch.Tensor:
    """Values of keys[M], all ones where missing (reference hashmap_lookup_cuda)."""
    _check_rows(keys, None, "keys")
    kb, vb = _hash_table(hashmap_keys, hashmap_values)
    dev = hashmap_keys.device
    keys = _hash_input(keys, "keys", hashmap_keys.dtype, f"Keys must be {_type_name(kb)}", dev)
    M = keys.size(0)
    out = torch.empty((M,), dtype=hashmap_values.dtype, device=dev)
    with torch.cuda.device(dev):
        _check(_load().ovx_hashmap_lookup(hashmap_keys.data_ptr(), kb, hashmap_values.data_ptr(), vb,
                                          hashmap_keys.numel(), keys.data_ptr(), out.data_ptr(), M,
                                          _stream_handle(dev)), "hashmap_lookup")
    return out


def hashmap_insert_3d_cuda(hashmap_keys: torch.Tensor, hashmap_values: torch.Tensor, coords: torch.Tensor,
                           values: torch.Tensor, W: int, H: int, D: int) -> None:
    """Insert int32 coords[M,4] = (b, x, y, z) under key b*W*H*D + x*H*D + y*D + z
    (reference hashmap_insert_3d_cuda)."""
    _check_rows(coords, values, "coords", cols=4)
    kb, vb = _hash_table(hashmap_keys, hashmap_values)
    dev = hashmap_keys.device
    coords = _hash_input(coords, "coords", torch.int32, "Coords must be int32", dev, cols=4)
    values = _hash_input(values, "values", hashmap_values.dtype, f"Values must be {_type_name(vb)}", dev)
    with torch.cuda.device(dev):
        _check(_load().ovx_hashmap_insert_3d(hashmap_keys.data_ptr(), kb, hashmap_values.data_ptr(), vb,
                                             hashmap_keys.numel(), coords.data_ptr(), values.data_ptr(),
                                             coords.size(0), int(W), int(H), int(D), _stream_handle(dev)),
               "hashmap_insert_3d")


def hashmap_lookup_3d_cuda(hashmap_keys: torch.Tensor, hashmap_values: torch.Tensor, coords: torch.Tensor, W: int,
                           H: int, D: int) -> torch.Tensor:
    """Values of int32 coords[M,4]; all ones where missing or where (x, y, z)
    lies outside [0,W) x [0,H) x [0,D) (reference hashmap_lookup_3d_cuda)."""
    _check_rows(coords, None, "coords", cols=4)
    kb, vb = _hash_table(hashmap_keys, hashmap_values)
    dev = hashmap_keys.device
    coords = _hash_input(coords, "coords", torch.int32, "Coords must be int32", dev, cols=4)
    out = torch.empty((coords.size(0),), dtype=hashmap_values.dtype, device=dev)
    with torch.cuda.device(dev):
        _check(_load().ovx_hashmap_lookup_3d(hashmap_keys.data_ptr(), kb, hashmap_values.data_ptr(), vb,
                                             hashmap_keys.numel(), coords.data_ptr(), out.data_ptr(),
                                             coords.size(0), int(W), int(H), int(D), _stream_handle(dev)),
               "hashmap_lookup_3d")
    return out


def hashmap_insert_3d_idx_as_val_cuda(hashmap_keys: torch.Tensor, hashmap_values: torch.Tensor, coords: torch.Tensor,
                                      W: int, H: int, D: int) -> None:
    """Insert int32 coords[M,4] with each row's index as its value
    (reference hashmap_insert_3d_idx_as_val_cuda)."""
    _check_rows(coords, None, "coords", cols=4)
    kb, vb = _hash_table(hashmap_keys, hashmap_values)
    dev = hashmap_keys.device
    coords = _hash_input(coords, "coords", torch.int32, "Coords must be int32", dev, cols=4)
    with torch.cuda.device(dev):
        _check(_load().ovx_hashmap_insert_3d_idx_as_val(hashmap_keys.data_ptr(), kb, hashmap_values.data_ptr(), vb,
                                                        hashmap_keys.numel(), coords.data_ptr(), coords.size(0),
                                                        int(W), int(H), int(D), _stream_handle(dev)),
               "hashmap_insert_3d_idx_as_val")


# ---------------------------------------------------------------------------
# 30-bit voxel codes (reference: src/serialize/api.cu, bound src/ext.cpp:24-31)
# ---------------------------------------------------------------------------
_SEQ_MODES = {"z_order": 0, "hilbert": 1}


def seq_encode_columns(cols, mode: str) -> torch.Tensor:
    """int32 codes [N] of three coordinate columns (1-D int32 / int64 device
    tensors of one length, any strides; read in place)."""
    dev = cols[0].device
    n = cols[0].numel()
    dt = cols[0].dtype
    for c in cols:
        if not isinstance(c, torch.Tensor) or not c.is_cuda or c.device != dev:
            raise RuntimeError("coordinate columns must be device tensors on one device")
        if c.dim() != 1 or c.numel() != n or c.dtype != dt or dt not in (torch.int32, torch.int64):
            raise RuntimeError("coordinate columns must be 1-D int32 or int64 tensors of one length")
    codes = torch.empty((n,), dtype=torch.int32, device=dev)
    ptrs = (ctypes.c_void_p * 3)(*[c.data_ptr() for c in cols])
    strides = (ctypes.c_int64 * 3)(*[c.stride(0) for c in cols])
    with torch.cuda.device(dev):
        _check(_load().ovx_seq_encode(ptrs, strides, cols[0].element_size(), n, _SEQ_MODES[mode], codes.data_ptr(),
                                      _stream_handle(dev)), f"{mode}_encode")
    return codes


def seq_decode_into(codes: torch.Tensor, mode: str, out_cols) -> None:
    """Write the three int32 coordinate columns of int32 codes [N] into
    ``out_cols`` (1-D int32 device tensor views of length N, any strides)."""
    dev = codes.device
    ptrs = (ctypes.c_void_p * 3)(*[c.data_ptr() for c in out_cols])
    strides = (ctypes.c_int64 * 3)(*[c.stride(0) for c in out_cols])
    with torch.cuda.device(dev):
        _check(_load().ovx_seq_decode(codes.data_ptr(), codes.numel(), _SEQ_MODES[mode], ptrs, strides,
                                      _stream_handle(dev)), f"{mode}_decode")


def _seq_codes_in(codes: torch.Tensor) -> torch.Tensor:
    if not isinstance(codes, torch.Tensor) or not codes.is_cuda:
        raise RuntimeError("codes must be a device tensor")
    if codes.dtype not in (torch.int32, torch.int64):
        raise RuntimeError(f"codes must be int32, got {codes.dtype}")
    return codes.reshape(-1).to(torch.int32).contiguous()


def _seq_decode(codes: torch.Tensor, mode: str):
    codes = _seq_codes_in(codes)
    out = torch.empty((3, codes.numel()), dtype=torch.int32, device=codes.device)
    seq_decode_into(codes, mode, (out[0], out[1], out[2]))
    return out[0], out[1], out[2]


def z_order_encode_cuda(x: torch.Tensor, y: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
    """Morton codes of (x, y, z) (reference src/serialize/api.cu z_order_encode_cuda)."""
    return seq_encode_columns((x, y, z), "z_order")


def z_order_decode_cuda(codes: torch.Tensor):
    """(x, y, z) int32 of Morton codes (reference z_order_decode_cuda)."""
    return _seq_decode(codes, "z_order")


def hilbert_encode_cuda(x: torch.Tensor, y: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
    """Hilbert codes of (x, y, z) (reference hilbert_encode_cuda)."""
    return seq_encode_columns((x, y, z), "hilbert")


def hilbert_decode_cuda(codes: torch.Tensor):
    """(x, y, z) int32 of Hilbert codes (reference hilbert_decode_cuda)."""
    return _seq_decode(codes, "hilbert")


def _seq_host(fn, *ts):
    dev = _gpu_for_host_call()
    out = fn(*[t.to(dev) for t in ts])
    return tuple(t.cpu() for t in out) if isinstance(out, tuple) else out.cpu()


def z_order_encode_cpu(x, y, z):
    """Host-tensor twin (reference src/ext.cpp:24), run on the current GPU."""
    return _seq_host(z_order_encode_cuda, x, y, z)


def z_order_decode_cpu(codes):
    return _seq_host(z_order_decode_cuda, codes)


def hilbert_encode_cpu(x, y, z):
    return _seq_host(hilbert_encode_cuda, x, y, z)


def hilbert_decode_cpu(codes):
    return _seq_host(hilbert_decode_cuda, codes)


# ---------------------------------------------------------------------------
# sparse voxel octrees of the .vxz format (reference: src/io/svo.cpp,
# src/io/filter_parent.cpp, src/io/filter_neighbor.cpp, src/io/api.h)
# ---------------------------------------------------------------------------
def svo_encode_chunks(codes: torch.Tensor, chunk: torch.Tensor, chunk_off: torch.Tensor, n_chunks: int, depth: int):
    """Octrees of all chunks at once on the device.  codes int32 [n] (Morton
    codes of chunk-local coordinates, ascending inside each chunk), chunk
    int32 [n], chunk_off int64 [n_chunks + 1] device tensors.  Returns
    (octree bytes uint8 [total] on the device, per-chunk byte offsets as a
    host int64 tensor [n_chunks + 1]).  The context keeps the codes'
    octree state for svo_parent_encode."""
    dev = codes.device
    n = codes.numel()
    codes, chunk, chunk_off = codes.contiguous(), chunk.contiguous(), chunk_off.contiguous()
    ctx = _ctx(dev)
    L = _load()
    with torch.cuda.device(dev), ctx.lock:
        stream = _stream_handle(dev)
        nb = ctypes.c_int64(0)
        _check(L.ovx_svo_encode_begin(ctx.h, codes.data_ptr(), chunk.data_ptr(), chunk_off.data_ptr(), n,
                                      int(n_chunks), int(depth), stream, ctypes.byref(nb)), "svo_encode")
        total = int(nb.value)
        buf = torch.empty(((total + 3) // 4) * 4, dtype=torch.uint8, device=dev)
        offs = torch.empty((int(n_chunks) + 1,), dtype=torch.int64, device=dev)
        _check(L.ovx_svo_encode_finish(ctx.h, buf.data_ptr() if total else None,
                                       offs.data_ptr() if n_chunks else None, stream), "svo_encode")
        ctx.svo_keep = (codes, chunk, chunk_off, buf)  # the context's octree state refers to these
    return buf[:total], (offs.cpu() if n_chunks else torch.zeros((1,), dtype=torch.int64))


def svo_parent_encode(attr: torch.Tensor) -> torch.Tensor:
    """Parent-filter rows (uint8 [n, C]) of the codes of the last
    svo_encode_chunks on attr's device, attr uint8 [n, C] in code order."""
    dev = attr.device
    attr = attr.contiguous()
    ctx = _ctx(dev)
    out = torch.empty_like(attr)
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_svo_parent_encode(ctx.h, attr.data_ptr(), attr.shape[1], out.data_ptr(),
                                             _stream_handle(dev)), "svo_parent_encode")
    return out


def svo_neighbor_filter(coords: torch.Tensor, chunk: torch.Tensor, n_chunks: int, res: int, data: torch.Tensor,
                        decode: bool) -> torch.Tensor:
    """Neighbour filter (or its inverse) of uint8 [n, C] data over chunk-local
    int32 [n, 3] coordinates in [0, res) and int32 chunk ids, on the device."""
    dev = coords.device
    coords, chunk, data = coords.contiguous(), chunk.contiguous(), data.contiguous()
    out = torch.empty_like(data)
    if data.numel() == 0:
        return out
    ctx = _ctx(dev)
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_svo_neighbor_filter(ctx.h, coords.data_ptr(), chunk.data_ptr(), coords.shape[0],
                                               int(n_chunks), int(res), data.data_ptr(), data.shape[1],
                                               int(bool(decode)), out.data_ptr(), _stream_handle(dev)),
               "svo_neighbor_filter")
    return out


def svo_decode_host(octree, depth: int):
    """Ascending codes (numpy int32) of one chunk's octree bytes (host parse)."""
    import numpy as np

    o = np.ascontiguousarray(np.asarray(octree, dtype=np.uint8))
    cap = 8 * max(1, o.size)
    codes = np.empty(cap, np.int32)
    n = ctypes.c_int64(0)
    _check(_load().ovx_svo_decode_host(o.ctypes.data, o.size, int(depth), codes.ctypes.data, cap, ctypes.byref(n)),
           "svo_decode")
    return codes[:int(n.value)]


def svo_parent_decode_host(octree, depth: int, delta):
    """Attributes (numpy uint8 [n, C]) of one chunk from its octree and
    parent-filter rows (host parse)."""
    import numpy as np

    o = np.ascontiguousarray(np.asarray(octree, dtype=np.uint8))
    d = np.ascontiguousarray(np.asarray(delta, dtype=np.uint8))
    out = np.empty_like(d)
    _check(_load().ovx_svo_parent_decode_host(o.ctypes.data, o.size, int(depth), d.ctypes.data, d.shape[0],
                                              d.shape[1], out.ctypes.data), "svo_parent_decode")
    return out


def _single_chunk(n: int, dev):
    return (torch.zeros((n,), dtype=torch.int32, device=dev),
            torch.tensor([0, n], dtype=torch.int64, device=dev))


def _check_codes(codes: torch.Tensor, depth: int):
    if not 1 <= int(depth) <= 10:
        raise RuntimeError(f"depth must be in [1, 10], got {depth}")
    if codes.numel():
        c = codes.long()
        if bool((c < 0).any()) or bool((c >= (1 << (3 * int(depth)))).any()):
            raise RuntimeError(f"codes must lie in [0, 2^{3 * int(depth)})")
        if bool((c[1:] < c[:-1]).any()):
            raise RuntimeError("codes must be sorted in ascending order")


def encode_sparse_voxel_octree_cpu(codes: torch.Tensor, depth: int) -> torch.Tensor:
    """Octree bytes of ascending Morton codes (reference src/io/svo.cpp:17-73);
    computed on the current GPU."""
    dev = _gpu_for_host_call()
    c = codes.to(device=dev, dtype=torch.int32).reshape(-1)
    _check_codes(c, depth)
    if c.numel() == 0:
        return torch.zeros((1,), dtype=torch.uint8)  # the root alone, as the reference writes it
    chunk, off = _single_chunk(c.numel(), dev)
    svo, _ = svo_encode_chunks(c, chunk, off, 1, int(depth))
    return svo.cpu()


def decode_sparse_voxel_octree_cpu(octree: torch.Tensor, depth: int) -> torch.Tensor:
    """Ascending codes of an octree (reference src/io/svo.cpp:76-137)."""
    return torch.from_numpy(svo_decode_host(octree.cpu().numpy(), depth).copy())


def encode_sparse_voxel_octree_attr_parent_cpu(octree: torch.Tensor, depth: int, attr: torch.Tensor) -> torch.Tensor:
    """Parent-filter rows (reference src/io/filter_parent.cpp:10-87)."""
    dev = _gpu_for_host_call()
    codes = torch.from_numpy(svo_decode_host(octree.cpu().numpy(), depth).copy()).to(dev)
    a = attr.to(device=dev, dtype=torch.uint8)
    if a.shape[0] != codes.numel():
        raise RuntimeError(f"attr has {a.shape[0]} rows for {codes.numel()} octree leaves")
    chunk, off = _single_chunk(codes.numel(), dev)
    svo_encode_chunks(codes, chunk, off, 1, int(depth))
    return svo_parent_encode(a).cpu()


def decode_sparse_voxel_octree_attr_parent_cpu(octree: torch.Tensor, depth: int, delta: torch.Tensor) -> torch.Tensor:
    """Inverse of the parent filter (reference src/io/filter_parent.cpp:90-164)."""
    return torch.from_numpy(svo_parent_decode_host(octree.cpu().numpy(), depth, delta.cpu().numpy()))


def _neighbor_cpu(coord, res, data, decode):
    dev = _gpu_for_host_call()
    c = coord.to(device=dev, dtype=torch.int32).reshape(-1, 3)
    chunk, _ = _single_chunk(c.shape[0], dev)
    return svo_neighbor_filter(c, chunk, 1, int(res), data.to(device=dev, dtype=torch.uint8), decode).cpu()


def encode_sparse_voxel_octree_attr_neighbor_cpu(coord: torch.Tensor, res: int, attr: torch.Tensor) -> torch.Tensor:
    """Neighbour-filter rows (reference src/io/filter_neighbor.cpp:18-99)."""
    return _neighbor_cpu(coord, res, attr, False)


def decode_sparse_voxel_octree_attr_neighbor_cpu(coord: torch.Tensor, res: int, delta: torch.Tensor) -> torch.Tensor:
    """Inverse of the neighbour filter (reference src/io/filter_neighbor.cpp:110-178)."""
    return _neighbor_cpu(coord, res, delta, True)


def _gpu_for_host_call() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError(
            "o_voxel (MI355X build) has no CPU implementation: CPU tensors are converted on the GPU, "
            "and no GPU is available"
        )
    return torch.device("cuda", torch.cuda.current_device())


def mesh_to_flexible_dual_grid_cpu(
    vertices: torch.Tensor,
    faces: torch.Tensor,
    voxel_size: List[float],
    grid_range: List[int],
    face_weight: float,
    boundary_weight: float,
    regularization_weight: float,
    timing: bool,
    precision: str = "reference",
) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Host-tensor entry point (reference: src/ext.cpp:18 ->
    src/convert/flexible_dual_grid.cpp:660-992), run on the current GPU.

    Same outputs as the device path, returned as CPU tensors.  With
    ``timing`` the per-stage device times are printed.
    """
    dev = _gpu_for_host_call()
    v = vertices.to(device=dev, dtype=torch.float32).contiguous()
    f = faces.to(device=dev, dtype=torch.int32).contiguous()
    if timing:
        set_timing(dev, True)
    out = mesh_to_flexible_dual_grid_cuda(v, f, voxel_size, grid_range, face_weight, boundary_weight,
                                          regularization_weight, precision=precision)
    if timing:
        names = ("Intersect QEF computation", "Face QEF computation", "Boundary QEF computation",
                 "Dual vertices computation")
        for k, ms in zip(names, stage_times(dev)):
            print(f"{k} took {ms / 1e3} seconds.")
    return tuple(t.cpu() for t in out)


def intersect_occ_cpu(triangles: torch.Tensor, voxel_size: List[float], grid_range: List[int]) -> torch.Tensor:
    """Host-tensor occupancy (reference: src/ext.cpp:20 -> flexible_dual_grid.cpp:605-644), run on the GPU."""
    dev = _gpu_for_host_call()
    t = triangles.to(device=dev, dtype=torch.float32).contiguous()
    return intersect_occ_cuda(t, voxel_size, grid_range).cpu()


# ---------------------------------------------------------------------------
# instrumentation
# ---------------------------------------------------------------------------
def last_stats(device=None) -> dict:
    """Counts of the last conversion on ``device`` (synchronises its stream)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    keys = ("voxels", "events", "records", "face_records", "boundary_edges", "boundary_records",
            "workspace_bytes", "giant_voxels", "active_bricks", "chunks", "large_bricks", "large_brick_records",
            "box_voxels", "fp64_voxels", "record_attempts")
    out = (ctypes.c_int64 * len(keys))()
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_fdg_stats(ctx.h, out, len(keys)), "ovx_fdg_stats")
    return {k: int(v) for k, v in zip(keys, out)}


def set_timing(device=None, enable: bool = True):
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_fdg_set_timing(ctx.h, int(bool(enable))), "ovx_fdg_set_timing")


def stage_times(device=None) -> List[float]:
    """Device ms of the reference's timed stages [intersection QEF, face QEF,
    boundary QEF, dual vertices] (FDG.cpp:711-986) of the last conversion."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    out = (ctypes.c_float * 4)()
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_fdg_stage_times(ctx.h, out), "ovx_fdg_stage_times")
    return [float(x) for x in out]


def kernel_times(device=None) -> dict:
    """Device ms per kernel slot of the last conversion (timing must be enabled)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    L = _load()
    n = L.ovx_fdg_kernel_count()
    out = (ctypes.c_float * n)()
    with torch.cuda.device(dev), ctx.lock:
        _check(L.ovx_fdg_kernel_times(ctx.h, out, n), "ovx_fdg_kernel_times")
    return {L.ovx_fdg_kernel_name(k).decode(): float(out[k]) for k in range(n) if out[k] >= 0.0}


def set_tile_classes(device=None, mark_fpc=None, face_vpf=None):
    """Launch-class thresholds for the next conversions on ``device`` (no
    reference counterpart): three scanline thresholds on faces per grid
    cross-section and two face-test thresholds on voxels per face, each
    non-decreasing (ovx_fdg_set_tile_classes); None keeps a set."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    if (mark_fpc is not None and len(mark_fpc) != 3) or (face_vpf is not None and len(face_vpf) != 2):
        raise RuntimeError("set_tile_classes: mark_fpc needs 3 thresholds and face_vpf 2")
    m = None if mark_fpc is None else (ctypes.c_double * 3)(*[float(x) for x in mark_fpc])
    f = None if face_vpf is None else (ctypes.c_double * 2)(*[float(x) for x in face_vpf])
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_fdg_set_tile_classes(ctx.h, m, f), "set_tile_classes")


def get_tile_classes(device=None) -> Tuple[List[float], List[float]]:
    """(scanline thresholds [3], face-test thresholds [2]) of ``device``'s context."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    m = (ctypes.c_double * 3)()
    f = (ctypes.c_double * 2)()
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_fdg_get_tile_classes(ctx.h, m, f), "get_tile_classes")
    return [float(x) for x in m], [float(x) for x in f]


def set_brick_class(device=None, faces_per_voxel: float = 3.0):
    """Meshes with at least ``faces_per_voxel`` faces per occupied voxel group
    and fold their records in 4x4x4 bricks, others in 8x8x8 ones, for the next
    conversions on ``device`` (ovx_fdg_set_brick_class; results do not change)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_fdg_set_brick_class(ctx.h, float(faces_per_voxel)), "set_brick_class")


def get_brick_class(device=None) -> float:
    """The brick-size threshold (faces per voxel) of ``device``'s context."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    out = ctypes.c_double()
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_fdg_get_brick_class(ctx.h, ctypes.byref(out)), "get_brick_class")
    return float(out.value)


def set_graph(device=None, enable: bool = True):
    """Graph mode of ``device``'s context (ovx_fdg_set_graph): each
    conversion's device work after the voxel set is captured as a HIP graph
    the first time its plan occurs and replayed for conversions of the same
    plan.  Outputs do not change."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_fdg_set_graph(ctx.h, int(bool(enable))), "set_graph")


def graph_stats(device=None) -> dict:
    """{"replays": graph launches of a kept graph, "captures": graphs captured} on ``device``."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    r, c = ctypes.c_int64(0), ctypes.c_int64(0)
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_fdg_graph_stats(ctx.h, ctypes.byref(r), ctypes.byref(c)), "graph_stats")
    return {"replays": int(r.value), "captures": int(c.value)}


def set_capacity(device=None, events: int = 0, hits: int = 0):
    """Capacity of the crossing / face-hit buffers for the next conversion (0 keeps)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_fdg_set_capacity(ctx.h, int(events), int(hits)), "ovx_fdg_set_capacity")


def workspace_bytes(device=None) -> int:
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    return int(_load().ovx_fdg_workspace_bytes(_ctx(dev).h))
