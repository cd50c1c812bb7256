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
    "workspace_bytes",
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
        L.ovx_fdg_begin.argtypes = [vp, vp, i64, vp, i64, ctypes.POINTER(_Grid), f32, f32, f32, vp,
                                    ctypes.POINTER(i64)]
        L.ovx_fdg_finish.argtypes = [vp, vp, vp, vp, vp]
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
        L.ovx_fdg_set_capacity.restype = ctypes.c_int
        L.ovx_fdg_kernel_count.restype = ctypes.c_int
        L.ovx_fdg_kernel_name.argtypes = [ctypes.c_int]
        L.ovx_fdg_kernel_name.restype = ctypes.c_char_p
        L.ovx_fdg_kernel_times.argtypes = [vp, ctypes.POINTER(f32), ctypes.c_int]
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
    dev = vertices[0].device
    nv = [int(v.size(0)) for v in vertices]
    nf = [int(f.size(0)) for f in faces]
    v_all = torch.cat(vertices, 0)
    f_all = torch.cat(faces, 0)
    base = torch.tensor([0] + nv[:-1], dtype=torch.int64).cumsum(0).to(torch.int32).to(dev)
    f_all = f_all + torch.repeat_interleave(base, torch.tensor(nf, device=dev), output_size=sum(nf)).unsqueeze(1)
    return mesh_to_flexible_dual_grid_batch_packed_cuda(v_all, f_all.to(torch.int32), nv, nf, voxel_sizes, grid_ranges,
                                                        face_weight, boundary_weight, regularization_weight,
                                                        precision=precision, cond_max=cond_max)


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
            "box_voxels", "fp64_voxels")
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


def set_capacity(device=None, events: int = 0, hits: int = 0):
    """Capacity of the crossing / face-hit buffers for the next conversion (0 keeps)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    ctx = _ctx(dev)
    with torch.cuda.device(dev), ctx.lock:
        _check(_load().ovx_fdg_set_capacity(ctx.h, int(events), int(hits)), "ovx_fdg_set_capacity")


def workspace_bytes(device=None) -> int:
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    return int(_load().ovx_fdg_workspace_bytes(_ctx(dev).h))
