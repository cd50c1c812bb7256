"""Public flexible-dual-grid conversion API (MI355X build).

Same signatures, grid set-up and error behaviour as the reference module
``o_voxel/convert/flexible_dual_grid.py``; the work is done by the HIP
library behind ``o_voxel._C``.
"""

from typing import Any, List, Optional, Tuple, Union

import numpy as np
import torch

from .. import _C

FloatArrayLike = Union[float, List[Any], Tuple[Any, ...], np.ndarray, torch.Tensor]
IntArrayLike = Union[int, List[Any], Tuple[Any, ...], np.ndarray, torch.Tensor]
ArrayLike = Union[List[Any], Tuple[Any, ...], np.ndarray, torch.Tensor]

__all__ = [
    "mesh_to_flexible_dual_grid",
    "mesh_to_flexible_dual_grid_batch",
    "autotune_tile_classes",
    "intersect_occ",
    "flexible_dual_grid_to_mesh",
]


def _host_vector(value, name: str, dtype: torch.dtype, kind: str) -> torch.Tensor:
    """Scalar / sequence / ndarray / CPU tensor -> CPU tensor of 3 entries
    (reference: flexible_dual_grid.py:95-127)."""
    scalar_type = float if dtype == torch.float32 else int
    if isinstance(value, scalar_type):
        value = [value, value, value]
    if isinstance(value, (list, tuple)):
        value = np.array(value)
    if isinstance(value, np.ndarray):
        t = torch.tensor(value, dtype=dtype)
    else:
        assert isinstance(value, torch.Tensor), (
            f"{name} must be a {kind}, list, tuple, np.ndarray, or torch.Tensor, but got {type(value)}"
        )
        assert not value.is_cuda, (
            f"{name} Tensor metadata must be on CPU; pass Python values to avoid implicit CUDA sync"
        )
        t = value.to(dtype=dtype).contiguous()
    assert t.dim() == 1, f"{name} must be a 1D tensor, but got {t.shape}"
    assert t.size(0) == 3, f"{name} must have 3 elements, but got {t.size(0)}"
    return t


def _aabb_host(aabb) -> torch.Tensor:
    """Explicit aabb -> CPU float32 [2, 3] tensor (reference: flexible_dual_grid.py:129-144)."""
    if isinstance(aabb, (list, tuple)):
        aabb = np.array(aabb)
    if isinstance(aabb, np.ndarray):
        aabb_cpu = torch.tensor(aabb, dtype=torch.float32)
    else:
        assert isinstance(aabb, torch.Tensor), (
            f"aabb must be a list, tuple, np.ndarray, or torch.Tensor, but got {type(aabb)}"
        )
        assert not aabb.is_cuda, "aabb Tensor metadata must be on CPU; pass Python values to avoid implicit CUDA sync"
        aabb_cpu = aabb.to(dtype=torch.float32).contiguous()
    assert aabb_cpu.dim() == 2, f"aabb must be a 2D tensor, but got {aabb_cpu.shape}"
    assert aabb_cpu.size(0) == 2, f"aabb must have 2 rows, but got {aabb_cpu.size(0)}"
    assert aabb_cpu.size(1) == 3, f"aabb must have 3 columns, but got {aabb_cpu.size(1)}"
    return aabb_cpu


def _grid_host(voxel_size, grid_size, aabb):
    """Grid of an explicit aabb, derived on the host exactly as the reference
    wrapper does (flexible_dual_grid.py:88-178): (voxel_size list, grid_range
    list, aabb CPU [2, 3])."""
    assert voxel_size is not None or grid_size is not None, "Either voxel_size or grid_size must be provided"
    vs_cpu = _host_vector(voxel_size, "voxel_size", torch.float32, "float") if voxel_size is not None else None
    gs_cpu = _host_vector(grid_size, "grid_size", torch.int32, "int") if grid_size is not None else None
    aabb_cpu = _aabb_host(aabb)
    if vs_cpu is None:
        vs_cpu = ((aabb_cpu[1] - aabb_cpu[0]) / gs_cpu).to(dtype=torch.float32).contiguous()
    if gs_cpu is None:
        gs_cpu = ((aabb_cpu[1] - aabb_cpu[0]) / vs_cpu).round().to(dtype=torch.int32).contiguous()
    return [float(x) for x in vs_cpu.tolist()], [0, 0, 0] + [int(x) for x in gs_cpu.tolist()], aabb_cpu


def _to_device(t: torch.Tensor, device: torch.device) -> torch.Tensor:
    """Small host table -> device without waiting for the work already queued
    there (pinned staging, non-blocking copy)."""
    if device.type != "cuda":
        return t.to(device)
    return t.pin_memory().to(device=device, non_blocking=True)


def _grid_setup(vertices: torch.Tensor, voxel_size, grid_size, aabb):
    """Derive (voxel_size list, grid_range list, aabb_min tensor on device) exactly as
    the reference wrapper does (flexible_dual_grid.py:88-178)."""
    assert voxel_size is not None or grid_size is not None, "Either voxel_size or grid_size must be provided"
    device = vertices.device
    if aabb is not None:
        vs_l, gr_l, aabb_cpu = _grid_host(voxel_size, grid_size, aabb)
        return vs_l, gr_l, _to_device(aabb_cpu, device).contiguous()
    # automatic aabb from the vertices, padded to whole voxels (flexible_dual_grid.py:146-159)
    vs_cpu = gs_cpu = vs_dev = gs_dev = None
    if voxel_size is not None:
        vs_cpu = _host_vector(voxel_size, "voxel_size", torch.float32, "float")
        vs_dev = vs_cpu.to(device=device)
    if grid_size is not None:
        gs_cpu = _host_vector(grid_size, "grid_size", torch.int32, "int")
        gs_dev = gs_cpu.to(device=device)
    lo = vertices.min(dim=0).values
    hi = vertices.max(dim=0).values
    if vs_dev is not None:
        pad = torch.ceil((hi - lo) / vs_dev) * vs_dev - (hi - lo)
        lo = lo - pad * 0.5
        hi = hi + pad * 0.5
    if gs_dev is not None:
        pad = (hi - lo) / (gs_dev - 1)
        lo = lo - pad * 0.5
        hi = hi + pad * 0.5
    aabb_dev = torch.stack([lo, hi], dim=0).float().contiguous()
    if vs_cpu is None:
        assert not aabb_dev.is_cuda, "CUDA inputs require CPU aabb or explicit voxel_size to avoid implicit CUDA sync"
        vs_cpu = ((aabb_dev[1] - aabb_dev[0]) / gs_cpu).to(dtype=torch.float32).contiguous()
    if gs_cpu is None:
        assert not aabb_dev.is_cuda, "CUDA inputs require CPU aabb or explicit grid_size to avoid implicit CUDA sync"
        gs_cpu = ((aabb_dev[1] - aabb_dev[0]) / vs_cpu).round().to(dtype=torch.int32).contiguous()
    voxel_size_arg = [float(x) for x in vs_cpu.tolist()]
    grid_range_arg = [0, 0, 0] + [int(x) for x in gs_cpu.tolist()]
    return voxel_size_arg, grid_range_arg, aabb_dev


def _check_mesh(vertices, faces):
    assert isinstance(vertices, torch.Tensor), f"vertices must be a torch.Tensor, but got {type(vertices)}"
    assert isinstance(faces, torch.Tensor), f"faces must be a torch.Tensor, but got {type(faces)}"
    assert vertices.dim() == 2, f"vertices must be a 2D tensor, but got {vertices.shape}"
    assert vertices.size(1) == 3, f"vertices must have 3 columns, but got {vertices.size(1)}"
    assert faces.dim() == 2, f"faces must be a 2D tensor, but got {faces.shape}"
    assert faces.size(1) == 3, f"faces must have 3 columns, but got {faces.size(1)}"
    assert vertices.device == faces.device, "vertices and faces must be on the same device"


@torch.no_grad()
def mesh_to_flexible_dual_grid(
    vertices: torch.Tensor,
    faces: torch.Tensor,
    voxel_size: Optional[FloatArrayLike] = None,
    grid_size: Optional[IntArrayLike] = None,
    aabb: Optional[ArrayLike] = None,
    face_weight: float = 1.0,
    boundary_weight: float = 1.0,
    regularization_weight: float = 0.1,
    timing: bool = False,
    precision: str = "reference",
) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Convert a triangle mesh into a sparse flexible dual grid.

    Reference: o_voxel/convert/flexible_dual_grid.py:38-205.  Intersection,
    face and boundary QEF terms are accumulated for the occupied voxels and
    each voxel's dual vertex is the constrained QEF minimiser inside it.

    Returns (coords int32 [N,3], dual_vertices float32 [N,3] in grid-local
    world units, intersected bool [N,3]).  Voxels come in ascending
    x + Gx*(y + Gy*z) order (the reference CPU path returns first-insertion
    order and its CUDA path an unspecified order; the sets and the per-voxel
    values are the same).

    ``precision`` is an addition to the reference signature: "reference"
    (default) reproduces the reference CPU float arithmetic bit for bit;
    "compensated" accumulates each voxel's quadric in the voxel's own frame
    with compensated FP32 sums and solves in FP32 with an FP64 fallback for
    ill-conditioned voxels, which brings the dual vertices close to the exact
    (FP64) solution.  Occupancy and flags do not depend on it.
    """
    _check_mesh(vertices, faces)
    device = vertices.device
    vertices = vertices.to(device=device, dtype=torch.float32).contiguous()
    if vertices.is_cuda:
        # the int32 cast and the shift by the aabb's min corner in one library
        # kernel (no framework kernel, whose first launch in a process loads
        # its code): the same values as the reference's torch operations
        if faces.dtype not in (torch.int32, torch.int64):
            faces = faces.to(dtype=torch.int64)
        faces = faces.contiguous()
        voxel_size_arg, grid_range_arg, aabb_dev = _grid_setup(vertices, voxel_size, grid_size, aabb)
        vertices, faces = _C._prepare_mesh(vertices, faces, aabb_dev[0].contiguous())
        return _C.mesh_to_flexible_dual_grid_cuda(vertices, faces, voxel_size_arg, grid_range_arg, face_weight,
                                                  boundary_weight, regularization_weight, precision=precision)
    faces = faces.to(device=device, dtype=torch.int32).contiguous()
    voxel_size_arg, grid_range_arg, aabb_dev = _grid_setup(vertices, voxel_size, grid_size, aabb)
    vertices = vertices - aabb_dev[0].reshape(1, 3)
    return _C.mesh_to_flexible_dual_grid_cpu(vertices, faces, voxel_size_arg, grid_range_arg, face_weight,
                                             boundary_weight, regularization_weight, timing, precision=precision)


@torch.no_grad()
def mesh_to_flexible_dual_grid_batch(
    meshes,
    voxel_size: Optional[FloatArrayLike] = None,
    grid_size: Optional[IntArrayLike] = None,
    aabb: Optional[ArrayLike] = None,
    face_weight: float = 1.0,
    boundary_weight: float = 1.0,
    regularization_weight: float = 0.1,
    precision: str = "reference",
    voxel_sizes=None,
    grid_sizes=None,
    aabbs=None,
) -> List[Tuple[torch.Tensor, torch.Tensor, torch.Tensor]]:
    """Convert many meshes at once (an addition to the reference API, whose
    ``mesh_to_flexible_dual_grid`` converts one mesh per call).

    ``meshes`` is a sequence of ``(vertices, faces)`` on one device.
    ``voxel_size`` / ``grid_size`` / ``aabb`` apply to every mesh; the plural
    ``voxel_sizes`` / ``grid_sizes`` / ``aabbs`` give one value per mesh
    instead.  Each mesh's grid is set up exactly as by
    ``mesh_to_flexible_dual_grid`` and each returned ``(coords, dual_vertices,
    intersected)`` equals that function's result for the mesh alone; the
    conversion runs every pipeline stage once over all meshes.
    """
    meshes = list(meshes)
    M = len(meshes)
    if M == 0:
        return []

    def pick(single, plural, k, name):
        if plural is not None:
            assert len(plural) == M, f"{name}s must have one entry per mesh"
            return plural[k]
        return single

    # grid set-up per mesh, shared by meshes with the same explicit parameters;
    # explicit aabbs are handled on the host (no device round trip)
    cache = {}
    keys = {}

    def key_of(value, plural, k):
        # parameters given once for the whole batch are keyed once
        if plural is not None:
            return _hashable(value)
        if id(value) not in keys:
            keys[id(value)] = _hashable(value)
        return keys[id(value)]

    vs_list, gr_list, shift_rows, v_list, f_list = [], [], [], [], []
    device = meshes[0][0].device if isinstance(meshes[0][0], torch.Tensor) else None
    host_shift = True
    for k, (vertices, faces) in enumerate(meshes):
        _check_mesh(vertices, faces)
        assert vertices.device == device, "all meshes must be on the same device"
        if vertices.dtype != torch.float32:
            vertices = vertices.to(dtype=torch.float32)
        vs_k, gs_k, ab_k = (pick(voxel_size, voxel_sizes, k, "voxel_size"), pick(grid_size, grid_sizes, k, "grid_size"),
                            pick(aabb, aabbs, k, "aabb"))
        key = None
        if ab_k is not None:
            try:
                key = (key_of(vs_k, voxel_sizes, k), key_of(gs_k, grid_sizes, k), key_of(ab_k, aabbs, k))
            except TypeError:
                key = None
        if key is not None and key in cache:
            vs_l, gr_l, a0 = cache[key]
        else:
            if ab_k is not None:
                vs_l, gr_l, aabb_cpu = _grid_host(vs_k, gs_k, ab_k)
                a0 = aabb_cpu[0].reshape(1, 3)
            else:
                vs_l, gr_l, aabb_dev = _grid_setup(vertices, vs_k, gs_k, ab_k)
                a0 = aabb_dev[0].reshape(1, 3)
            if key is not None:
                cache[key] = (vs_l, gr_l, a0)
        host_shift = host_shift and a0.device.type == "cpu"
        vs_list.append(vs_l)
        gr_list.append(gr_l)
        shift_rows.append(a0)
        v_list.append(vertices)
        f_list.append(faces)
    if device.type == "cuda":
        # packed grid-local vertices and batch-global int32 face indices in
        # one library kernel (no framework kernel); each mesh's shift row goes
        # up with the others in one non-blocking copy
        if host_shift:
            shift = _to_device(torch.cat(shift_rows, 0), device)
        else:
            shift = torch.cat([r.to(device=device) for r in shift_rows], 0).contiguous()
        f_list = [(f if f.dtype in (torch.int32, torch.int64) else f.to(dtype=torch.int64)).contiguous()
                  for f in f_list]
        v_all, f_all, nv, nf = _C._batch_prepare([v.contiguous() for v in v_list], f_list, shift)
        return _C.mesh_to_flexible_dual_grid_batch_packed_cuda(v_all, f_all, nv, nf, vs_list, gr_list, face_weight,
                                                               boundary_weight, regularization_weight,
                                                               precision=precision)
    nv = [int(v.size(0)) for v in v_list]
    nf = [int(f.size(0)) for f in f_list]
    # host meshes: packed with whole-batch tensor operations, then converted
    # on the current GPU
    nv_t = torch.tensor(nv, dtype=torch.int64)
    nf_t = torch.tensor(nf, dtype=torch.int64)
    counts = _to_device(torch.cat([nv_t, nf_t, torch.cumsum(nv_t, 0) - nv_t]), device)
    nv_d, nf_d, base = counts[:M], counts[M:2 * M], counts[2 * M:]
    if host_shift:
        shift = _to_device(torch.cat(shift_rows, 0), device)
    else:
        shift = torch.cat([r.to(device=device) for r in shift_rows], 0)
    v_all = torch.cat(v_list, 0) - torch.repeat_interleave(shift, nv_d, dim=0, output_size=sum(nv))
    f_all = torch.cat([f if f.dtype == torch.int64 else f.to(torch.int64) for f in f_list], 0)
    f_all = (f_all + torch.repeat_interleave(base, nf_d, output_size=sum(nf)).unsqueeze(1)).to(torch.int32).contiguous()
    dev = _C._gpu_for_host_call()
    out = _C.mesh_to_flexible_dual_grid_batch_packed_cuda(v_all.to(dev).contiguous(), f_all.to(dev), nv, nf, vs_list,
                                                          gr_list, face_weight, boundary_weight,
                                                          regularization_weight, precision=precision)
    return [tuple(t.cpu() for t in o) for o in out]


def autotune_tile_classes(meshes, scales=(0.25, 0.5, 1.0, 2.0, 4.0), repeats: int = 3, **batch_kwargs) -> dict:
    """Choose the launch-class thresholds of the scanline and face-test kernels
    and the brick-size threshold of the record grouping for meshes like
    ``meshes`` (an addition to the reference API, whose launch shapes and
    8x8x8 bricks are fixed).

    Runs ``mesh_to_flexible_dual_grid_batch(meshes, **batch_kwargs)`` with the
    current thresholds scaled by each of ``scales`` (one warm-up, then
    ``repeats`` timed conversions per candidate), first for the scanline
    thresholds (timing the scanline kernels), then for the face-test thresholds
    (timing the face kernels), then for the faces-per-voxel threshold above
    which a mesh's records are grouped in 4x4x4 instead of 8x8x8 bricks
    (timing the record generation and fold kernels), keeps the fastest of each
    on the device's context and returns them with the measured times.  The
    outputs of a conversion do not depend on the thresholds, only its launch
    shapes do.
    """
    meshes = list(meshes)
    assert len(meshes) > 0, "autotune_tile_classes needs at least one mesh"
    device = meshes[0][0].device
    assert device.type == "cuda", "autotune_tile_classes times device conversions: pass device tensors"
    mark0, face0 = _C.get_tile_classes(device)
    brick0 = _C.get_brick_class(device)
    trials = []

    def timed(slots):
        best = float("inf")
        mesh_to_flexible_dual_grid_batch(meshes, **batch_kwargs)
        for _ in range(max(1, int(repeats))):
            mesh_to_flexible_dual_grid_batch(meshes, **batch_kwargs)
            kt = _C.kernel_times(device)
            best = min(best, sum(kt.get(s, 0.0) for s in slots) if any(s in kt for s in slots) else float("inf"))
        return best

    _C.set_timing(device, True)
    try:
        picks = {}
        for kind, base, slots in (("mark_fpc", mark0, ("mark",)), ("face_vpf", face0, ("face",))):
            best = None
            for sc in scales:
                cand = [float(x) * float(sc) for x in base]
                _C.set_tile_classes(device, **{kind: cand})
                ms = timed(slots)
                trials.append({kind: cand, slots[0] + "_ms": ms})
                if best is None or ms < best[0]:
                    best = (ms, cand)
            _C.set_tile_classes(device, **{kind: best[1]})
            picks[kind] = best
        best = None
        slots = ("gen_events", "gen_fhits", "gen_bhits", "active", "fold", "fold_large")
        for sc in scales:
            cand = float(brick0) * float(sc)
            _C.set_brick_class(device, cand)
            ms = timed(slots)
            trials.append({"brick_fpv": cand, "fold_ms": ms})
            if best is None or ms < best[0]:
                best = (ms, cand)
        _C.set_brick_class(device, best[1])
        picks["brick_fpv"] = best
    finally:
        _C.set_timing(device, False)
    return {"mark_fpc": picks["mark_fpc"][1], "mark_ms": picks["mark_fpc"][0],
            "face_vpf": picks["face_vpf"][1], "face_ms": picks["face_vpf"][0],
            "brick_fpv": picks["brick_fpv"][1], "fold_ms": picks["brick_fpv"][0], "trials": trials}


def _hashable(x):
    """Nested plain-value key of grid parameters; None if a value is a device tensor."""
    if x is None or isinstance(x, (int, float, str)):
        return x
    if isinstance(x, torch.Tensor):
        if x.is_cuda:
            raise TypeError
        return ("t", tuple(x.reshape(-1).tolist()), tuple(x.shape), str(x.dtype))
    if isinstance(x, np.ndarray):
        return ("a", tuple(x.reshape(-1).tolist()), x.shape, str(x.dtype))
    return tuple(_hashable(e) for e in x)


@torch.no_grad()
def intersect_occ(
    vertices: torch.Tensor,
    faces: torch.Tensor,
    voxel_size: Optional[FloatArrayLike] = None,
    grid_size: Optional[IntArrayLike] = None,
    aabb: Optional[ArrayLike] = None,
) -> torch.Tensor:
    """Voxel coordinates intersected by a triangle mesh (occupancy only).

    Reference: o_voxel/convert/flexible_dual_grid.py:208-329.  Same grid set-up
    as :func:`mesh_to_flexible_dual_grid`; returns coords int32 [N,3] in
    ascending x + Gx*(y + Gy*z) order.
    """
    _check_mesh(vertices, faces)
    device = vertices.device
    vertices = vertices.to(device=device, dtype=torch.float32).contiguous()
    faces = faces.to(device=device, dtype=torch.int32).contiguous()
    voxel_size_arg, grid_range_arg, aabb_dev = _grid_setup(vertices, voxel_size, grid_size, aabb)
    vertices = vertices - aabb_dev[0].reshape(1, 3)
    triangles = vertices[faces.to(dtype=torch.long)].contiguous()
    if vertices.is_cuda:
        return _C.intersect_occ_cuda(triangles, voxel_size_arg, grid_range_arg)
    return _C.intersect_occ_cpu(triangles, voxel_size_arg, grid_range_arg)


# split patterns over a quad's 4 rows plus (train mode) its mid vertex in
# column 4: the reference's quad_split_1 / quad_split_train
# (flexible_dual_grid.py:376-381)
_TRAIN_FAN = (0, 1, 4, 1, 2, 4, 2, 3, 4, 3, 0, 4)


def _dual_grid_params(aabb, voxel_size, grid_size, device):
    """aabb / voxel_size / grid_size as the reference's
    flexible_dual_grid_to_mesh derives them (flexible_dual_grid.py:386-417):
    float32 aabb [2,3], float32 voxel_size [3] and int32 grid_size [3]
    tensors on ``device``, with its assertion messages."""
    if isinstance(aabb, (list, tuple)):
        aabb = np.array(aabb)
    if isinstance(aabb, np.ndarray):
        aabb = torch.tensor(aabb, dtype=torch.float32, device=device)
    assert isinstance(aabb, torch.Tensor), f"aabb must be a list, tuple, np.ndarray, or torch.Tensor, but got {type(aabb)}"
    assert aabb.dim() == 2, f"aabb must be a 2D tensor, but got {aabb.shape}"
    assert aabb.size(0) == 2, f"aabb must have 2 rows, but got {aabb.size(0)}"
    assert aabb.size(1) == 3, f"aabb must have 3 columns, but got {aabb.size(1)}"
    aabb = aabb.to(device=device, dtype=torch.float32)
    extent = aabb[1] - aabb[0]
    if voxel_size is not None:
        if isinstance(voxel_size, float):
            voxel_size = [voxel_size] * 3
        if isinstance(voxel_size, (list, tuple)):
            voxel_size = np.array(voxel_size)
        if isinstance(voxel_size, np.ndarray):
            voxel_size = torch.tensor(voxel_size, dtype=torch.float32, device=device)
        if isinstance(voxel_size, torch.Tensor):
            voxel_size = voxel_size.to(device=device)
            grid_size = (extent / voxel_size).round().int()
    else:
        assert grid_size is not None, "Either voxel_size or grid_size must be provided"
        if isinstance(grid_size, int):
            grid_size = [grid_size] * 3
        if isinstance(grid_size, (list, tuple)):
            grid_size = np.array(grid_size)
        if isinstance(grid_size, np.ndarray):
            grid_size = torch.tensor(grid_size, dtype=torch.int32, device=device)
        if isinstance(grid_size, torch.Tensor):
            grid_size = grid_size.to(device=device)
            voxel_size = extent / grid_size
    assert isinstance(voxel_size, torch.Tensor), (
        f"voxel_size must be a float, list, tuple, np.ndarray, or torch.Tensor, but got {type(voxel_size)}"
    )
    assert voxel_size.dim() == 1, f"voxel_size must be a 1D tensor, but got {voxel_size.shape}"
    assert voxel_size.size(0) == 3, f"voxel_size must have 3 elements, but got {voxel_size.size(0)}"
    assert isinstance(grid_size, torch.Tensor), (
        f"grid_size must be an int, list, tuple, np.ndarray, or torch.Tensor, but got {type(grid_size)}"
    )
    assert grid_size.dim() == 1, f"grid_size must be a 1D tensor, but got {grid_size.shape}"
    assert grid_size.size(0) == 3, f"grid_size must have 3 elements, but got {grid_size.size(0)}"
    return aabb, voxel_size.to(torch.float32), grid_size


def flexible_dual_grid_to_mesh(
    coords: torch.Tensor,
    dual_vertices: torch.Tensor,
    intersected_flag: torch.Tensor,
    split_weight: Optional[torch.Tensor],
    aabb: ArrayLike,
    voxel_size: Optional[FloatArrayLike] = None,
    grid_size: Optional[IntArrayLike] = None,
    train: bool = False,
):
    """Triangle mesh of a flexible dual grid: one quad around every
    intersected grid edge whose four voxels are all present, split into two
    triangles (or, with ``train``, four triangles around a split-weighted mid
    vertex).

    Reference: o_voxel/convert/flexible_dual_grid.py:331-477 (same signature,
    grid set-up and assertions).  ``dual_vertices`` are in voxel units
    (the voxel's own [0, 1]^3).  Returns (vertices float32 [N,3] (train:
    [N+L,3]), faces int32 [2L,3] (train: int64 [4L,3])).  The voxel table,
    the neighbour lookups, the quads, the vertices and the split choice are
    one HIP library call; train mode forms its mid vertices with
    differentiable torch operations on the quads the library returns.
    Differences: vertices are float32 whatever the aabb dtype, a grid of
    2^32 cells or more uses 64-bit keys (the reference's int32 cell count
    wraps), duplicate coordinates resolve to the last row, and CPU inputs
    run on the current GPU (results returned on the CPU).
    """
    host = not coords.is_cuda
    dev = _C._gpu_for_host_call() if host else coords.device
    if host:
        coords = coords.to(dev)
        dual_vertices = dual_vertices.to(dev)
        intersected_flag = intersected_flag.to(dev)
        if split_weight is not None:
            split_weight = split_weight.to(dev)
    # the grid parameters are derived on the host (the same float32 operations
    # as on the device, without three device round trips per call)
    aabb_t, vs_t, gs_t = _dual_grid_params(aabb, voxel_size, grid_size, torch.device("cpu"))
    if train:
        assert split_weight is not None, "split_weight must be provided in training mode"
    N = dual_vertices.shape[0]
    gs = [int(x) for x in gs_t.tolist()]
    vs = [float(x) for x in vs_t.tolist()]
    a0 = [float(x) for x in aabb_t[0].tolist()]
    verts, faces, quads = _C.flexible_dual_grid_to_mesh_cuda(
        coords, dual_vertices, intersected_flag, None if train else split_weight, gs, vs, a0,
        faces=not train, quads=train)
    if train or (torch.is_grad_enabled() and dual_vertices.requires_grad):
        # the same values as the library's, through autograd
        verts = (coords.float() + dual_vertices) * vs_t.to(dev) + aabb_t[0].reshape(1, 3).to(dev)
    if train:
        L = quads.size(0)
        q = quads.long()
        corner = verts[q]                                                  # [L, 4, 3]
        w = split_weight.reshape(N, 1)[q]                                  # [L, 4, 1]
        w02 = w[:, 0] * w[:, 2]
        w13 = w[:, 1] * w[:, 3]
        mid02 = (corner[:, 0] + corner[:, 2]) / 2
        mid13 = (corner[:, 1] + corner[:, 3]) / 2
        mid = (w02 * mid02 + w13 * mid13) / (w02 + w13)
        verts = torch.cat([verts, mid], dim=0)
        q5 = torch.cat([q, torch.arange(N, N + L, device=dev).unsqueeze(1)], dim=1)
        faces = q5[:, list(_TRAIN_FAN)].reshape(-1, 3)
    if host:
        return verts.cpu(), faces.cpu()
    return verts, faces
