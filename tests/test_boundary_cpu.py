"""CPU-tier checks of the drop-in boundary: the C-ABI library loads and
exports every symbol include/ovx_fdg.h declares, the Python mirror of the
reference interface keeps its signatures, grid set-up and error behaviour,
and no path silently falls back to a CPU implementation."""

import ctypes
import inspect
import os
import re
import subprocess
import threading

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ovx_fdg.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ovx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from o_voxel import _C

    path = _C.library_path()
    assert os.path.exists(path), "libovx_fdg.so not built"
    lib = ctypes.CDLL(path)
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    for s in syms:
        assert re.search(rf"\bT {s}$", out, re.M), s
    assert lib.ovx_version  # callable without a device
    lib.ovx_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.ovx_version()


def test_library_is_built_for_gfx950_only():
    from o_voxel import _C

    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", _C.library_path()], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    assert "gfx950" in out.stdout
    for arch in ("gfx90a", "gfx942", "gfx1100"):
        assert arch not in out.stdout


def test_public_signatures_match_reference():
    import o_voxel

    sig = inspect.signature(o_voxel.convert.mesh_to_flexible_dual_grid)
    # the reference's parameters in the reference's order, then the one
    # addition (the precision variant) with a default that keeps the
    # reference behaviour
    assert list(sig.parameters) == ["vertices", "faces", "voxel_size", "grid_size", "aabb", "face_weight",
                                    "boundary_weight", "regularization_weight", "timing", "precision"]
    d = {k: v.default for k, v in sig.parameters.items()}
    assert d["face_weight"] == 1.0 and d["boundary_weight"] == 1.0 and d["regularization_weight"] == 0.1
    assert d["timing"] is False and d["precision"] == "reference"
    sig = inspect.signature(o_voxel.convert.intersect_occ)
    assert list(sig.parameters) == ["vertices", "faces", "voxel_size", "grid_size", "aabb"]
    from o_voxel import _C

    for name in ("mesh_to_flexible_dual_grid_cuda", "mesh_to_flexible_dual_grid_cpu", "intersect_occ_cuda",
                 "intersect_occ_cpu"):
        assert callable(getattr(_C, name))
    assert list(inspect.signature(_C.mesh_to_flexible_dual_grid_cpu).parameters)[-2:] == ["timing", "precision"]


def test_assertions_like_reference():
    import o_voxel

    f = torch.zeros((1, 3), dtype=torch.int64)
    with pytest.raises(AssertionError, match="vertices must be a torch.Tensor"):
        o_voxel.convert.mesh_to_flexible_dual_grid(np.zeros((3, 3)), f, grid_size=8)
    with pytest.raises(AssertionError, match="vertices must have 3 columns"):
        o_voxel.convert.mesh_to_flexible_dual_grid(torch.zeros((3, 2)), f, grid_size=8)
    with pytest.raises(AssertionError, match="faces must be a 2D tensor"):
        o_voxel.convert.mesh_to_flexible_dual_grid(torch.zeros((3, 3)), torch.zeros(3), grid_size=8)
    with pytest.raises(AssertionError, match="Either voxel_size or grid_size must be provided"):
        o_voxel.convert.mesh_to_flexible_dual_grid(torch.zeros((3, 3)), f)
    with pytest.raises(AssertionError, match="grid_size must have 3 elements"):
        o_voxel.convert.mesh_to_flexible_dual_grid(torch.zeros((3, 3)), f, grid_size=[8, 8])
    with pytest.raises(AssertionError, match="aabb must have 2 rows"):
        o_voxel.convert.mesh_to_flexible_dual_grid(torch.zeros((3, 3)), f, grid_size=8, aabb=[[0, 0, 0]])


def test_grid_setup_matches_reference_formulas():
    """voxel_size = f32(aabb1 - aabb0) / G; grid_range = [0,0,0,G]; the
    auto-aabb padding of the reference wrapper (flexible_dual_grid.py:146-178)."""
    from o_voxel.convert.flexible_dual_grid import _grid_setup

    v = torch.tensor([[-0.3, -0.2, -0.1], [0.4, 0.25, 0.3]], dtype=torch.float32)
    for G in (64, 100, 512, 1024):
        vs, gr, a0 = _grid_setup(v, None, G, [[-0.5] * 3, [0.5] * 3])
        want = float(np.float32(np.float32(1.0) / np.float32(G)))
        assert vs == [want] * 3 and gr == [0, 0, 0, G, G, G]
        assert torch.equal(a0[0], torch.tensor([-0.5] * 3))
    vs, gr, a0 = _grid_setup(v, 0.05, None, [[-0.5] * 3, [0.5] * 3])
    assert gr == [0, 0, 0, 20, 20, 20] and np.allclose(vs, 0.05)
    # automatic aabb with grid_size: padding (hi - lo) / (G - 1) / 2 on each side
    vs, gr, a0 = _grid_setup(v, None, 10, None)
    lo, hi = v.min(0).values, v.max(0).values
    pad = (hi - lo) / (torch.tensor([10, 10, 10], dtype=torch.int32) - 1)
    assert torch.allclose(a0[0], lo - pad * 0.5)
    assert np.allclose(vs, ((hi + pad * 0.5) - (lo - pad * 0.5)).numpy() / 10)
    # automatic aabb with voxel_size: padded to whole voxels
    vs, gr, a0 = _grid_setup(v, [0.1, 0.1, 0.1], None, None)
    ext = hi - lo
    vs_t = torch.tensor([0.1] * 3, dtype=torch.float32)
    pad = torch.ceil(ext / vs_t) * vs_t - ext
    g = (((hi + pad * 0.5) - (lo - pad * 0.5)) / vs_t).round().int().tolist()
    assert gr[3:] == g and all(x >= e for x, e in zip(g, [7, 5, 4]))


def test_no_cpu_fallback_without_gpu():
    """Host tensors run the HIP path on a GPU; with no GPU they raise."""
    import o_voxel

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    v = torch.rand((10, 3))
    f = torch.randint(0, 10, (5, 3))
    with pytest.raises(RuntimeError, match="no CPU implementation"):
        o_voxel.convert.mesh_to_flexible_dual_grid(v, f, grid_size=8, aabb=[[0] * 3, [1] * 3])
    with pytest.raises(RuntimeError, match="no CPU implementation"):
        o_voxel.convert.intersect_occ(v, f, grid_size=8, aabb=[[0] * 3, [1] * 3])


def test_context_creation_returns_without_gpu():
    """Creating a device context reports the missing device instead of
    blocking (the library loader and the context table use separate locks)."""
    from o_voxel import _C

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    box = {}

    def run():
        try:
            _C._ctx(torch.device("cuda", 0))
        except RuntimeError as e:
            box["err"] = str(e)

    t = threading.Thread(target=run, daemon=True)
    t.start()
    t.join(20)
    assert not t.is_alive(), "context creation blocked"
    assert "ovx_fdg_create" in box.get("err", "")


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "o-voxel-gpu_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".cpp")):
                txt = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in txt.lower(), os.path.join(dirpath, fn)
