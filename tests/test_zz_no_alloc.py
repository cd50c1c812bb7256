"""GPU: the allocation-free conversion (ovx_fdg_convert_noalloc), called
through the C-ABI, against ovx_fdg_begin / ovx_fdg_finish on the same inputs:
bit-identical voxel coordinates, dual vertices and flags on two calls in a
row, the second on a smaller mesh and grid (stale scratch contents of the
first would show), in both precisions.  Conversions that need more than the
scratch or have more voxels than the output rows are refused with
OVX_ERR_CAPACITY and write no output; the context's own workspace and
results are untouched by these calls."""

import ctypes

import numpy as np
import pytest
import torch

import meshgen

pytestmark = pytest.mark.gpu

OVX_ERR_CAPACITY = 3
KW = (1.0, 0.2, 1e-2)  # face, boundary, regularization weights
COND_MAX = 1.0e6


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda", 0)


def _case(mesh, grid):
    v, f = mesh
    vs, gr, a0 = meshgen.grid_args(grid)
    return (v - a0).astype(np.float32), f.astype(np.int32), [float(x) for x in vs], [int(x) for x in gr]


def _noalloc(dev, case, precision, rows):
    """One ovx_fdg_convert_noalloc call with `rows` output rows, prefilled
    with sentinels: (status, n_voxels, coords, dual, flags)."""
    from o_voxel import _C

    vl, f, vs, gr = case
    L = _C._load()
    ctx = _C._ctx(dev)
    vt = torch.from_numpy(vl).to(dev)
    ft = torch.from_numpy(f).to(dev)
    coords = torch.full((rows, 3), -7, dtype=torch.int32, device=dev)
    dual = torch.full((rows, 3), float("nan"), dtype=torch.float32, device=dev)
    flags = torch.full((rows, 3), 0xAB, dtype=torch.uint8, device=dev)
    n = ctypes.c_int64(-1)
    g = _C._grid(vs, gr)
    with torch.cuda.device(dev), ctx.lock:
        stream = _C._stream_handle(dev)
        assert L.ovx_fdg_set_precision(ctx.h, _C._PRECISIONS[precision], COND_MAX) == 0
        st = L.ovx_fdg_convert_noalloc(ctx.h, vt.data_ptr(), vt.size(0), ft.data_ptr(), ft.size(0), ctypes.byref(g),
                                       KW[0], KW[1], KW[2], coords.data_ptr(), dual.data_ptr(), flags.data_ptr(),
                                       rows, ctypes.c_void_p(stream), ctypes.byref(n))
        msg = L.ovx_last_error().decode(errors="replace") if st else ""
    torch.cuda.synchronize(dev)
    return st, int(n.value), coords, dual, flags, msg


def _begin_finish(dev, case, precision):
    from o_voxel import _C

    vl, f, vs, gr = case
    vt = torch.from_numpy(vl).to(dev)
    ft = torch.from_numpy(f).to(dev)
    c, d, i = _C.mesh_to_flexible_dual_grid_cuda(vt, ft, vs, gr, *KW, precision=precision, cond_max=COND_MAX)
    torch.cuda.synchronize(dev)
    return c.cpu(), d.cpu(), i.cpu().view(torch.uint8)


def _assert_same(noalloc, ref, what):
    st, n, coords, dual, flags, msg = noalloc
    c, d, i = ref
    assert st == 0, f"{what}: status {st}: {msg}"
    assert n == c.shape[0], f"{what}: {n} voxels vs {c.shape[0]}"
    assert torch.equal(coords[:n].cpu(), c), what
    assert torch.equal(dual[:n].cpu().view(torch.int32), d.view(torch.int32)), what
    assert torch.equal(flags[:n].cpu(), i), what
    # rows past the voxel count are left alone
    assert bool((coords[n:] == -7).all()) and bool((flags[n:] == 0xAB).all()), what


@pytest.mark.parametrize("precision", ["reference", "compensated"])
def test_noalloc_equals_begin_finish_twice(precision):
    dev = _dev()
    big = _case(meshgen.ico_cap(), 64)
    small = _case(meshgen.scene(seed=4, target_faces=1500), 32)
    print(f"[case] noalloc vs begin/finish, ico_cap@64 then scene-1500@32, {precision}", flush=True)
    ref_big = _begin_finish(dev, big, precision)
    ref_small = _begin_finish(dev, small, precision)
    assert ref_small[0].shape[0] < ref_big[0].shape[0]
    rows = ref_big[0].shape[0] + 100
    _assert_same(_noalloc(dev, big, precision, rows), ref_big, "first call")
    _assert_same(_noalloc(dev, small, precision, rows), ref_small, "second, smaller call")
    # the context's own path afterwards: same results as before
    again = _begin_finish(dev, small, precision)
    for a, b in zip(again, ref_small):
        assert torch.equal(a, b)


def test_noalloc_refuses_too_few_rows():
    dev = _dev()
    case = _case(meshgen.ico_cap(), 64)
    ref = _begin_finish(dev, case, "reference")
    N = ref[0].shape[0]
    print(f"[case] noalloc with {N - 1} rows for {N} voxels", flush=True)
    st, n, coords, dual, flags, msg = _noalloc(dev, case, "reference", N - 1)
    assert st == OVX_ERR_CAPACITY and "output capacity" in msg, (st, msg)
    assert n == 0
    assert bool((coords == -7).all()) and bool(torch.isnan(dual).all()) and bool((flags == 0xAB).all())
    # and the next call with room succeeds
    _assert_same(_noalloc(dev, case, "reference", N), ref, "after a refused call")


def test_noalloc_refuses_beyond_scratch():
    from o_voxel import _C

    dev = _dev()
    # a 1024^3 grid's occupancy bitset alone (128 MiB) exceeds the scratch
    case = _case(meshgen.ico_cap(), 1024)
    print("[case] noalloc on a 1024^3 grid (beyond the scratch)", flush=True)
    st, n, coords, _, flags, msg = _noalloc(dev, case, "reference", 1000)
    assert st == OVX_ERR_CAPACITY and "scratch" in msg, (st, msg)
    assert n == 0 and bool((coords == -7).all()) and bool((flags == 0xAB).all())
    # the context still converts normally, and the allocation-free path still works
    small = _case(meshgen.ico_cap(), 48)
    ref = _begin_finish(dev, small, "reference")
    _assert_same(_noalloc(dev, small, "reference", ref[0].shape[0]), ref, "after a scratch refusal")
    assert _C.workspace_bytes(dev) > 0
