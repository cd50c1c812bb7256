"""GPU parity of the .vxz format: octree bytes, both attribute filters and
whole files against the Python oracle (oracle/vxz.py).  Bar: byte-identical
octrees, filter rows and files; exact round trips."""

import io

import numpy as np
import pytest
import torch

from oracle import serialize as S
from oracle import vxz as X

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.fixture(scope="module")
def ov():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import o_voxel

    return o_voxel


def _codes(rng, n, depth):
    return np.unique(rng.integers(0, 1 << (3 * depth), n)).astype(np.int32)


@pytest.mark.parametrize("depth", [1, 2, 4, 6, 8])
def test_svo_and_parent_filter_match_oracle(ov, depth):
    print(f"svo depth {depth}", flush=True)
    _C = ov._C
    rng = np.random.default_rng(depth)
    codes = _codes(rng, 4000, depth)
    svo = _C.encode_sparse_voxel_octree_cpu(torch.from_numpy(codes), depth)
    ref = X.svo_encode(codes, depth)
    np.testing.assert_array_equal(svo.numpy(), ref)
    np.testing.assert_array_equal(_C.decode_sparse_voxel_octree_cpu(svo, depth).numpy(), codes)
    a = rng.integers(0, 256, (len(codes), 5)).astype(np.uint8)
    d = _C.encode_sparse_voxel_octree_attr_parent_cpu(svo, depth, torch.from_numpy(a))
    np.testing.assert_array_equal(d.numpy(), X.parent_encode(ref, depth, a))
    np.testing.assert_array_equal(_C.decode_sparse_voxel_octree_attr_parent_cpu(svo, depth, d).numpy(), a)


def test_svo_edge_cases(ov):
    _C = ov._C
    assert _C.encode_sparse_voxel_octree_cpu(torch.zeros(0, dtype=torch.int32), 3).tolist() == [0]
    assert _C.decode_sparse_voxel_octree_cpu(torch.tensor([0], dtype=torch.uint8), 3).numel() == 0
    one = _C.encode_sparse_voxel_octree_cpu(torch.tensor([0o765], dtype=torch.int32), 3)
    assert one.tolist() == [1 << 7, 1 << 6, 1 << 5]
    with pytest.raises(RuntimeError, match="ascending"):
        _C.encode_sparse_voxel_octree_cpu(torch.tensor([5, 3], dtype=torch.int32), 2)
    with pytest.raises(RuntimeError, match="must lie in"):
        _C.encode_sparse_voxel_octree_cpu(torch.tensor([1 << 9], dtype=torch.int32), 3)
    with pytest.raises(RuntimeError):
        _C.decode_sparse_voxel_octree_cpu(torch.tensor([3, 1], dtype=torch.uint8), 3)  # truncated


@pytest.mark.parametrize("res", [8, 32, 256])
def test_neighbor_filter_matches_oracle(ov, res):
    print(f"neighbor res {res}", flush=True)
    _C = ov._C
    rng = np.random.default_rng(res)
    n = min(5000, res ** 3 // 3)
    cells = rng.choice(res ** 3, n, replace=False)
    c = np.stack([cells % res, cells // res % res, cells // (res * res)], 1).astype(np.int32)
    a = rng.integers(0, 256, (n, 3)).astype(np.uint8)
    d = _C.encode_sparse_voxel_octree_attr_neighbor_cpu(torch.from_numpy(c), res, torch.from_numpy(a))
    np.testing.assert_array_equal(d.numpy(), X.neighbor_encode(c, res, a))
    back = _C.decode_sparse_voxel_octree_attr_neighbor_cpu(torch.from_numpy(c), res, d)
    np.testing.assert_array_equal(back.numpy(), a)
    # a dense block: long neighbour chains
    g = np.stack(np.meshgrid(*[np.arange(min(res, 16))] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.int32)
    ag = rng.integers(0, 256, (len(g), 2)).astype(np.uint8)
    dg = _C.encode_sparse_voxel_octree_attr_neighbor_cpu(torch.from_numpy(g), res, torch.from_numpy(ag))
    np.testing.assert_array_equal(dg.numpy(), X.neighbor_encode(g, res, ag))
    np.testing.assert_array_equal(_C.decode_sparse_voxel_octree_attr_neighbor_cpu(torch.from_numpy(g), res, dg).numpy(),
                                  ag)


def _voxels(rng, G, n):
    cells = rng.choice(G ** 3, n, replace=False)
    coord = np.stack([cells % G, cells // G % G, cells // (G * G)], 1).astype(np.int32)
    return coord, {"color": rng.integers(0, 256, (n, 3)).astype(np.uint8),
                   "flag": rng.integers(0, 8, (n, 1)).astype(np.uint8)}


CASES = [("none", "as_is", "deflate", 16), ("parent", "none", "deflate", 16), ("neighbor", "all", "deflate", 32),
         ("parent", "as_is", "none", 64), ("neighbor", "as_is", "lzma", 16), ("none", "all", "none", 128)]


@pytest.mark.parametrize("filt,inter,comp,cs", CASES)
def test_write_vxz_bytes_match_oracle(ov, filt, inter, comp, cs):
    print(f"vxz {filt} {inter} {comp} {cs}", flush=True)
    rng = np.random.default_rng(cs + len(filt))
    coord, attr = _voxels(rng, 96, 3000)
    buf = io.BytesIO()
    ov.io.write_vxz(buf, torch.from_numpy(coord), {k: torch.from_numpy(v) for k, v in attr.items()}, chunk_size=cs,
                    filter=filt, compression=comp, attr_interleave=inter)
    data = buf.getvalue()
    assert data == X.write_vxz_bytes(coord, attr, chunk_size=cs, filter=filt, compression=comp,
                                     attr_interleave=inter)
    c2, a2 = ov.io.read_vxz(io.BytesIO(data))
    rc, ra = X.read_vxz_bytes(data)
    np.testing.assert_array_equal(c2.numpy(), rc)
    for k in attr:
        np.testing.assert_array_equal(a2[k].numpy(), ra[k])
    assert ov.io.read_vxz_info(io.BytesIO(data))["num_voxel"] == 3000


def test_vxz_round_trip_of_a_conversion(ov, tmp_path):
    """The reference example's pipeline: convert, sort by Morton code, pack
    duals to uint8 and flags to 3 bits, write, read back."""
    import meshgen

    v, f = meshgen.icosphere(4)
    G = 256
    coords, dual, inter = ov.convert.mesh_to_flexible_dual_grid(torch.from_numpy(v * 0.4).to(DEV),
                                                                torch.from_numpy(f).to(DEV), grid_size=G,
                                                                aabb=[[-0.5] * 3, [0.5] * 3])
    order = torch.argsort(ov.serialize.encode_seq(coords))
    coords, dual, inter = coords[order], dual[order], inter[order]
    dq = (torch.clamp(dual * G - coords, 0, 1) * 255).type(torch.uint8)
    fl = (inter[:, 0:1] + 2 * inter[:, 1:2] + 4 * inter[:, 2:3]).type(torch.uint8)
    path = str(tmp_path / "sphere.vxz")
    ov.io.write(path, coords, {"dual_vertices": dq, "intersected": fl}, filter="parent", compression="deflate")
    c2, a2 = ov.io.read(path)
    key = lambda c: np.lexsort((c[:, 2], c[:, 1], c[:, 0]))  # noqa: E731
    c1 = coords.cpu().numpy()
    o1, o2 = key(c1), key(c2.numpy())
    np.testing.assert_array_equal(c2.numpy()[o2], c1[o1])
    np.testing.assert_array_equal(a2["dual_vertices"].numpy()[o2], dq.cpu().numpy()[o1])
    np.testing.assert_array_equal(a2["intersected"].numpy()[o2], fl.cpu().numpy()[o1])


def test_vxz_errors(ov):
    c = torch.tensor([[1, 2, 3], [1, 2, 3]], dtype=torch.int32)
    with pytest.raises(ValueError, match="distinct"):
        ov.io.write_vxz(io.BytesIO(), c, {"a": torch.zeros((2, 1), dtype=torch.uint8)}, compression="none")
    with pytest.raises(ImportError, match="zstandard"):
        ov.io.write_vxz(io.BytesIO(), c[:1], {"a": torch.zeros((1, 1), dtype=torch.uint8)}, compression="zstd")
    with pytest.raises(AssertionError, match="Attributes must be uint8"):
        ov.io.write_vxz(io.BytesIO(), c[:1], {"a": torch.zeros((1, 1), dtype=torch.int32)})
