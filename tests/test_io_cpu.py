"""CPU checks of the file I/O layer that needs no GPU (.npz, dispatch,
header parsing) and of the reader's refusal to run without one."""

import io

import numpy as np
import pytest
import torch


def test_npz_round_trip(tmp_path):
    import o_voxel

    coord = torch.tensor([[1, 2, 3], [400, 5, 6]], dtype=torch.int32)
    attr = {"c": torch.arange(6, dtype=torch.uint8).reshape(2, 3)}
    p = str(tmp_path / "v.npz")
    o_voxel.io.write(p, coord, attr)
    c2, a2 = o_voxel.io.read(p)
    assert torch.equal(c2, coord) and torch.equal(a2["c"], attr["c"])


def test_dispatch_and_header():
    import o_voxel
    from oracle import vxz as X

    with pytest.raises(ValueError, match="Unsupported file type"):
        o_voxel.io.read("x.obj")
    data = X.write_vxz_bytes(np.array([[0, 0, 0], [3, 1, 2]]), {"a": np.array([[7], [9]], np.uint8)},
                             chunk_size=4, compression="none")
    info = o_voxel.io.read_vxz_info(io.BytesIO(data))
    assert info["num_voxel"] == 2 and info["chunk_size"] == 4 and info["attr"] == [["a", 1]]
    with pytest.raises(AssertionError, match="Invalid file type"):
        o_voxel.io.read_vxz_info(io.BytesIO(b"XYZ\x00\x00\x00\x00\x08"))
    with pytest.raises(ImportError, match="plyfile"):
        o_voxel.io.read_ply("x.ply")


def test_vxz_reader_needs_a_gpu():
    import o_voxel
    from oracle import vxz as X

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    data = X.write_vxz_bytes(np.array([[0, 0, 0]]), {"a": np.array([[7]], np.uint8)}, chunk_size=4,
                             compression="none")
    with pytest.raises(RuntimeError, match="no CPU implementation"):
        o_voxel.io.read_vxz(io.BytesIO(data))


@pytest.mark.parametrize("depth", [1, 2, 3, 6, 8])
def test_host_octree_parsers_match_oracle(depth):
    """The reader's pre-order parsers run on the host (library, no GPU call)."""
    from o_voxel import _C
    from oracle import vxz as X

    rng = np.random.default_rng(depth)
    codes = np.unique(rng.integers(0, 1 << (3 * depth), 3000)).astype(np.int32)
    svo = X.svo_encode(codes, depth)
    np.testing.assert_array_equal(_C.svo_decode_host(svo, depth), codes)
    a = rng.integers(0, 256, (len(codes), 4)).astype(np.uint8)
    np.testing.assert_array_equal(_C.svo_parent_decode_host(svo, depth, X.parent_encode(svo, depth, a)), a)
    if len(svo) > 1:
        with pytest.raises(RuntimeError, match="truncated|malformed"):
            _C.svo_decode_host(svo[: len(svo) // 2], depth)
