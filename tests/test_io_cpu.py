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


def _ply_bytes(fmt, props, rows):
    """A PLY file written independently of the codec: header text plus the
    body packed with numpy (binary) or as text lines (ascii)."""
    head = ["ply", f"format {fmt} 1.0", "comment written by the test", f"element vertex {len(rows)}"]
    head += [f"property {t} {n}" for n, t, _ in props]
    head += ["element face 0", "property list uchar int vertex_indices", "end_header"]
    out = ("\n".join(head) + "\n").encode("ascii")
    if fmt == "ascii":
        return out + "".join(" ".join(str(x) for x in r) + "\n" for r in rows).encode("ascii")
    bo = "<" if fmt == "binary_little_endian" else ">"
    dt = np.dtype([(n, bo + d) for n, _, d in props])
    return out + np.array([tuple(r) for r in rows], dtype=dt).tobytes()


@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian", "ascii"])
def test_ply_read_known_files(fmt):
    """Known files in each encoding: coordinates rounded from float x, y, z,
    properties <name>_<j> gathered into attribute channels, sized type aliases."""
    import o_voxel

    props = [("x", "float", "f4"), ("y", "float32", "f4"), ("z", "float", "f4"),
             ("color_0", "uchar", "u1"), ("color_1", "uint8", "u1"), ("color_2", "uchar", "u1"),
             ("density_0", "double", "f8"), ("label_0", "short", "i2")]
    rows = [[1.0, 2.0, 3.0, 10, 20, 30, 0.5, -7], [400.0, 0.0, 5.0, 255, 0, 1, -1.25, 300]]
    coord, attr = o_voxel.io.read_ply(io.BytesIO(_ply_bytes(fmt, props, rows)))
    assert coord.tolist() == [[1, 2, 3], [400, 0, 5]]
    assert set(attr) == {"color", "density", "label"}
    assert attr["color"].dtype == torch.uint8 and attr["color"].tolist() == [[10, 20, 30], [255, 0, 1]]
    assert attr["density"].dtype == torch.float64 and attr["density"].flatten().tolist() == [0.5, -1.25]
    assert attr["label"].dtype == torch.int16 and attr["label"].flatten().tolist() == [-7, 300]


def test_ply_write_layout_and_round_trip(tmp_path):
    """The written header names float x, y, z and one typed property per
    channel (what plyfile writes for the reference's structured array); the
    body is the packed rows; reading it back gives the same voxels."""
    import o_voxel

    coord = torch.tensor([[0, 1, 2], [511, 3, 7], [5, 5, 5]], dtype=torch.int32)
    attr = {"rgb": torch.tensor([[1, 2, 3], [4, 5, 6], [7, 8, 9]], dtype=torch.uint8),
            "w": torch.tensor([[0.25], [1.5], [-2.0]], dtype=torch.float32),
            "id": torch.tensor([[1], [70000], [3]], dtype=torch.int32)}
    p = str(tmp_path / "v.ply")
    o_voxel.io.write(p, coord, attr)
    raw = open(p, "rb").read()
    head, body = raw.split(b"end_header\n", 1)
    lines = head.decode().splitlines()
    assert lines[:3] == ["ply", "format binary_little_endian 1.0", "element vertex 3"]
    assert lines[3:] == ["property float x", "property float y", "property float z",
                         "property uchar rgb_0", "property uchar rgb_1", "property uchar rgb_2",
                         "property float w_0", "property int id_0"]
    rec = np.frombuffer(body, dtype=np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("r", "u1", 3),
                                              ("w", "<f4"), ("id", "<i4")]))
    assert rec["x"].tolist() == [0, 511, 5] and rec["id"].tolist() == [1, 70000, 3]
    c2, a2 = o_voxel.io.read(p)
    assert torch.equal(c2, coord.long())
    assert all(torch.equal(a2[k], attr[k]) for k in attr)
    with pytest.raises(AssertionError, match="Unsupported data type"):
        o_voxel.io.write_ply(io.BytesIO(), coord, {"b": torch.zeros((3, 1), dtype=torch.bool)})
    with pytest.raises(ValueError, match="not a PLY file"):
        o_voxel.io.read_ply(io.BytesIO(b"PLX\n"))
    with pytest.raises(ValueError, match="shorter"):
        o_voxel.io.read_ply(io.BytesIO(raw[:-3]))


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
