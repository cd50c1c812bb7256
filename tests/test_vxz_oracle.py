"""CPU pins of the .vxz oracle (oracle/vxz.py).

The reference ships no .vxz files, so the oracle is pinned by hand-built
octrees (bit order and pre-order layout), by structural invariants of the
octree (one byte per distinct code prefix, child counts), by closed forms of
the two attribute filters (sibling and neighbour differences on small
layouts, the neighbour priority) and by exact round trips of every
encode / decode pair and of whole files.
"""

import json
import struct

import numpy as np
import pytest

from oracle import serialize as S
from oracle import vxz as X


def _sorted_codes(rng, n, depth):
    return np.unique(rng.integers(0, 1 << (3 * depth), n)).astype(np.int32)


def test_svo_hand_built():
    # depth 2: code 0 = digits (0, 0), code 9 = digits (1, 1)
    assert X.svo_encode([0, 9], 2).tolist() == [0b11, 0b1, 0b10]
    # depth 3, three leaves under one grandchild plus one far leaf: pre-order
    codes = [0b000_000_001, 0b000_000_011, 0b000_000_111, 0b101_010_000]
    assert X.svo_encode(codes, 3).tolist() == [0b100001, 0b1, 0b10001010, 0b100, 0b1]
    assert X.svo_decode(X.svo_encode(codes, 3), 3).tolist() == codes


@pytest.mark.parametrize("depth", [1, 2, 3, 5, 8])
def test_svo_structure_and_round_trip(depth):
    rng = np.random.default_rng(depth)
    codes = _sorted_codes(rng, 3000, depth)
    svo = X.svo_encode(codes, depth)
    # one byte per distinct prefix of every length 0 .. depth-1
    n_nodes = sum(len(np.unique(codes.astype(np.int64) >> (3 * (depth - j)))) for j in range(depth))
    assert len(svo) == n_nodes
    # every byte's children: the other nodes plus the leaves
    assert int(sum(bin(b).count("1") for b in svo.tolist())) == n_nodes - 1 + len(codes)
    np.testing.assert_array_equal(X.svo_decode(svo, depth), codes)


def test_svo_duplicate_codes_set_the_same_bit():
    assert X.svo_encode([5, 5, 7], 2).tolist() == X.svo_encode([5, 7], 2).tolist()


def test_parent_filter_closed_form():
    # depth 1: the root's children are the leaves
    a = np.array([[10, 200], [13, 100], [12, 0]], np.uint8)
    svo = X.svo_encode([1, 4, 6], 1)
    d = X.parent_encode(svo, 1, a)
    assert d.tolist() == [[10, 200], [3, 156], [255, 156]]
    np.testing.assert_array_equal(X.parent_decode(svo, 1, d), a)
    # depth 2, two inner nodes: rows 1 = second inner node minus first (their
    # first leaves), then each inner node's sibling differences in pre-order
    codes = [0o01, 0o03, 0o21, 0o24, 0o27]
    a = np.arange(5, dtype=np.uint8)[:, None] * 10
    d = X.parent_encode(X.svo_encode(codes, 2), 2, a)
    assert d[:, 0].tolist() == [0, 20, 10, 10, 10]


@pytest.mark.parametrize("depth", [1, 3, 6])
def test_parent_filter_round_trip(depth):
    rng = np.random.default_rng(10 + depth)
    codes = _sorted_codes(rng, 2000, depth)
    a = rng.integers(0, 256, (len(codes), 4)).astype(np.uint8)
    svo = X.svo_encode(codes, depth)
    np.testing.assert_array_equal(X.parent_decode(svo, depth, X.parent_encode(svo, depth, a)), a)


def test_neighbor_filter_closed_form_and_priority():
    # a row along x: each voxel minus its x-neighbour
    c = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]])
    a = np.array([[5], [9], [4]], np.uint8)
    assert X.neighbor_encode(c, 4, a)[:, 0].tolist() == [5, 4, 251]
    # (1,1,1) has both its x-neighbour (0,1,1) and its xyz-neighbour (0,0,0):
    # xyz wins; (0,1,1) has only its yz-neighbour (0,0,0)
    c = np.array([[1, 1, 1], [0, 1, 1], [0, 0, 0]])
    a = np.array([[100], [7], [30]], np.uint8)
    d = X.neighbor_encode(c, 4, a)
    assert d[:, 0].tolist() == [70, 233, 30]
    np.testing.assert_array_equal(X.neighbor_decode(c, 4, d), a)


@pytest.mark.parametrize("res", [8, 32])
def test_neighbor_filter_round_trip(res):
    rng = np.random.default_rng(res)
    cells = rng.choice(res ** 3, min(3000, res ** 3 // 3), replace=False)
    c = np.stack([cells % res, cells // res % res, cells // (res * res)], 1)
    a = rng.integers(0, 256, (len(c), 3)).astype(np.uint8)
    np.testing.assert_array_equal(X.neighbor_decode(c, res, X.neighbor_encode(c, res, a)), a)


CASES = [(f, i, "deflate") for f in ("none", "parent", "neighbor") for i in ("none", "as_is", "all")] + [
    ("parent", "as_is", "none"), ("neighbor", "all", "lzma")]


@pytest.mark.parametrize("filt,inter,comp", CASES)
def test_file_round_trip(filt, inter, comp):
    rng = np.random.default_rng(len(filt) * 100 + len(inter) * 10 + len(comp))
    G, n = 48, 1500
    cells = rng.choice(G ** 3, n, replace=False)
    coord = np.stack([cells % G, cells // G % G, cells // (G * G)], 1)
    attr = {"color": rng.integers(0, 256, (n, 3)).astype(np.uint8),
            "flag": rng.integers(0, 8, (n, 1)).astype(np.uint8)}
    data = X.write_vxz_bytes(coord, attr, chunk_size=16, filter=filt, compression=comp, attr_interleave=inter)
    assert data[:4] == b"VXZ\x00"
    start = struct.unpack(">I", data[4:8])[0]
    info = json.loads(data[8:start])
    assert info["num_voxel"] == n and len(data) == start + info["compressed_size"]
    assert [k for k in info["chunks"][0]][:2] == ["idx", "svo"] and list(info["chunks"][0])[-1] == "ptr"
    c2, a2 = X.read_vxz_bytes(data)
    # chunks in ascending (x, y, z) chunk order, Morton order inside
    key = lambda c: np.lexsort((c[:, 2], c[:, 1], c[:, 0]))  # noqa: E731
    o1, o2 = key(coord), key(c2)
    np.testing.assert_array_equal(c2[o2], coord[o1])
    for k in attr:
        np.testing.assert_array_equal(a2[k][o2], attr[k][o1])
    ch = c2.astype(np.int64) // 16
    lo = c2.astype(np.int64) % 16
    m = S.z_order_encode(lo[:, 0], lo[:, 1], lo[:, 2]).astype(np.int64)
    order_key = ((ch[:, 0] * 1000 + ch[:, 1]) * 1000 + ch[:, 2]) * (1 << 30) + m
    assert (np.diff(order_key) > 0).all()
