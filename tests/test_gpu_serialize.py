"""GPU parity of o_voxel.serialize (encode_seq / decode_seq) and the
_C.{z_order,hilbert}_{encode,decode}_{cuda,cpu} boundary against the numpy
oracle (oracle/serialize.py): bit-exact codes and coordinates."""

import sys

import numpy as np
import pytest
import torch

from oracle import serialize as S

pytestmark = pytest.mark.gpu
DEV = "cuda"
MODES = {"z_order": (S.z_order_encode, S.z_order_decode), "hilbert": (S.hilbert_encode, S.hilbert_decode)}


@pytest.fixture(scope="module")
def ov():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import o_voxel

    return o_voxel


def _coords(seed, n=200000):
    rng = np.random.default_rng(seed)
    c = rng.integers(0, 1024, (n, 3)).astype(np.int32)
    odd = np.array([[-1, 0, 5], [1024, 3, 2], [2 ** 31 - 1, -2 ** 31, 7], [0, 0, 0], [1023, 1023, 1023]], np.int32)
    return np.concatenate([c, odd])


@pytest.mark.parametrize("mode", ["z_order", "hilbert"])
@pytest.mark.parametrize("permute", [[0, 1, 2], [2, 0, 1], [1, 2, 0]])
def test_encode_decode_matches_oracle(ov, mode, permute):
    print(f"serialize {mode} {permute}", flush=True)
    enc, dec = MODES[mode]
    c = _coords(len(mode) + permute[0])
    t = torch.from_numpy(c).to(DEV)
    codes = ov.serialize.encode_seq(t, permute, mode)
    want = enc(c[:, permute[0]], c[:, permute[1]], c[:, permute[2]])
    assert codes.dtype == torch.int32
    np.testing.assert_array_equal(codes.cpu().numpy(), want)
    back = ov.serialize.decode_seq(codes, permute, mode)
    d = dec(want)
    ref = np.stack([d[permute.index(j)] for j in range(3)], 1)
    assert back.dtype == torch.int32 and back.shape == (len(c), 3)
    np.testing.assert_array_equal(back.cpu().numpy(), ref)
    # in-range coordinates come back unchanged
    np.testing.assert_array_equal(back[:-5].cpu().numpy(), c[:-5])


@pytest.mark.parametrize("mode", ["z_order", "hilbert"])
def test_int64_strided_and_host_inputs(ov, mode):
    enc, _ = MODES[mode]
    c = _coords(7, 5000).astype(np.int64)
    c[:10] += 1 << 33  # the low 32 bits are used, as torch's .int() does
    t64 = torch.from_numpy(c).to(DEV)
    want = enc(c[:, 0], c[:, 1], c[:, 2])
    np.testing.assert_array_equal(ov.serialize.encode_seq(t64, mode=mode).cpu().numpy(), want)
    # a strided view of a wider tensor
    wide = torch.zeros((len(c), 5), dtype=torch.int32, device=DEV)
    wide[:, 1:4] = torch.from_numpy(c.astype(np.int32)).to(DEV)
    np.testing.assert_array_equal(ov.serialize.encode_seq(wide[:, 1:4], mode=mode).cpu().numpy(), want)
    # float coordinates go through .int() as in the reference
    fc = c[10:110].astype(np.float32) + np.float32(0.7)  # rows in [0, 1024)
    f = torch.from_numpy(fc).to(DEV)
    np.testing.assert_array_equal(ov.serialize.encode_seq(f, mode=mode).cpu().numpy(),
                                  enc(*(np.trunc(fc).astype(np.int64).T)))
    # host tensors run on the GPU and come back on the host
    h = ov.serialize.encode_seq(torch.from_numpy(c[:, :3]), mode=mode)
    assert not h.is_cuda
    np.testing.assert_array_equal(h.numpy(), want)


def test_C_boundary_functions(ov):
    _C = ov._C
    c = _coords(11, 3000)
    x, y, z = (torch.from_numpy(np.ascontiguousarray(c[:, k])).to(DEV) for k in range(3))
    for mode in ("z_order", "hilbert"):
        enc, dec = MODES[mode]
        codes = getattr(_C, f"{mode}_encode_cuda")(x, y, z)
        np.testing.assert_array_equal(codes.cpu().numpy(), enc(c[:, 0], c[:, 1], c[:, 2]))
        xs = getattr(_C, f"{mode}_decode_cuda")(codes)
        for got, want in zip(xs, dec(codes.cpu().numpy())):
            np.testing.assert_array_equal(got.cpu().numpy(), want)
        hc = getattr(_C, f"{mode}_encode_cpu")(x.cpu(), y.cpu(), z.cpu())
        assert not hc.is_cuda and torch.equal(hc, codes.cpu())
        hx = getattr(_C, f"{mode}_decode_cpu")(codes.cpu())
        assert all(torch.equal(a, b.cpu()) for a, b in zip(hx, xs))


def test_empty_and_errors(ov):
    e = ov.serialize.encode_seq(torch.zeros((0, 3), dtype=torch.int32, device=DEV))
    assert e.shape == (0,)
    assert ov.serialize.decode_seq(e).shape == (0, 3)
    with pytest.raises(AssertionError, match="Input coordinates must be of shape"):
        ov.serialize.encode_seq(torch.zeros((4, 2), dtype=torch.int32, device=DEV))
    with pytest.raises(ValueError, match="Unknown encoding mode"):
        ov.serialize.encode_seq(torch.zeros((4, 3), dtype=torch.int32, device=DEV), mode="peano")
    with pytest.raises(AssertionError, match="Input code must be of shape"):
        ov.serialize.decode_seq(torch.zeros((4, 3), dtype=torch.int32, device=DEV))


def test_sorted_fdg_voxels(ov):
    """The example's use: sort a conversion's voxels by their codes."""
    import meshgen

    v, f = meshgen.icosphere(4)
    coords, _, _ = ov.convert.mesh_to_flexible_dual_grid(torch.from_numpy(v * 0.4).to(DEV),
                                                         torch.from_numpy(f).to(DEV), grid_size=256,
                                                         aabb=[[-0.5] * 3, [0.5] * 3])
    for mode in ("z_order", "hilbert"):
        codes = ov.serialize.encode_seq(coords, mode=mode)
        order = torch.argsort(codes)
        assert torch.equal(ov.serialize.decode_seq(codes[order], mode=mode), coords[order])
        assert len(torch.unique(codes)) == len(codes)
