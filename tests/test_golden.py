"""Golden fixtures: the oracle reproduces them bit for bit (CPU tier) and the
HIP path matches them after canonical voxel ordering (GPU tier)."""

import os

import numpy as np
import pytest
import torch

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["ico_cap_g64.npz", "square_g32.npz"]


def load(name):
    return dict(np.load(os.path.join(HERE, name), allow_pickle=False))


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(oracle, name):
    g = load(name)
    w = g["weights"].tolist()
    a = oracle.flexible_dual_grid(g["vertices"], g["faces"], g["voxel_size"], g["grid_range"], *w)
    assert np.array_equal(a.coords, g["coords"])
    assert np.array_equal(a.flags, g["flags"])
    assert np.array_equal(a.dual, g["dual"])
    b = oracle.flexible_dual_grid(g["vertices"], g["faces"], g["voxel_size"], g["grid_range"], *w, fp64=True)
    assert np.array_equal(b.dual64, g["dual64"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_golden(oracle, name):
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from o_voxel import _C

    g = load(name)
    dev = torch.device("cuda", 0)
    w = g["weights"].tolist()
    c, d, i = _C.mesh_to_flexible_dual_grid_cuda(torch.from_numpy(g["vertices"]).to(dev),
                                                  torch.from_numpy(g["faces"]).to(dev),
                                                  g["voxel_size"].tolist(), g["grid_range"].tolist(), *w)
    perm = oracle.canonical_order(g["coords"], g["grid_range"])
    assert np.array_equal(c.cpu().numpy(), g["coords"][perm])
    assert np.array_equal(i.cpu().numpy(), g["flags"][perm])
    assert np.array_equal(d.cpu().numpy(), g["dual"][perm])
    # error of the float path against the FP64 oracle, in voxels (reported)
    e = (d.cpu().numpy().astype(np.float64) - g["dual64"][perm]) / g["voxel_size"].astype(np.float64)
    print(f"{name}: float-path dual vs FP64 rms={np.sqrt((e ** 2).sum(1).mean()):.3e} max={np.abs(e).max():.3e}")
