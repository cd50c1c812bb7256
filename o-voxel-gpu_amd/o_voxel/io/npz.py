"""Voxels in .npz files (reference o_voxel/io/npz.py, same behaviour)."""

from typing import Dict

import numpy as np
import torch

__all__ = ["read_npz", "write_npz"]


def read_npz(file):
    """(coord int32 [N, 3], {name: array}) of an .npz file written by write_npz."""
    data = np.load(file)  # no pickled objects
    coord = torch.from_numpy(data["coord"]).int()
    return coord, {k: torch.from_numpy(v) for k, v in data.items() if k != "coord"}


def write_npz(file, coord: torch.Tensor, attr: Dict[str, torch.Tensor], compress=True):
    """Coordinates (stored as uint16) and attributes as arrays of an .npz file."""
    data = {"coord": coord.cpu().numpy().astype(np.uint16)}
    data.update({k: v.cpu().numpy() for k, v in attr.items()})
    (np.savez_compressed if compress else np.savez)(file, **data)
