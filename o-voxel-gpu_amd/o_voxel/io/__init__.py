"""Voxel file I/O (reference o_voxel/io/__init__.py): .vxz (the O-voxel
format), .npz, and .ply (own PLY codec, no plyfile needed)."""

from typing import Dict

import torch

from .npz import *  # noqa: F401,F403
from .ply import *  # noqa: F401,F403
from .vxz import *  # noqa: F401,F403
from .npz import read_npz, write_npz
from .ply import read_ply, write_ply
from .vxz import read_vxz, write_vxz


def read(file_path: str):
    """(coord, {name: attribute}) of a .npz, .ply or .vxz file."""
    if file_path.endswith(".npz"):
        return read_npz(file_path)
    if file_path.endswith(".ply"):
        return read_ply(file_path)
    if file_path.endswith(".vxz"):
        return read_vxz(file_path)
    raise ValueError(f"Unsupported file type {file_path}")


def write(file_path: str, coord: torch.Tensor, attr: Dict[str, torch.Tensor], **kwargs):
    """Write voxels to a .npz, .ply or .vxz file."""
    if file_path.endswith(".npz"):
        return write_npz(file_path, coord, attr, **kwargs)
    if file_path.endswith(".ply"):
        return write_ply(file_path, coord, attr, **kwargs)
    if file_path.endswith(".vxz"):
        return write_vxz(file_path, coord, attr, **kwargs)
    raise ValueError(f"Unsupported file type {file_path}")
