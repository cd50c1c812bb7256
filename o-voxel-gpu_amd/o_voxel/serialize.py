"""30-bit voxel serialization codes (MI355X build).

Same signatures, defaults and errors as the reference module
``o_voxel/serialize.py``.  Codes are computed by the HIP library behind
``o_voxel._C``, reading the columns of the [N, 3] coordinate tensor in place
(int32 or int64) in the requested permutation, and decoding writes the
permuted columns of the [N, 3] result directly.  CPU tensors are processed on
the current GPU.
"""

from typing import List, Literal

import torch

from . import _C

__all__ = ["encode_seq", "decode_seq"]

_MODES = ("z_order", "hilbert")


@torch.no_grad()
def encode_seq(coords: torch.Tensor, permute: List[int] = [0, 1, 2],
               mode: Literal["z_order", "hilbert"] = "z_order") -> torch.Tensor:
    """int32 [N] codes of coords [N, 3] (reference o_voxel/serialize.py:6-35):
    axis ``permute[0]`` supplies the most significant bit of each triple."""
    assert coords.shape[-1] == 3 and coords.ndim == 2, "Input coordinates must be of shape [N, 3]"
    if coords.device.type not in ("cpu", "cuda"):
        raise ValueError(f"Unsupported device type: {coords.device.type}")
    if mode not in _MODES:
        raise ValueError(f"Unknown encoding mode: {mode}")
    host = coords.device.type == "cpu"
    c = coords.to(_C._gpu_for_host_call()) if host else coords
    if c.dtype not in (torch.int32, torch.int64):
        c = c.int()  # the reference's .int() conversion
    # columns used in place: entries of permute index the columns as in the reference
    codes = _C.seq_encode_columns((c[:, permute[0]], c[:, permute[1]], c[:, permute[2]]), mode)
    return codes.cpu() if host else codes


@torch.no_grad()
def decode_seq(code: torch.Tensor, permute: List[int] = [0, 1, 2],
               mode: Literal["z_order", "hilbert"] = "z_order") -> torch.Tensor:
    """int32 [N, 3] coordinates of codes [N] (reference o_voxel/serialize.py:38-69)."""
    assert code.ndim == 1, "Input code must be of shape [N]"
    if code.device.type not in ("cpu", "cuda"):
        raise ValueError(f"Unsupported device type: {code.device.type}")
    if mode not in _MODES:
        raise ValueError(f"Unknown decoding mode: {mode}")
    # column j of the result is decoded axis permute.index(j) (as the reference)
    src = [list(permute).index(j) for j in range(3)]
    host = code.device.type == "cpu"
    c = _C._seq_codes_in(code.to(_C._gpu_for_host_call()) if host else code)
    out = torch.empty((c.numel(), 3), dtype=torch.int32, device=c.device)
    cols = [None] * 3
    for j in range(3):
        cols[src[j]] = out[:, j]
    _C.seq_decode_into(c, mode, cols)
    return out.cpu() if host else out
