"""Argument checks of the extension boundary that run before any device work
(no GPU needed): a malformed call raises RuntimeError instead of reaching a
kernel that would read past the end of a tensor."""

import pytest
import torch


def _C():
    from o_voxel import _C

    return _C


def test_hashmap_insert_rejects_short_values():
    table_k = torch.empty(64, dtype=torch.uint32)
    table_v = torch.empty(64, dtype=torch.uint32)
    keys = torch.arange(10, dtype=torch.int64).to(torch.uint32)
    with pytest.raises(RuntimeError, match="values has 9 entries for 10 keys"):
        _C().hashmap_insert_cuda(table_k, table_v, keys, torch.zeros(9, dtype=torch.uint32))


def test_hashmap_rejects_non_1d_keys():
    table_k = torch.empty(64, dtype=torch.uint64)
    table_v = torch.empty(64, dtype=torch.uint64)
    keys = torch.zeros((2, 5), dtype=torch.uint64)
    with pytest.raises(RuntimeError, match="1-D"):
        _C().hashmap_insert_cuda(table_k, table_v, keys, torch.zeros(10, dtype=torch.uint64))
    with pytest.raises(RuntimeError, match="1-D"):
        _C().hashmap_lookup_cuda(table_k, table_v, keys)
    with pytest.raises(RuntimeError, match="1-D"):
        _C().hashmap_lookup_cuda(table_k, table_v, torch.tensor(3, dtype=torch.uint64))


def test_hashmap_3d_rejects_bad_coords_and_short_values():
    table_k = torch.empty(64, dtype=torch.uint32)
    table_v = torch.empty(64, dtype=torch.uint32)
    coords = torch.zeros((6, 4), dtype=torch.int32)
    with pytest.raises(RuntimeError, match="values has 5 entries for 6 keys"):
        _C().hashmap_insert_3d_cuda(table_k, table_v, coords, torch.zeros(5, dtype=torch.uint32), 8, 8, 8)
    with pytest.raises(RuntimeError, match=r"shape \[M, 4\]"):
        _C().hashmap_insert_3d_cuda(table_k, table_v, coords[:, :3], torch.zeros(6, dtype=torch.uint32), 8, 8, 8)
    with pytest.raises(RuntimeError, match=r"shape \[M, 4\]"):
        _C().hashmap_lookup_3d_cuda(table_k, table_v, coords.reshape(-1), 8, 8, 8)
    with pytest.raises(RuntimeError, match=r"shape \[M, 4\]"):
        _C().hashmap_insert_3d_idx_as_val_cuda(table_k, table_v, coords[:, :2], 8, 8, 8)


def test_hashmap_table_must_be_on_device():
    """Well-formed keys but a host table: refused before any library call."""
    table_k = torch.empty(64, dtype=torch.uint32)
    table_v = torch.empty(64, dtype=torch.uint32)
    keys = torch.arange(4, dtype=torch.int64).to(torch.uint32)
    with pytest.raises(RuntimeError, match="device tensor"):
        _C().hashmap_insert_cuda(table_k, table_v, keys, keys.clone())


def test_input_preparation_rejects_bad_arguments():
    """The wrapper's preparation entry points check their tensors on the host:
    host tensors, wrong dtypes, an empty batch."""
    v = torch.zeros((4, 3), dtype=torch.float32)
    f = torch.zeros((2, 3), dtype=torch.int64)
    o = torch.zeros(3, dtype=torch.float32)
    with pytest.raises(RuntimeError):
        _C()._prepare_mesh(v, f, o)  # host tensors
    with pytest.raises(RuntimeError, match="at least one mesh"):
        _C()._batch_prepare([], [])
    with pytest.raises(RuntimeError, match="one entry per mesh"):
        _C()._batch_prepare([v], [])
    with pytest.raises(RuntimeError):
        _C()._batch_prepare([v], [f])  # host tensors
