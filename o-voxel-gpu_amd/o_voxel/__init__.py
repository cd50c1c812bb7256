"""o_voxel (MI355X build): the flexible-dual-grid conversion path.

Provided: ``o_voxel.convert`` (mesh_to_flexible_dual_grid, its batched form,
intersect_occ, flexible_dual_grid_to_mesh), ``o_voxel.serialize``
(encode_seq / decode_seq), ``o_voxel.io`` (the .vxz format, .npz) and the
``o_voxel._C`` boundary they call.
"""
from . import _C, convert, io, serialize

__all__ = ["convert", "io", "serialize", "_C"]
