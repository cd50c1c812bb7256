"""o_voxel (MI355X build): the flexible-dual-grid conversion path.

Only ``o_voxel.convert.mesh_to_flexible_dual_grid`` / ``intersect_occ`` and
the ``o_voxel._C`` boundary they call are provided.
"""
from . import _C, convert

__all__ = ["convert", "_C"]
