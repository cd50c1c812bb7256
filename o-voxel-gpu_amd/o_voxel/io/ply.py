"""Voxels in .ply files (reference o_voxel/io/ply.py).  Needs the plyfile
package, which this environment does not provide: the functions raise
ImportError when it is missing."""

__all__ = ["read_ply", "write_ply"]


def _plyfile():
    try:
        import plyfile
    except ImportError as e:
        raise ImportError("reading / writing .ply voxels needs the plyfile package, which is not installed") from e
    return plyfile


def read_ply(file):
    _plyfile()
    raise NotImplementedError("read_ply is not provided by this build")


def write_ply(file, coord, attr):
    _plyfile()
    raise NotImplementedError("write_ply is not provided by this build")
