"""Writes the regression fixtures in this directory from the CPU oracle.

  python tests/golden/make_golden.py

Fixtures (numpy .npz, compressed):
  ico_cap_g64.npz   the icosphere-4 + open cap mesh at grid 64, weights
                    face 1.0 / boundary 0.2 / regularization 1e-2:
                    inputs (grid-local vertices, faces, voxel_size,
                    grid_range) and the oracle outputs in first-insertion
                    order (coords, dual float path, dual FP64 path, flags)
  square_g32.npz    an axis-aligned square at a non-grid z (closed-form case)

The reference repository ships no golden vectors for this path, and its CPU
implementation needs Eigen, which is absent; these fixtures pin the oracle
against regressions (they are produced by the oracle, whose building blocks
are checked independently in tests/test_oracle.py).
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import meshgen  # noqa: E402
from oracle import oracle as O  # noqa: E402

W = (1.0, 0.2, 1e-2)


def write(name, v, f, vs, gr, weights=W):
    a = O.flexible_dual_grid(v, f, vs, gr, *weights)
    b = O.flexible_dual_grid(v, f, vs, gr, *weights, fp64=True)
    assert np.array_equal(a.coords, b.coords)
    np.savez_compressed(os.path.join(HERE, name), vertices=v.astype(np.float32), faces=f.astype(np.int32),
                        voxel_size=np.asarray(vs, np.float32), grid_range=np.asarray(gr, np.int64),
                        weights=np.asarray(weights, np.float32), coords=a.coords, dual=a.dual, dual64=b.dual64,
                        flags=a.flags)
    print(name, len(a.coords))


def main():
    O.build()
    v, f = meshgen.ico_cap()
    vs, gr, a0 = meshgen.grid_args(64)
    write("ico_cap_g64.npz", (v - a0).astype(np.float32), f, vs, gr)
    G = 32
    vsq = np.array([[0.2131, 0.3047, 0.4443], [0.6789, 0.3047, 0.4443], [0.6789, 0.7211, 0.4443],
                    [0.2131, 0.7211, 0.4443]], np.float32)
    write("square_g32.npz", vsq, np.array([[0, 1, 2], [0, 2, 3]], np.int32), [1.0 / G] * 3, [0, 0, 0, G, G, G],
          (1.0, 0.0, 0.1))


if __name__ == "__main__":
    main()
