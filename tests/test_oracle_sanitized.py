"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/Makefile target liboracle_fdg_asan.so, host code only) runs a set of
conversions and building blocks without a report.  The sanitized library is
loaded in a child interpreter with the ASan runtime preloaded; any report
aborts the child (halt_on_error), which fails the test."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
import numpy as np
sys.path[:0] = [ROOT]
import meshgen
from oracle import oracle as O
O._LIB_PATH = LIB
v, f = meshgen.ico_cap()
vs, gr, a0 = meshgen.grid_args(64)
vl = (v - a0).astype(np.float32)
for fp64 in (False, True):
    r = O.flexible_dual_grid(vl, f, vs, gr, 1.0, 0.2, 1e-2, fp64=fp64, keep_qef=True)
    assert len(r.coords) > 1000
v, f = meshgen.scene(seed=5, target_faces=3000)
vs, gr, a0 = meshgen.grid_args(48)
vl = (v - a0).astype(np.float32)
O.flexible_dual_grid(vl, f, vs, gr, 2.0, 0.7, 0.0)
O.intersect_occ(vl[f], vs, gr)
# a triangle partly outside the grid, a degenerate one, a segment leaving the grid
O.flexible_dual_grid(np.array([[-0.2, 0.3, 0.4], [1.3, 0.5, 0.45], [0.4, 1.2, 0.5], [0.1, 0.1, 0.1]], np.float32),
                     np.array([[0, 1, 2], [3, 3, 3]], np.int32), [1 / 16] * 3, [0, 0, 0, 16, 16, 16], 1.0, 1.0, 0.1)
O.boundary_voxels(np.array([[0.5, 0.5, 0.5], [1.7, -0.3, 0.2]], np.float32), [1 / 16] * 3, [0, 0, 0, 16, 16, 16])
O.cpqr_solve(np.zeros((3, 3), np.float32), np.ones(3, np.float32))
O.cpqr_solve(np.random.default_rng(0).normal(size=(3, 3)), np.ones(3))
print("sanitized oracle ok")
"""


def test_oracle_under_asan_ubsan():
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", odir, "liboracle_fdg_asan.so"], check=True)
    lib = os.path.join(odir, "liboracle_fdg_asan.so")
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    code = f"ROOT = {ROOT!r}\nLIB = {lib!r}\n" + CHILD
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "sanitized oracle ok" in r.stdout, (r.returncode, r.stdout[-2000:],
                                                                     r.stderr[-4000:])
