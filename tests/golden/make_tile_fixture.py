"""Writes tests/golden/uniform16_solution.ctf: the uniform 16^3 problem
(levels 0-1, solid sphere r = 0.2 NEUMANN, top face DIRICHLET, seeded random
divergence), its right-hand side (channel 1) and the oracle's FP32 pressure
(channel 0; RB-GS 1+1, 10 bottom iterations, tol 1e-6), in the portable tile
format.  Run from the repository root: python tests/golden/make_tile_fixture.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import cirrus_amd as ca  # noqa: E402
import pyoracle as po  # noqa: E402


def main():
    o = po.Oracle(uniform_level=1)
    lev, crd, typ, nbr = o.topology()
    tree = ca.Tree(o.params.h0, lev, crd, typ, o.celltypes())
    b = o.get(po.CH_B).astype(np.float32)
    o.solve(max_iters=200, tol=1e-6, level_iters=1, coarsest_iters=10)
    x = o.get(po.CH_X).astype(np.float32)
    out = os.path.join(ROOT, "tests", "golden", "uniform16_solution.ctf")
    ca.save_tiles(out, tree, {0: x, 1: b})
    print(out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
