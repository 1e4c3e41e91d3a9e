"""Device-side build of the scenes' octrees (cg_scene_tree_device): refinement
rounds with 2:1 face balance (RefineLeafsOneStep / IterativeRefine,
src/PoissonGrid.h:35-162,317-326), ghost spawning (SpawnGhostTiles,
src/PoissonGrid.cu:30-60) and cell typing (CalcCellTypesFromLeafs,
src/PoissonSolver.cu:7-61), every step a kernel.  The tree must equal the
oracle's (a loop-for-loop restatement of the reference's refinement) tile for
tile -- level, coordinate, tile type and all 512 cell types -- and a solve on
it must give the oracle's solve bit for bit.
"""
import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po
from cirrus_amd.workloads import ADAPTIVE256, PLAIN64, batch_scene_params

pytestmark = pytest.mark.gpu

SCENES = {
    "adaptive256": ADAPTIVE256,
    "plain64": PLAIN64,
    "plane": dict(kind=2, plane_level=[3, 2]),
    "two_roots": dict(kind=0, nroot=[1, 1, 2], uniform_level=2, solid_r=0.1),
    "band_deep": dict(kind=1, solid_r=0.12, band_r=0.12, band_w=[0.01, 0.05], band_level=[6, 5, 3]),
}
SCENES.update({f"batch{q}": batch_scene_params(q)[0] for q in (0, 1, 2, 3, 5, 17, 96, 242)})


def keyed(level, coord):
    coord = np.asarray(coord).reshape(-1, 3)
    return {(int(l), *map(int, c)): i for i, (l, c) in enumerate(zip(level, coord))}


@pytest.mark.parametrize("name", list(SCENES))
def test_device_tree_equals_oracle_tree(name):
    kw = SCENES[name]
    dt = ca.scene_tree_device(**kw)
    o = po.Oracle(**kw)
    lev, crd, typ, _ = o.topology()
    ct = o.celltypes()
    kd, ko = keyed(dt.level, dt.coord), keyed(lev, crd)
    assert kd.keys() == ko.keys(), f"{name}: {len(kd)} device tiles, {len(ko)} oracle tiles"
    d_idx = np.array([kd[k] for k in ko])
    o_idx = np.array(list(ko.values()))
    assert np.array_equal(dt.type[d_idx], typ[o_idx])
    assert np.array_equal(dt.cell_type.reshape(-1, 512)[d_idx], ct[o_idx])


def test_device_tree_solves_like_the_oracle():
    kw = SCENES["adaptive256"]
    dt = ca.scene_tree_device(**kw)
    o = po.Oracle(**kw)
    lev, crd, typ, _ = o.topology()
    ko = keyed(lev, crd)
    to_o = np.array([ko[(int(l), *map(int, c))] for l, c in zip(dt.level, np.asarray(dt.coord).reshape(-1, 3))])
    P = ca.Problem(dt)
    b = o.get(po.CH_B).astype(np.float32)
    P.set(ca.CH_B, b[to_o])
    S = ca.GMGSolver()
    S.solve(P, False, 1000, 1e-6, 2, 10, 1, False)
    ro = o.solve(max_iters=1000, tol=1e-6, level_iters=2, coarsest_iters=10)
    assert (S.last["iters"], S.last["rr"], S.last["bb"]) == (ro["iters"], ro["rr"], ro["bb"])
    leaf = dt.type == po.LEAF
    assert np.array_equal(P.get(ca.CH_X)[leaf], o.get(po.CH_X).astype(np.float32)[to_o][leaf])


def test_device_build_limits():
    with pytest.raises(ca.CirrusError, match="uniform, sphere band and plane"):
        ca.scene_tree_device(kind=3)
    with pytest.raises(ca.CirrusError, match="max_tiles"):
        ca.scene_tree_device(max_tiles=100, **SCENES["adaptive256"])
