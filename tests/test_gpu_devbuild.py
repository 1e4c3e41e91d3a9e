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


# Adaptation (cg_tree_adapt_device): the tree of one scene adapted to another
# scene's target levels, as a simulation step adapts its grid
# (src/FluidEuler.h:820-835): refinement rounds, then CoarsenStep until none
# deletes (src/PoissonGrid.h:165-314), then ghosts and cell types.  The
# oracle's AdaptScene runs the reference's steps on the host.
MOVED = dict(ADAPTIVE256, solid_c=[0.56, 0.45, 0.5], band_c=[0.56, 0.45, 0.5])
ADAPT = {
    "moving_sphere": (ADAPTIVE256, MOVED),
    "shrinking_band": (ADAPTIVE256, dict(ADAPTIVE256, band_level=[4, 3, 2])),
    "uniform_to_plane_up": (dict(kind=0, uniform_level=4), dict(kind=2, plane_level=[1, 4], plane_z=0.6)),
    "band_to_plane_down": (SCENES["band_deep"], dict(kind=2, plane_level=[4, 1], plane_z=0.3)),
    "plane_to_band": (SCENES["plane"], ADAPTIVE256),
    "batch5_to_batch17": (batch_scene_params(5)[0], batch_scene_params(17)[0]),
    "batch96_to_batch0": (batch_scene_params(96)[0], batch_scene_params(0)[0]),
}
OPS = {"refine": ca.ADAPT_REFINE, "coarsen": ca.ADAPT_COARSEN, "both": ca.ADAPT_REFINE | ca.ADAPT_COARSEN}


def assert_same_tree(dt, o, what):
    lev, crd, typ, _ = o.topology()
    ct = o.celltypes()
    kd, ko = keyed(dt.level, dt.coord), keyed(lev, crd)
    assert kd.keys() == ko.keys(), f"{what}: {len(kd)} device tiles, {len(ko)} oracle tiles"
    d_idx = np.array([kd[k] for k in ko])
    o_idx = np.array(list(ko.values()))
    assert np.array_equal(dt.type[d_idx], typ[o_idx])
    assert np.array_equal(dt.cell_type.reshape(-1, 512)[d_idx], ct[o_idx])


@pytest.mark.parametrize("ops", list(OPS))
@pytest.mark.parametrize("name", list(ADAPT))
def test_device_adaptation_equals_oracle(name, ops):
    frm, to = ADAPT[name]
    start = ca.scene_tree_device(**frm)
    dt = ca.tree_adapt_device(start, ops=OPS[ops], **to)
    assert_same_tree(dt, po.Oracle.adapted(frm, to, OPS[ops]), f"{name}/{ops}")


def test_device_adaptation_of_a_host_tree_and_a_solve_on_it():
    frm, to = ADAPT["moving_sphere"]
    o0 = po.Oracle(**frm)
    lev, crd, typ, _ = o0.topology()
    host = ca.Tree(o0.params.h0, lev, crd, typ, o0.celltypes())
    dt = ca.tree_adapt_device(host, **to)
    o = po.Oracle.adapted(frm, to)
    assert_same_tree(dt, o, "host tree")
    lev, crd, typ, _ = o.topology()
    ko = keyed(lev, crd)
    to_o = np.array([ko[(int(l), *map(int, c))] for l, c in zip(dt.level, np.asarray(dt.coord).reshape(-1, 3))])
    P = ca.Problem(dt)
    P.set(ca.CH_B, o.get(po.CH_B).astype(np.float32)[to_o])
    S = ca.GMGSolver()
    S.solve(P, False, 1000, 1e-6, 2, 10, 1, False)
    ro = o.solve(max_iters=1000, tol=1e-6, level_iters=2, coarsest_iters=10)
    assert (S.last["iters"], S.last["rr"], S.last["bb"]) == (ro["iters"], ro["rr"], ro["bb"])
    leaf = dt.type == po.LEAF
    assert np.array_equal(P.get(ca.CH_X)[leaf], o.get(po.CH_X).astype(np.float32)[to_o][leaf])


def test_device_adaptation_limits():
    start = ca.scene_tree_device(**ADAPTIVE256)
    with pytest.raises(ca.CirrusError, match="bad ops"):
        ca.tree_adapt_device(start, ops=4, **ADAPTIVE256)
    with pytest.raises(ca.CirrusError, match="h0"):
        ca.tree_adapt_device(start, h0=0.25, **{k: v for k, v in ADAPTIVE256.items()})
    with pytest.raises(ca.CirrusError, match="root tiles"):
        ca.tree_adapt_device(ca.scene_tree_device(kind=0, nroot=[1, 1, 2], uniform_level=1), kind=0, uniform_level=1)
    with pytest.raises(ca.CirrusError, match="max_tiles"):
        ca.tree_adapt_device(start, max_tiles=1000, **MOVED)


def test_device_build_limits():
    with pytest.raises(ca.CirrusError, match="uniform, sphere band and plane"):
        ca.scene_tree_device(kind=3)
    with pytest.raises(ca.CirrusError, match="max_tiles"):
        ca.scene_tree_device(max_tiles=100, **SCENES["adaptive256"])
