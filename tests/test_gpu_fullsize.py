"""Full-size parity through the library's own builders.

The problems are built by ``ca.Scene`` -- the library's host builder:
refinement with 2:1 face balance (RefineLeafsOneStep / IterativeRefine,
src/PoissonGrid.h:35-162,317-326), ghost spawning (SpawnGhostTiles,
src/PoissonGrid.cu:30-60), cell typing (CalcCellTypesFromLeafs,
src/PoissonSolver.cu:7-61) and the divergence right-hand side
(src/PoissonSolver.cu:66-142) -- at the sizes the benchmark times: the adaptive
256^3 scene (5,505 tiles) and eight problems of the 256-problem batched pool,
among them its largest (13,756 leaf tiles).  Each scene must equal the oracle's
scene bit for bit (tiles, types, cell types, right-hand side); each GPU solve
must equal the oracle's solve bit for bit (iterations, |r|^2, pressure); the
FP32 pressure of the adaptive scene must lie within REL_L2_TOL of the FP64
oracle's converged pressure.  The pool is solved under each level-form setting
(the eager prolongation/restriction, the fused first 16-bit sweep and the
one-launch red-black iterations never, at the defaults, always), so every
kernel form meets the oracle on levels of every size, small and large (the
pool's finest level has more than 16,384 tiles).
"""
import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po
from cirrus_amd.workloads import ADAPTIVE256, batch_scene_params, build_scenes

pytestmark = pytest.mark.gpu

# relative L2 error over LEAF INTERIOR cells of the FP32 pressure solved to a
# recursive residual of 1e-6, against the FP64 oracle solved to 1e-12
REL_L2_TOL = 1e-4
# problems of the batched pool (seed 0): the two largest, a large one, the
# median, three mid-sized ones and the smallest (uniform level 2)
POOL = [242, 218, 96, 188, 124, 109, 5, 17]
LEVEL_FORMS = {"never": (0, 0, 0), "default": (-1, -1, -1), "always": (1 << 30, 1 << 30, 1)}
_cache = {}


def by_key(level, coord):
    return {(int(l), *map(int, c)): i for i, (l, c) in enumerate(zip(level, coord))}


def scene_vs_oracle(sc, o):
    """Asserts the library-built scene equals the oracle's; returns the map from
    scene tile index to oracle tile index."""
    lev, crd, typ, _ = o.topology()
    T = sc.tree
    assert T.n == o.n
    ko, ks = by_key(lev, crd), by_key(T.level, T.coord)
    assert ko.keys() == ks.keys()
    to_oracle = np.empty(T.n, np.int64)
    for k, i in ks.items():
        to_oracle[i] = ko[k]
    assert np.array_equal(T.type, typ[to_oracle])
    assert np.array_equal(T.cell_type, o.celltypes()[to_oracle])
    b = o.get(po.CH_B).astype(np.float32)
    assert np.array_equal(sc.rhs.view(np.uint32), b[to_oracle].view(np.uint32))
    return to_oracle


def adaptive():
    if "a" not in _cache:
        sc = ca.Scene(**ADAPTIVE256)
        o = po.Oracle(**ADAPTIVE256)
        _cache["a"] = (sc, o, scene_vs_oracle(sc, o))
    return _cache["a"]


def test_adaptive256_scene_equals_oracle_scene():
    sc, o, _ = adaptive()
    assert sc.tree.n == 5505 and int((sc.tree.type == po.LEAF).sum()) == 3424


def test_adaptive256_solve_bitexact_and_fp64_error():
    sc, o, to_o = adaptive()
    P = ca.Problem(sc.tree)
    b = np.ascontiguousarray(sc.rhs, np.float32)
    P.set(ca.CH_B, b)
    S = ca.GMGSolver()
    S.solve(P, False, 1000, 1e-6, 2, 10, 1, False, verify_f64=True)
    r = S.last
    x = P.get(ca.CH_X)
    b_or = o.get(po.CH_B).copy()  # the oracle's solve overwrites b with the residual
    ro = o.solve(max_iters=1000, tol=1e-6, level_iters=2, coarsest_iters=10)
    assert r["converged"] == 1
    assert (r["iters"], r["rr"], r["bb"]) == (ro["iters"], ro["rr"], ro["bb"])
    leaf = sc.tree.type == po.LEAF
    xo = o.get(po.CH_X).astype(np.float32)[to_o]
    assert np.array_equal(x[leaf], xo[leaf])
    # against the FP64 oracle's converged pressure
    o64 = po.Oracle(params=o.params, fp64=True)
    o64.set(po.CH_B, b_or)
    r64 = o64.solve(max_iters=1000, tol=1e-12, level_iters=2, coarsest_iters=10)
    assert r64["rel_res"] < 1e-12
    x64 = o64.get(po.CH_X)[to_o]
    m = leaf[:, None] & (sc.tree.cell_type == po.INTERIOR)
    err = float(np.linalg.norm((x.astype(np.float64) - x64)[m]) / np.linalg.norm(x64[m]))
    print(f"adaptive 256^3: {r['iters']} iterations, FP64 |b-Ax|/|b| {r['rel_res_f64']:.3e}, "
          f"rel. L2 error vs FP64 oracle {err:.3e}")
    assert err < REL_L2_TOL
    assert r["rel_res_f64"] < 5e-6


def pool():
    if "p" not in _cache:
        kws = [batch_scene_params(q)[0] for q in POOL]
        scenes = build_scenes(ca, kws, workers=8)
        items = []
        for kw, sc in zip(kws, scenes):
            o = po.Oracle(**kw)
            o.b0 = o.get(po.CH_B).copy()  # every solve overwrites b with its residual
            items.append((sc, o, scene_vs_oracle(sc, o)))
        B = ca.Problem.batch([sc.tree for sc, _, _ in items])
        _cache["p"] = (items, B, {})
    return _cache["p"]


def test_pool_scenes_equal_oracle_scenes():
    items, B, _ = pool()
    leaf = [int((sc.tree.type == po.LEAF).sum()) for sc, _, _ in items]
    assert max(leaf) == 13756
    # the pool's finest level is large enough for the level-size limit of the eager forms
    lev = np.concatenate([sc.tree.level[sc.tree.type != po.GHOST] for sc, _, _ in items])
    assert np.bincount(lev).max() > 16384


def oracle_solves(storage):
    items, B, memo = pool()
    if storage not in memo:
        out = []
        for sc, o, to_o in items:
            o.set(po.CH_B, o.b0)
            ro = o.solve(max_iters=1000, tol=1e-6, level_iters=2, coarsest_iters=10, storage=storage)
            out.append((ro, o.get(po.CH_X).astype(np.float32)[to_o]))
        memo[storage] = out
    return memo[storage]


@pytest.mark.parametrize("forms", list(LEVEL_FORMS))
@pytest.mark.parametrize("storage", [ca.STORE_F32, ca.STORE_BF16])
def test_pool_batch_solve_bitexact(storage, forms):
    items, B, _ = pool()
    ref = oracle_solves(storage)
    B.set_level_forms(*LEVEL_FORMS[forms])
    try:
        B.set(ca.CH_B, np.concatenate([np.ascontiguousarray(sc.rhs, np.float32) for sc, _, _ in items]))
        res = ca.GMGSolver().solve_batch(B, 1000, 1e-6, 2, 10, storage=storage)
        xb = B.split(B.get(ca.CH_X))
    finally:
        B.set_level_forms(-1, -1, -1)
    for q, ((sc, _, _), (ro, xo)) in enumerate(zip(items, ref)):
        assert res[q]["converged"] == 1, POOL[q]
        assert (res[q]["iters"], res[q]["rr"], res[q]["bb"]) == (ro["iters"], ro["rr"], ro["bb"]), POOL[q]
        leaf = sc.tree.type == po.LEAF
        assert np.array_equal(xb[q][leaf], xo[leaf]), f"pool problem {POOL[q]}: pressure differs from the oracle"


@pytest.mark.parametrize("storage", [ca.STORE_F32, ca.STORE_BF16])
def test_pool_looped_colour_sweeps_bitexact(storage):
    """Colour sweeps with each wavefront walking its share of a level's tiles
    and the next tile's loads issued ahead (tile block 8, used on levels with
    more tiles than the device holds wavefronts): the same solves bit for bit."""
    items, B, _ = pool()
    ref = oracle_solves(storage)
    B.set_schedule(0, True, 8)
    try:
        B.set(ca.CH_B, np.concatenate([np.ascontiguousarray(sc.rhs, np.float32) for sc, _, _ in items]))
        res = ca.GMGSolver().solve_batch(B, 1000, 1e-6, 2, 10, storage=storage)
        xb = B.split(B.get(ca.CH_X))
    finally:
        B.set_schedule(0, True, 1)
    for q, ((sc, _, _), (ro, xo)) in enumerate(zip(items, ref)):
        assert (res[q]["iters"], res[q]["rr"], res[q]["bb"]) == (ro["iters"], ro["rr"], ro["bb"]), POOL[q]
        leaf = sc.tree.type == po.LEAF
        assert np.array_equal(xb[q][leaf], xo[leaf]), f"pool problem {POOL[q]}: pressure differs from the oracle"
