"""Seeded random scenes through every solver path, GPU against the oracle bit
for bit: tree shapes (one or two roots, uniform, band, plane and vortex
refinement, 2-6 levels), boundary types (any mix of NEUMANN and DIRICHLET
walls, all-NEUMANN with re-centring), solids on or off, and the GMG solve in
FP32, bfloat16 and float16 storage, the AMG solve, and a batch of all of them.
The scene parameters are drawn from a fixed seed, so every run checks the same
cases."""
import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po

pytestmark = pytest.mark.gpu


def random_scene(seed):
    rng = np.random.default_rng([2026, seed])
    kind = int(rng.integers(4))
    kw = dict(seed=int(rng.integers(1 << 30)), has_solid=int(rng.integers(2)),
              solid_r=float(rng.uniform(0.08, 0.25)), solid_c=list(rng.uniform(0.35, 0.65, 3)))
    if rng.random() < 0.3:
        kw["nroot"] = [1, 1, 2]
    pure = rng.random() < 0.2
    kw["face_bc"] = [po.NEUMANN] * 6 if pure else [int(po.NEUMANN if rng.random() < 0.5 else po.DIRICHLET) for _ in range(6)]
    if not pure and po.DIRICHLET not in kw["face_bc"]:
        kw["face_bc"][int(rng.integers(6))] = po.DIRICHLET
    if kind == 0:
        kw.update(kind=po.SCENE_UNIFORM, uniform_level=int(rng.integers(1, 3)))
    elif kind == 1:
        top = int(rng.integers(3, 5))
        kw.update(kind=po.SCENE_SPHERE_BAND, band_r=float(rng.uniform(0.1, 0.2)), band_level=[top, top - 1, top - 2],
                  band_c=list(rng.uniform(0.4, 0.6, 3)))
    elif kind == 2:
        kw.update(kind=po.SCENE_PLANE, plane_z=float(rng.uniform(0.3, 0.7)), plane_level=[3, 2])
    else:
        kw.update(kind=po.SCENE_VORTEX, vortex_level=[int(rng.integers(3, 5)), 2], vortex_thresh=float(rng.uniform(0.3, 1.0)))
    return kw, pure


SEEDS = list(range(8))


def setup(seed):
    kw, pure = random_scene(seed)
    o = po.Oracle(**kw)
    lev, crd, typ, nbr = o.topology()
    tree = ca.Tree(o.params.h0, lev, crd, typ, o.celltypes())
    return o, tree, typ, o.get(po.CH_B), pure, kw


@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_every_storage(seed):
    o, tree, typ, b, pure, kw = setup(seed)
    P = ca.Problem(tree)
    m = typ == po.LEAF
    for storage in (ca.STORE_F32, ca.STORE_BF16, ca.STORE_F16):
        o.set(po.CH_B, b)
        P.set(ca.CH_B, b.astype(np.float32))
        ro = o.solve(max_iters=300, tol=1e-6, level_iters=2, coarsest_iters=10, pure_neumann=pure, storage=storage)
        S = ca.GMGSolver()
        S.solve(P, False, 300, 1e-6, 2, 10, 1, pure, storage=storage)
        assert (S.last["iters"], S.last["rr"], S.last["bb"]) == (ro["iters"], ro["rr"], ro["bb"]), (seed, storage, kw)
        assert np.array_equal(P.get(ca.CH_X)[m], o.get(po.CH_X).astype(np.float32)[m]), (seed, storage)


@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_amg(seed):
    o, tree, typ, b, pure, kw = setup(seed)
    P = ca.Problem(tree)
    o.set(po.CH_B, b)
    o.amg_prepare()
    ro = o.amg_solve(max_iters=300, tol=1e-6, level_iters=2, coarsest_iters=10, pure_neumann=pure)
    A = ca.AMGSolver()
    A.prepareTypesAndCoeffs(P)
    P.set(ca.CH_B, b.astype(np.float32))
    A.solve(P, False, 300, 1e-6, 2, 10, 1, pure)
    r = A.last_batch[0]
    assert (r["iters"], r["rr"]) == (ro["iters"], ro["rr"]), (seed, kw)
    m = typ == po.LEAF
    assert np.array_equal(P.get(ca.CH_X)[m], o.get(po.CH_X).astype(np.float32)[m]), seed


def test_random_scenes_as_one_batch():
    """The non-pure-Neumann scenes together in one handle (pure Neumann is a
    per-call flag): each result equals its oracle solve."""
    items = [setup(s) for s in SEEDS]
    items = [it for it in items if not it[4]]
    B = ca.Problem.batch([it[1] for it in items])
    B.set(ca.CH_B, np.concatenate([it[3].astype(np.float32) for it in items]))
    res = ca.GMGSolver().solve_batch(B, 300, 1e-6, 2, 10, storage=ca.STORE_BF16)
    xb = B.split(B.get(ca.CH_X))
    for q, (o, tree, typ, b, pure, kw) in enumerate(items):
        o.set(po.CH_B, b)
        ro = o.solve(max_iters=300, tol=1e-6, level_iters=2, coarsest_iters=10, storage=po.STORE_BF16)
        assert (res[q]["iters"], res[q]["rr"]) == (ro["iters"], ro["rr"]), kw
        m = typ == po.LEAF
        assert np.array_equal(xb[q][m], o.get(po.CH_X).astype(np.float32)[m])
