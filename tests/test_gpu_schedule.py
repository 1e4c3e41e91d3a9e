"""Schedules change how the kernels cover the tiles, never a result: every
combination of LDS-resident coarse depth, tile order and tile block (tiles per
wavefront of the colour sweeps; lists of odd length leave one wavefront a
single tile) gives the oracle's solve bit for bit (FP32, bfloat16 and float16 storage,
where the LDS kernel rounds exactly where the grid-wide kernels store).  The autotuner picks among such
schedules and among smoother settings by timing solves of the handle's own
right-hand sides."""
import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po

pytestmark = pytest.mark.gpu

SCENES = {
    "uniform32": dict(uniform_level=2),
    "band": dict(kind=po.SCENE_SPHERE_BAND, solid_r=0.15, band_level=[4, 3, 2]),
    "vortex2root": dict(kind=po.SCENE_VORTEX, nroot=[1, 1, 2], vortex_level=[4, 2], seed=3),
}


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("storage", [ca.STORE_F32, ca.STORE_BF16, ca.STORE_F16])
def test_every_schedule_gives_the_oracle_solve(name, storage):
    o = po.Oracle(**SCENES[name])
    lev, crd, typ, nbr = o.topology()
    P = ca.Problem(ca.Tree(o.params.h0, lev, crd, typ, o.celltypes()))
    b = o.get(po.CH_B)
    ro = o.solve(max_iters=200, tol=1e-6, level_iters=2, coarsest_iters=10, storage=storage)
    xo = o.get(po.CH_X).astype(np.float32)[typ == po.LEAF]
    for shift in (0, 1, 2, 3):
        for swz in (True, False):
            for block in (1, 2, 4):
                P.set_schedule(shift, swz, block)
                P.set(ca.CH_B, b.astype(np.float32))
                S = ca.GMGSolver()
                S.solve(P, False, 200, 1e-6, 2, 10, 1, False, storage=storage)
                assert (S.last["iters"], S.last["rr"]) == (ro["iters"], ro["rr"]), (shift, swz, block)
                assert np.array_equal(P.get(ca.CH_X)[typ == po.LEAF], xo), (shift, swz, block)


def test_tile_block_on_a_batch():
    """A batched pool with two tiles per wavefront: each problem's pressure
    equals its solve with one tile per wavefront, bit for bit."""
    trees, bs = [], []
    for kw in SCENES.values():
        sc = ca.Scene(**kw)
        trees.append(sc.tree)
        bs.append(sc.rhs.astype(np.float32))
    b = np.concatenate(bs)
    out = []
    for block in (1, 2, 4):
        B = ca.Problem.batch(trees)
        B.set_schedule(0, True, block)
        B.set(ca.CH_B, b)
        res = ca.GMGSolver().solve_batch(B, 1000, 1e-6, 2, 10, storage=ca.STORE_BF16)
        out.append(([(r["iters"], r["rr"]) for r in res], B.get(ca.CH_X)))
    for q in (1, 2):
        assert out[0][0] == out[q][0]
        assert np.array_equal(out[0][1], out[q][1])


def test_autotune_on_a_batch():
    trees, bs = [], []
    for kw in SCENES.values():
        sc = ca.Scene(**kw)
        trees.append(sc.tree)
        bs.append(sc.rhs.astype(np.float32))
    B = ca.Problem.batch(trees)
    b = np.concatenate(bs)
    B.set(ca.CH_B, b)
    S = ca.GMGSolver()
    kw, rep = S.autotune(B, 1000, 1e-6, 2, 10, allow_bf16=True)
    assert rep["candidates"] >= 8 and rep["best_seconds"] <= rep["base_seconds"]
    assert rep["tile_block"] in (1, 2, 4, 8)
    assert np.array_equal(B.get(ca.CH_B), b)  # the right-hand sides are restored
    res = S.solve_batch(B, 1000, 1e-6, coarsest_iters=10, verify_f64=True, **kw)
    assert all(r["converged"] == 1 and r["rel_res_f64"] < 5e-6 for r in res)


def test_autotune_refined_on_a_batch():
    """cg_autotune_refined times the FP64-refined solve; its choice then
    refines every problem to the tolerance and the right-hand sides are back."""
    trees, bs = [], []
    for kw in SCENES.values():
        sc = ca.Scene(**kw)
        trees.append(sc.tree)
        bs.append(sc.rhs.astype(np.float32))
    B = ca.Problem.batch(trees)
    b = np.concatenate(bs)
    B.set(ca.CH_B, b)
    S = ca.GMGSolver()
    kw, rep = S.autotune_refined(B, 1e-6, 5e-7, 20, 1000, 2, 10, storage=ca.STORE_BF16, allow_16bit=True)
    assert rep["candidates"] >= 8 and rep["best_seconds"] <= rep["base_seconds"]
    assert np.array_equal(B.get(ca.CH_B), b)
    res = S.solve_refined(B, 1e-6, 5e-7, 20, 1000, kw["level_iters"], 10, smoother=kw["smoother"],
                          storage=kw["storage"])
    assert all(r["converged"] == 1 and r["rel_res"] < 1e-6 for r in res)
