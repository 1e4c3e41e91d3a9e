"""GPU parity of the mixed-precision form.

BF16 storage: the MGPCG whose V-cycle vectors (x on every level, its rhs and
scratch) are stored as bfloat16 -- FP32 arithmetic, every store rounded to
nearest even -- equals the oracle's solve with the same rounding at the same
stores, bit for bit (iterations, |r|^2, pressure).  F16 storage (float16
stores, clamped to +-65504, with the right-hand side scaled by a power of two
for the solve) likewise.

Iterative refinement: x and the residual in FP64 around that MGPCG.  The whole
outer loop equals the oracle's SolveRefined bit for bit (steps, iterations, the
FP64 residual and the FP64 x), reaches its FP64 tolerance, and its error
against the FP64 oracle's converged pressure is stated for every scene.
"""
import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po

pytestmark = pytest.mark.gpu

SCENES = {
    "uniform32": dict(uniform_level=2),
    "plain64": dict(),
    "plane": dict(kind=po.SCENE_PLANE, plane_level=[3, 2]),
    "band": dict(kind=po.SCENE_SPHERE_BAND, solid_r=0.15, band_level=[4, 3, 2]),
    "vortex2root": dict(kind=po.SCENE_VORTEX, nroot=[1, 1, 2], vortex_level=[4, 2], seed=3),
}
ALL_NEUMANN = dict(kind=po.SCENE_SPHERE_BAND, solid_r=0.15, band_level=[4, 3, 2], face_bc=[po.NEUMANN] * 6)
# relative L2 error of the refined FP64 pressure against the FP64 oracle's
# (solved to 1e-13) at an outer tolerance of 1e-8
REFINED_ERR_TOL = 1e-5
_cache = {}


def pair(name, kw=None):
    if name not in _cache:
        o = po.Oracle(**(kw or SCENES[name]))
        lev, crd, typ, nbr = o.topology()
        tree = ca.Tree(o.params.h0, lev, crd, typ, o.celltypes())
        _cache[name] = (o, ca.Problem(tree), typ, o.get(po.CH_B), tree)
    return _cache[name]


def set_b(o, P, b):
    o.set(po.CH_B, b)
    P.set(ca.CH_B, b.astype(np.float32))


def leaf_eq(a, b, typ):
    m = typ == po.LEAF
    return np.array_equal(np.asarray(a)[m], np.asarray(b)[m])


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("smoother,nu", [(ca.SMOOTH_RBGS, 2), (ca.SMOOTH_JACOBI, 2), (ca.SMOOTH_RBGS, 1)])
@pytest.mark.parametrize("storage", [ca.STORE_BF16, ca.STORE_F16])
def test_16bit_storage_solve_bitexact(name, smoother, nu, storage):
    o, P, typ, b, _ = pair(name)
    set_b(o, P, b)
    ro = o.solve(max_iters=200, tol=1e-6, level_iters=nu, coarsest_iters=10, smoother=smoother, storage=storage)
    S = ca.GMGSolver()
    S.solve(P, False, 200, 1e-6, nu, 10, 1, False, smoother=smoother, storage=storage)
    r = S.last
    assert r["converged"] == 1
    assert (r["iters"], r["rr"], r["bb"]) == (ro["iters"], ro["rr"], ro["bb"])
    assert leaf_eq(P.get(ca.CH_X), o.get(po.CH_X).astype(np.float32), typ), "pressure differs from the oracle"


@pytest.mark.parametrize("use_graph,stride", [(False, 1), (True, 3)])
@pytest.mark.parametrize("storage", [ca.STORE_BF16, ca.STORE_F16])
def test_16bit_storage_launch_modes_and_pure_neumann(use_graph, stride, storage):
    o, P, typ, b, _ = pair("allneumann", ALL_NEUMANN)
    set_b(o, P, b)
    ro = o.solve(max_iters=200, tol=1e-6, level_iters=2, coarsest_iters=10, pure_neumann=True, sync_stride=stride,
                 storage=storage)
    S = ca.GMGSolver()
    S.solve(P, False, 200, 1e-6, 2, 10, stride, True, use_graph=use_graph, storage=storage)
    assert (S.last["iters"], S.last["rr"]) == (ro["iters"], ro["rr"])
    assert leaf_eq(P.get(ca.CH_X), o.get(po.CH_X).astype(np.float32), typ)


def test_f32_solve_unchanged_after_bf16_solve():
    """Both storage types on one handle: the FP32 solve still equals the
    reference transcription."""
    o, P, typ, b, _ = pair("band")
    for storage in (ca.STORE_BF16, ca.STORE_F32, ca.STORE_F16, ca.STORE_BF16, ca.STORE_F32):
        set_b(o, P, b)
        ro = o.solve(max_iters=200, tol=1e-6, level_iters=2, coarsest_iters=10, storage=storage)
        S = ca.GMGSolver()
        S.solve(P, False, 200, 1e-6, 2, 10, 1, False, storage=storage)
        assert (S.last["iters"], S.last["rr"]) == (ro["iters"], ro["rr"]), storage


def fp64_error(o, typ, b, x64):
    o64 = po.Oracle(params=o.params, fp64=True)
    o64.set(po.CH_B, b.astype(np.float64))
    r64 = o64.solve(max_iters=500, tol=1e-13, level_iters=2, coarsest_iters=10)
    assert r64["rel_res"] < 1e-12
    xr = o64.get(po.CH_X)
    m = (typ == po.LEAF)[:, None] & (o.celltypes() == po.INTERIOR)
    return np.linalg.norm((x64 - xr)[m]) / np.linalg.norm(xr[m])


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("storage", [ca.STORE_BF16, ca.STORE_F16])
def test_refined_bitexact_and_fp64_error(name, storage):
    o, P, typ, b, _ = pair(name)
    set_b(o, P, b)
    ro = o.solve_refined(tol=1e-8, inner_tol=1e-4, level_iters=2, coarsest_iters=10, storage=storage)
    S = ca.GMGSolver()
    r = S.solve_refined(P, 1e-8, 1e-4, 20, 1000, 2, 10, storage=storage)[0]
    assert r["converged"] == 1 and r["rel_res"] < 1e-8
    assert (r["outer"], r["iters"], r["rr"], r["bb"]) == (ro["outer"], ro["iters"], ro["rr"], ro["bb"])
    x64 = P.get_f64()
    assert leaf_eq(x64, o.refined_x(), typ), "FP64 pressure differs from the oracle"
    assert leaf_eq(P.get(ca.CH_X), x64.astype(np.float32), typ)
    err = fp64_error(o, typ, b, x64)
    print(f"\n{name} storage {storage}: {r['outer']} refinement steps, {r['iters']} CG iterations, FP64 residual {r['rel_res']:.2e}, "
          f"error vs FP64 oracle {err:.2e}")
    assert err < REFINED_ERR_TOL, err


@pytest.mark.parametrize("storage,smoother,pure", [(ca.STORE_F32, ca.SMOOTH_RBGS, False),
                                                   (ca.STORE_BF16, ca.SMOOTH_JACOBI, False),
                                                   (ca.STORE_BF16, ca.SMOOTH_RBGS, True)])
def test_refined_variants_bitexact(storage, smoother, pure):
    name, kw = ("allneumann", ALL_NEUMANN) if pure else ("band", None)
    o, P, typ, b, _ = pair(name, kw)
    set_b(o, P, b)
    ro = o.solve_refined(tol=1e-9, inner_tol=1e-3, level_iters=2, coarsest_iters=10, smoother=smoother,
                         pure_neumann=pure, storage=storage)
    r = ca.GMGSolver().solve_refined(P, 1e-9, 1e-3, 20, 1000, 2, 10, is_pure_neumann=pure, smoother=smoother,
                                     storage=storage)[0]
    assert r["converged"] == 1
    assert (r["outer"], r["iters"], r["rr"]) == (ro["outer"], ro["iters"], ro["rr"])
    assert leaf_eq(P.get_f64(), o.refined_x(), typ)


def test_refined_outer_cap_and_zero_rhs():
    o, P, typ, b, _ = pair("plane")
    set_b(o, P, b)
    ro = o.solve_refined(tol=1e-12, inner_tol=1e-2, max_outer=2, level_iters=2, coarsest_iters=10)
    r = ca.GMGSolver().solve_refined(P, 1e-12, 1e-2, 2, 1000, 2, 10)[0]
    assert (r["outer"], r["converged"]) == (2, 0)
    assert (r["iters"], r["rr"]) == (ro["iters"], ro["rr"])
    z = np.zeros_like(b)
    set_b(o, P, z)
    r = ca.GMGSolver().solve_refined(P, 1e-8, 1e-4, 20, 1000, 2, 10)[0]
    assert (r["outer"], r["iters"], r["converged"], r["rel_res"]) == (0, 0, 1, 0.0)
    assert not P.get_f64()[typ == po.LEAF].any()


@pytest.mark.parametrize("storage", [ca.STORE_BF16, ca.STORE_F16])
def test_refined_batch_equals_alone(storage):
    """Refinement on a batched handle: every problem runs its own outer loop and
    equals the oracle and its solve alone (with float16 storage each problem
    has its own right-hand-side scale)."""
    names = ["plane", "band", "vortex2root", "uniform32"]
    trees, bs = [], []
    for nm in names:
        o, P, typ, b, tree = pair(nm)
        trees.append(tree)
        bs.append(b.astype(np.float32))
    B = ca.Problem.batch(trees)
    B.set(ca.CH_B, np.concatenate(bs))
    res = ca.GMGSolver().solve_refined(B, 1e-8, 1e-4, 20, 1000, 2, 10, storage=storage)
    xb = B.split(B.get_f64())
    for q, nm in enumerate(names):
        o, P, typ, b, tree = pair(nm)
        set_b(o, P, b)
        ro = o.solve_refined(tol=1e-8, inner_tol=1e-4, level_iters=2, coarsest_iters=10, storage=storage)
        assert (res[q]["outer"], res[q]["iters"], res[q]["rr"]) == (ro["outer"], ro["iters"], ro["rr"]), nm
        assert leaf_eq(xb[q], o.refined_x(), typ), nm


def test_server_bf16_cold_requests_equal_solve():
    """The serving loop with BF16 storage: cold requests equal the BF16 solve alone."""
    kws = [SCENES["plane"], SCENES["band"]]
    scs = [ca.Scene(**kw) for kw in kws]
    S = ca.Server(1 << 30, 8, level_iters=2, coarsest_iters=10, storage=ca.STORE_BF16)
    inst = [S.admit(sc.tree) for sc in scs]
    for q, (i, sc) in enumerate(zip(inst, scs)):
        S.submit(i, sc.rhs, tag=q)
    got = {}
    for _ in range(2000):
        S.step()
        while True:
            r = S.next_result()
            if r is None:
                break
            got[r[0]] = r
        if len(got) == len(scs):
            break
    assert len(got) == len(scs)
    for q, sc in enumerate(scs):
        P = ca.Problem(sc.tree)
        P.set(ca.CH_B, sc.rhs)
        G = ca.GMGSolver()
        G.solve(P, False, 1000, 1e-6, 2, 10, 1, False, storage=ca.STORE_BF16)
        tag, instance, res, x = got[q]
        assert (res["iters"], res["rr"]) == (G.last["iters"], G.last["rr"]), q
        leaf = sc.tree.type == po.LEAF
        assert np.array_equal(x[leaf], P.get(ca.CH_X)[leaf]), q


def test_f16_batch_solve_equals_alone_and_server_refuses_f16():
    """The float16 solve on a batched handle: every problem (its own scale)
    equals the oracle; the serving loop does not offer float16 storage."""
    names = ["plane", "band", "uniform32"]
    trees, bs = [], []
    for nm in names:
        o, P, typ, b, tree = pair(nm)
        trees.append(tree)
        bs.append(b.astype(np.float32))
    B = ca.Problem.batch(trees)
    B.set(ca.CH_B, np.concatenate(bs))
    res = ca.GMGSolver().solve_batch(B, 200, 1e-6, 2, 10, storage=ca.STORE_F16, verify_f64=True)
    xb = B.split(B.get(ca.CH_X))
    for q, nm in enumerate(names):
        o, P, typ, b, tree = pair(nm)
        set_b(o, P, b)
        ro = o.solve(max_iters=200, tol=1e-6, level_iters=2, coarsest_iters=10, storage=po.STORE_F16)
        assert (res[q]["iters"], res[q]["rr"], res[q]["bb"]) == (ro["iters"], ro["rr"], ro["bb"]), nm
        assert leaf_eq(xb[q], o.get(po.CH_X).astype(np.float32), typ), nm
        assert 0 < res[q]["rel_res_f64"] < 2e-6
    with pytest.raises(ca.CirrusError, match="float16"):
        ca.Server(1 << 28, 4, storage=ca.STORE_F16)
