"""GPU parity of AMGSolver (src/AMGSolver.cu), the stored-coefficient
preconditioner: the prepared coefficients and AMG cell types, and whole
AMG-preconditioned solves (iterations, |r|^2, pressure) equal the oracle's
restatement bit for bit; a batched handle equals each problem alone."""
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


def prepared(kw):
    o = po.Oracle(**kw)
    lev, crd, typ, nbr = o.topology()
    tree = ca.Tree(o.params.h0, lev, crd, typ, o.celltypes())
    b = o.get(po.CH_B)
    o.amg_prepare()
    P = ca.Problem(tree)
    S = ca.AMGSolver(1.0, 1.0)
    S.prepareTypesAndCoeffs(P)
    return o, P, S, typ, b, tree


@pytest.mark.parametrize("name", list(SCENES))
def test_amg_coefficients_and_types_bitexact(name):
    o, P, S, typ, b, _ = prepared(SCENES[name])
    got = ca.AMGSolver.coefficients(P)
    want = o.amg_coeffs().astype(np.float32)
    for a in range(4):
        assert np.array_equal(got[a].view(np.uint32), want[a].view(np.uint32)), f"coefficient {a}"
    assert np.array_equal(got[4].astype(np.uint8), o.celltypes()), "AMG cell types"


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("nu", [1, 2])
def test_amg_solve_bitexact(name, nu):
    o, P, S, typ, b, _ = prepared(SCENES[name])
    o.set(po.CH_B, b)
    P.set(ca.CH_B, b.astype(np.float32))
    ro = o.amg_solve(max_iters=300, tol=1e-6, level_iters=nu, coarsest_iters=10)
    it, rel = S.solve(P, False, 300, 1e-6, nu, 10, 1, False)
    r = S.last_batch[0]
    assert r["converged"] == 1 and rel < 1e-6
    assert (r["iters"], r["rr"], r["bb"]) == (ro["iters"], ro["rr"], ro["bb"])
    m = typ == po.LEAF
    assert np.array_equal(P.get(ca.CH_X)[m], o.get(po.CH_X).astype(np.float32)[m]), "pressure differs from the oracle"


@pytest.mark.parametrize("use_graph,stride", [(False, 1), (True, 3)])
def test_amg_pure_neumann_and_launch_modes(use_graph, stride):
    o, P, S, typ, b, _ = prepared(ALL_NEUMANN)
    o.set(po.CH_B, b)
    P.set(ca.CH_B, b.astype(np.float32))
    ro = o.amg_solve(max_iters=300, tol=1e-6, level_iters=2, coarsest_iters=10, pure_neumann=True, sync_stride=stride)
    S.solve(P, False, 300, 1e-6, 2, 10, stride, True, use_graph=use_graph)
    r = S.last_batch[0]
    assert (r["iters"], r["rr"]) == (ro["iters"], ro["rr"])
    m = typ == po.LEAF
    assert np.array_equal(P.get(ca.CH_X)[m], o.get(po.CH_X).astype(np.float32)[m])


def test_amg_batch_equals_alone_and_gmg_unchanged():
    names = ["plane", "band", "vortex2root", "uniform32"]
    items = [prepared(SCENES[nm]) for nm in names]
    B = ca.Problem.batch([it[5] for it in items])
    B.set(ca.CH_B, np.concatenate([it[4].astype(np.float32) for it in items]))
    S = ca.AMGSolver()
    S.prepareTypesAndCoeffs(B)
    S.solve(B, False, 300, 1e-6, 2, 10, 1, False)
    res = S.last_batch
    xb = B.split(B.get(ca.CH_X))
    for q, (o, P, _, typ, b, _) in enumerate(items):
        o.set(po.CH_B, b)
        ro = o.amg_solve(max_iters=300, tol=1e-6, level_iters=2, coarsest_iters=10)
        assert (res[q]["iters"], res[q]["rr"]) == (ro["iters"], ro["rr"]), names[q]
        m = typ == po.LEAF
        assert np.array_equal(xb[q][m], o.get(po.CH_X).astype(np.float32)[m]), names[q]
    # the GMG solve on the same handle still uses the handle's own cell types
    g = po.Oracle(**SCENES["band"])
    lev, crd, typ, nbr = g.topology()
    P = items[1][1]
    P.set(ca.CH_B, items[1][4].astype(np.float32))
    rg = g.solve(max_iters=300, tol=1e-6, level_iters=2, coarsest_iters=10)
    G = ca.GMGSolver()
    G.solve(P, False, 300, 1e-6, 2, 10, 1, False)
    assert (G.last["iters"], G.last["rr"]) == (rg["iters"], rg["rr"])


def test_amg_solve_needs_prepare_and_fp32():
    o = po.Oracle(**SCENES["uniform32"])
    lev, crd, typ, nbr = o.topology()
    P = ca.Problem(ca.Tree(o.params.h0, lev, crd, typ, o.celltypes()))
    with pytest.raises(ca.CirrusError, match="prepare"):
        ca.AMGSolver().solve(P)
    p = ca.default_params(storage=ca.STORE_BF16)
    ca.AMGSolver().prepareTypesAndCoeffs(P)
    import ctypes
    r = (ca.Result * 1)()
    assert ca.lib().cg_amg_solve(P.handle, ctypes.byref(p), r) != 0
