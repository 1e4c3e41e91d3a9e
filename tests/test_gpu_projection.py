"""GPU parity of the projection around the solve (FluidEuler::project,
src/FluidEuler.h:581-612): the divergence of the face velocities
(VelocityVolumeDivergenceOnLeafs with the ghost velocities taken from the
parents), the pressure gradient added to them (AddGradientToFaceCenters), and
the whole projection, each against the oracle bit for bit."""
import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po

pytestmark = pytest.mark.gpu

SCENES = {
    "plain64": dict(),
    "plane": dict(kind=po.SCENE_PLANE, plane_level=[3, 2]),
    "band": dict(kind=po.SCENE_SPHERE_BAND, solid_r=0.15, band_level=[4, 3, 2]),
    "vortex2root": dict(kind=po.SCENE_VORTEX, nroot=[1, 1, 2], vortex_level=[4, 2], seed=3),
}
_cache = {}


def pair(name):
    if name not in _cache:
        o = po.Oracle(**SCENES[name])
        lev, crd, typ, nbr = o.topology()
        P = ca.Problem(ca.Tree(o.params.h0, lev, crd, typ, o.celltypes()))
        _cache[name] = (o, P, typ)
    return _cache[name]


def both_set(o, P, ch, a):
    a32 = np.asarray(a, np.float32)
    o.set(ch, a32.astype(np.float64))
    P.set(ch, a32)


def assert_same(dev, ora, what):
    ora = np.asarray(ora, np.float64).astype(np.float32)
    bad = np.asarray(dev) != ora
    assert not bad.any(), f"{what}: {bad.sum()} of {bad.size} values differ"


@pytest.mark.parametrize("name", list(SCENES))
def test_divergence_bitexact(name):
    o, P, typ = pair(name)
    rng = np.random.default_rng(11)
    for a in range(3):
        both_set(o, P, po.CH_U + a, rng.standard_normal((P.n, 512)))
    o.velocity_divergence(po.CH_TMP)
    ca.VelocityVolumeDivergenceOnLeafs(P, ca.CH_TMP)
    m = typ != po.NONLEAF
    assert_same(P.get(ca.CH_TMP)[m], o.get(po.CH_TMP)[m], "divergence")
    for a in range(3):  # ghost velocities from the parents
        assert_same(P.get(ca.CH_U + a), o.get(po.CH_U + a), f"velocity {a}")


@pytest.mark.parametrize("name", list(SCENES))
def test_gradient_bitexact(name):
    o, P, typ = pair(name)
    rng = np.random.default_rng(12)
    for a in range(3):
        both_set(o, P, po.CH_U + a, rng.standard_normal((P.n, 512)))
    pv = rng.standard_normal((P.n, 512))
    pv[typ != po.LEAF] = 0
    both_set(o, P, po.CH_X, pv)
    o.add_gradient(po.CH_X)
    ca.AddGradientToFaceCenters(P, ca.CH_X)
    for a in range(3):
        assert_same(P.get(ca.CH_U + a), o.get(po.CH_U + a), f"velocity {a}")
    assert_same(P.get(ca.CH_X), o.get(po.CH_X), "pressure with ghost values")


@pytest.mark.parametrize("name", ["plane", "band", "vortex2root"])
def test_projection_bitexact(name):
    """The scene's face velocities projected: divergence, MGPCG solve, gradient."""
    o, P, typ = pair(name)
    o_fresh = po.Oracle(**SCENES[name])  # its channels 6-8 hold the scene's velocities
    uvw = [o_fresh.get(po.CH_U + a) for a in range(3)]
    for a in range(3):
        both_set(o, P, po.CH_U + a, uvw[a])
    o.velocity_divergence(po.CH_B)
    ro = o.solve(max_iters=200, tol=1e-6, level_iters=2, coarsest_iters=10)
    o.add_gradient(po.CH_X)
    r = ca.GMGSolver().project(P, 200, 1e-6, 2, 10)[0]
    assert r["converged"] == 1 and (r["iters"], r["rr"]) == (ro["iters"], ro["rr"])
    for a in range(3):
        assert_same(P.get(ca.CH_U + a), o.get(po.CH_U + a), f"projected velocity {a}")
    # the divergence of the projected field, on the LEAF cells that do not touch
    # a ghost tile, is the solve's residual
    ca.VelocityVolumeDivergenceOnLeafs(P, ca.CH_TMP, propagate_ghosts=False)
    d = P.get(ca.CH_TMP).astype(np.float64)
    b = o_fresh.get(po.CH_B)
    assert np.linalg.norm(d[typ == po.LEAF]) < np.linalg.norm(b[typ == po.LEAF])


def test_projection_on_a_batch():
    """cg_project on a batched handle: every problem's velocities equal its
    projection alone."""
    names = ["plane", "band"]
    trees, us = [], []
    for nm in names:
        o = po.Oracle(**SCENES[nm])
        lev, crd, typ, nbr = o.topology()
        trees.append(ca.Tree(o.params.h0, lev, crd, typ, o.celltypes()))
        us.append([o.get(po.CH_U + a).astype(np.float32) for a in range(3)])
    B = ca.Problem.batch(trees)
    for a in range(3):
        B.set(ca.CH_U + a, np.concatenate([u[a] for u in us]))
    res = ca.GMGSolver().project(B, 200, 1e-6, 2, 10)
    for q, (tree, u) in enumerate(zip(trees, us)):
        P = ca.Problem(tree)
        for a in range(3):
            P.set(ca.CH_U + a, u[a])
        r = ca.GMGSolver().project(P, 200, 1e-6, 2, 10)[0]
        assert (res[q]["iters"], res[q]["rr"]) == (r["iters"], r["rr"])
        for a in range(3):
            assert np.array_equal(B.split(B.get(ca.CH_U + a))[q], P.get(ca.CH_U + a)), (q, a)
