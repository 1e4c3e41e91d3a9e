"""GPU parity of the batched form: several independent trees in one handle,
solved together by one sequence of launches per CG iteration.

Each problem's arithmetic touches only its own tiles and its own scalars, and
its reductions run in the same fixed order as when it is alone, so every result
must equal -- bit for bit -- both the oracle's solve of that problem and the
library's solve of the same problem alone.
"""
import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po

pytestmark = pytest.mark.gpu

SCENES = [
    dict(uniform_level=2),
    dict(kind=po.SCENE_PLANE, plane_level=[3, 2]),
    dict(kind=po.SCENE_SPHERE_BAND, solid_r=0.15, band_level=[4, 3, 2]),
    dict(kind=po.SCENE_VORTEX, nroot=[1, 1, 2], vortex_level=[4, 2], seed=3),
    dict(),
    dict(kind=po.SCENE_SPHERE_BAND, solid_r=0.2, band_level=[4, 3, 2], seed=11,
         face_bc=[po.DIRICHLET, po.NEUMANN, po.NEUMANN, po.NEUMANN, po.DIRICHLET, po.NEUMANN]),
]
_cache = {}


def oracles():
    if "o" not in _cache:
        os_, trees, bs = [], [], []
        for kw in SCENES:
            o = po.Oracle(**kw)
            lev, crd, typ, nbr = o.topology()
            trees.append(ca.Tree(o.params.h0, lev, crd, typ, o.celltypes()))
            bs.append(o.get(po.CH_B).astype(np.float32))
            os_.append(o)
        _cache["o"] = (os_, trees, bs)
    return _cache["o"]


def leaf_mask(tree):
    return tree.type == po.LEAF


@pytest.mark.parametrize("smoother,nu,stride", [(ca.SMOOTH_RBGS, 2, 1), (ca.SMOOTH_JACOBI, 2, 1), (ca.SMOOTH_RBGS, 1, 3)])
def test_batch_equals_oracle_and_alone(smoother, nu, stride):
    os_, trees, bs = oracles()
    B = ca.Problem.batch(trees)
    assert B.num_problems == len(trees)
    B.set(ca.CH_B, np.concatenate(bs))
    S = ca.GMGSolver()
    res = S.solve_batch(B, 200, 1e-6, nu, 10, stride, False, smoother=smoother)
    xb = B.split(B.get(ca.CH_X))
    for q, (o, tree, b) in enumerate(zip(os_, trees, bs)):
        o.set(po.CH_B, b.astype(np.float64))
        ro = o.solve(max_iters=200, tol=1e-6, level_iters=nu, coarsest_iters=10, smoother=smoother, sync_stride=stride)
        assert res[q]["converged"] == 1, q
        assert res[q]["iters"] == ro["iters"] and res[q]["rr"] == ro["rr"] and res[q]["bb"] == ro["bb"], q
        m = leaf_mask(tree)
        ox = o.get(po.CH_X).astype(np.float32)
        assert np.array_equal(xb[q][m], ox[m]), f"problem {q}: pressure differs from the oracle"
        # the same problem alone
        P = ca.Problem(tree)
        P.set(ca.CH_B, b)
        S1 = ca.GMGSolver()
        S1.solve(P, False, 200, 1e-6, nu, 10, stride, False, smoother=smoother)
        assert S1.last["iters"] == res[q]["iters"] and S1.last["rr"] == res[q]["rr"], q
        assert np.array_equal(P.get(ca.CH_X)[m], xb[q][m]), f"problem {q}: batch and alone differ"


def test_batch_with_a_zero_rhs_and_repeats():
    """A problem with b = 0 stops at once without holding up the others; a
    second solve of the same handle repeats the first exactly."""
    os_, trees, bs = oracles()
    B = ca.Problem.batch(trees[:3])
    b = [bs[0], np.zeros_like(bs[1]), bs[2]]
    S = ca.GMGSolver()
    B.set(ca.CH_B, np.concatenate(b))
    r1 = S.solve_batch(B, 200, 1e-6, 2, 10)
    x1 = B.get(ca.CH_X)
    assert (r1[1]["iters"], r1[1]["rel_res"]) == (0, 0.0)
    assert r1[0]["converged"] == 1 and r1[2]["converged"] == 1
    B.set(ca.CH_B, np.concatenate(b))
    r2 = S.solve_batch(B, 200, 1e-6, 2, 10)
    assert [r["iters"] for r in r1] == [r["iters"] for r in r2]
    assert [r["rr"] for r in r1] == [r["rr"] for r in r2]
    assert np.array_equal(x1, B.get(ca.CH_X))


def test_batch_operator_api_covers_every_problem():
    """FullNegativeLaplacian on a batched handle equals the oracle per problem."""
    os_, trees, bs = oracles()
    B = ca.Problem.batch(trees)
    xs = [np.random.default_rng(q).standard_normal((t.n, 512)).astype(np.float32) for q, t in enumerate(trees)]
    B.set(ca.CH_X, np.concatenate(xs))
    ca.FullNegativeLaplacian(B, ca.CH_X, ca.CH_AP)
    ab = B.split(B.get(ca.CH_AP))
    for q, (o, x) in enumerate(zip(os_, xs)):
        o.set(po.CH_X, x.astype(np.float64))
        o.set(po.CH_AP, np.zeros((o.n, 512)))
        o.full_negative_laplacian(po.CH_X, po.CH_AP)
        assert np.array_equal(ab[q], o.get(po.CH_AP).astype(np.float32)), q
