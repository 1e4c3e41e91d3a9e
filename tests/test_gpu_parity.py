"""GPU parity: every HIP kernel and the full MGPCG solve against the oracle.

The FP32 oracle (the CPU transcription) and the HIP path use the same operation
order, so results are compared for exact equality; the pressure is also checked
against the FP64 oracle (relative L2 error) and its true residual recomputed in
FP64.  Inputs are the oracle's scenes (tree, cell types, divergence RHS) handed
to the library through the C ABI in canonical tile order.
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
_cache = {}


def pair(name):
    """(oracle FP32, device problem, tile types, neighbours, the scene's b).
    Both are cached per scene; kernel tests overwrite their channels, so b is
    kept as built and set again by each solve."""
    if name not in _cache:
        o = po.Oracle(**SCENES[name])
        lev, crd, typ, nbr = o.topology()
        tree = ca.Tree(o.params.h0, lev, crd, typ, o.celltypes())
        P = ca.Problem(tree)
        _cache[name] = (o, P, typ, nbr, o.get(po.CH_B))
    o, P, typ, nbr, b = _cache[name]
    return o, P, typ, nbr, b.copy()


def both_set(o, P, ch, a):
    a32 = np.asarray(a, np.float32)
    o.set(ch, a32.astype(np.float64))
    P.set(ch, a32)


def rnd(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


def assert_same(dev, ora, what, mask=None):
    dev = np.asarray(dev, np.float32)
    ora = np.asarray(ora, np.float64).astype(np.float32)
    if mask is not None:
        dev, ora = dev[mask], ora[mask]
    bad = dev != ora
    assert not bad.any(), f"{what}: {bad.sum()} of {bad.size} values differ, max |d| {np.abs(dev - ora).max():.3e}"


@pytest.mark.parametrize("name", list(SCENES))
def test_topology_matches_oracle(name):
    o, P, typ, nbr, b = pair(name)
    dnbr, dpar = P.topology()
    assert np.array_equal(dnbr, nbr)
    lev, crd, _, _ = o.topology()
    for t in np.nonzero(lev > 0)[0][:2000]:
        p = dpar[t]
        assert lev[p] == lev[t] - 1 and np.array_equal(crd[p], crd[t] // 2)


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("diag", [False, True])
def test_full_negative_laplacian_bitexact(name, diag):
    o, P, typ, nbr, b = pair(name)
    x = rnd((P.n, 512), 1)
    both_set(o, P, po.CH_X, x)
    both_set(o, P, po.CH_AP, np.zeros_like(x))
    o.full_negative_laplacian(po.CH_X, po.CH_AP, diag)
    ca.FullNegativeLaplacian(P, ca.CH_X, ca.CH_AP, diag)
    assert_same(P.get(ca.CH_AP), o.get(po.CH_AP), "Ax")
    assert_same(P.get(ca.CH_X), o.get(po.CH_X), "ghost-propagated x")


@pytest.mark.parametrize("name", ["plane", "band"])
def test_negative_laplacian_same_level_bitexact(name):
    o, P, typ, nbr, b = pair(name)
    for level in range(1, P.max_level + 1):
        x = rnd((P.n, 512), 10 + level)
        both_set(o, P, po.CH_X, x)
        both_set(o, P, po.CH_AP, np.zeros_like(x))
        o.negative_laplacian_same_level(level, level, po.LEAF | po.NONLEAF | po.GHOST, po.CH_X, po.CH_AP)
        ca.NegativeLaplacianSameLevel(P, level, level, ca.LEAF | ca.NONLEAF | ca.GHOST, ca.CH_X, ca.CH_AP)
        assert_same(P.get(ca.CH_AP), o.get(po.CH_AP), f"level {level} Ax")


@pytest.mark.parametrize("name", list(SCENES))
def test_gauss_seidel_bitexact(name):
    o, P, typ, nbr, b = pair(name)
    for level in range(P.max_level + 1):
        both_set(o, P, po.CH_Z, rnd((P.n, 512), 20 + level))
        both_set(o, P, po.CH_TMP, rnd((P.n, 512), 40 + level))
        for color in (0, 1, 1, 0):
            o.gs_color(level, po.CH_Z, po.CH_TMP, color)
            ca.GaussSeidelColor(P, level, ca.CH_Z, ca.CH_TMP, color)
        assert_same(P.get(ca.CH_Z), o.get(po.CH_Z), f"GS level {level}")


@pytest.mark.parametrize("name", ["plain64", "band"])
def test_jacobi_bitexact(name):
    o, P, typ, nbr, b = pair(name)
    for level in range(P.max_level + 1):
        both_set(o, P, po.CH_Z, rnd((P.n, 512), 60 + level))
        both_set(o, P, po.CH_TMP, rnd((P.n, 512), 70 + level))
        for _ in range(2):
            o.jacobi_sweep(level, po.CH_Z, po.CH_TMP, 2.0 / 3.0)
            ca.JacobiSweep(P, level, ca.CH_Z, ca.CH_TMP, 2.0 / 3.0)
        assert_same(P.get(ca.CH_Z), o.get(po.CH_Z), f"Jacobi level {level}")


@pytest.mark.parametrize("name", list(SCENES))
def test_restrict_prolong_bitexact(name):
    o, P, typ, nbr, b = pair(name)
    for level in range(1, P.max_level + 1):
        both_set(o, P, po.CH_Z, rnd((P.n, 512), 80 + level))
        both_set(o, P, po.CH_TMP, rnd((P.n, 512), 90 + level))
        both_set(o, P, po.CH_AP, rnd((P.n, 512), 100 + level))
        o.residual_restrict(level, po.CH_Z, po.CH_TMP, po.CH_AP)
        ca.ResidualAndRestrict(P, level, ca.CH_Z, ca.CH_TMP, ca.CH_AP)
        assert_same(P.get(ca.CH_AP), o.get(po.CH_AP), f"restrict level {level}")
        o.prolongate(level, po.CH_AP, po.CH_Z)
        ca.ProlongateAndUpdate(P, level, ca.CH_AP, ca.CH_Z)
        assert_same(P.get(ca.CH_Z), o.get(po.CH_Z), f"prolong level {level}")


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("smoother,nu", [(ca.SMOOTH_RBGS, 1), (ca.SMOOTH_RBGS, 2), (ca.SMOOTH_JACOBI, 2),
                                         (ca.SMOOTH_JACOBI, 1)])
def test_vcycle_bitexact(name, smoother, nu):
    o, P, typ, nbr, b = pair(name)
    f = rnd((P.n, 512), 7)
    f[typ != po.LEAF] = 0
    both_set(o, P, po.CH_R, f)
    o.vcycle(po.CH_Z, po.CH_R, po.CH_AP, nu, 10, smoother, 2.0 / 3.0)
    ca.GMGSolver().VCycle(P, ca.CH_Z, ca.CH_R, nu, 10, smoother, 2.0 / 3.0)
    leaf = typ == po.LEAF
    assert_same(P.get(ca.CH_Z), o.get(po.CH_Z), "V-cycle z on LEAF tiles", mask=leaf)


@pytest.mark.parametrize("name", ["plain64", "band"])
def test_dot_bitexact(name):
    o, P, typ, nbr, b = pair(name)
    both_set(o, P, po.CH_P, rnd((P.n, 512), 5))
    both_set(o, P, po.CH_Z, rnd((P.n, 512), 6))
    for types in (po.LEAF, po.LEAF | po.NONLEAF | po.GHOST):
        assert P.dot(ca.CH_P, ca.CH_Z, types) == o.dot(po.CH_P, po.CH_Z, types)


def _solve_both(name, smoother, nu, use_graph=True, sync_stride=1, pure=False, scene=None, max_iters=200):
    if scene is None:
        o, P, typ, nbr, b = pair(name)
    else:
        o, P, typ, nbr, b = scene
    both_set(o, P, po.CH_B, b)
    ro = o.solve(max_iters=max_iters, tol=1e-6, level_iters=nu, coarsest_iters=10, pure_neumann=pure, smoother=smoother,
                 sync_stride=sync_stride)
    S = ca.GMGSolver(1.0, 1.0)
    it, rel = S.solve(P, False, max_iters, 1e-6, nu, 10, sync_stride, pure, smoother=smoother, use_graph=use_graph)
    assert (it, rel) == (S.last["iters"], S.last["rel_res"])
    S.last["b"] = b
    return o, P, typ, ro, S.last


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("smoother,nu", [(ca.SMOOTH_RBGS, 1), (ca.SMOOTH_RBGS, 2), (ca.SMOOTH_JACOBI, 2)])
def test_solve_bitexact(name, smoother, nu):
    o, P, typ, ro, rd = _solve_both(name, smoother, nu)
    assert rd["converged"] == 1
    assert rd["iters"] == ro["iters"]
    assert rd["rr"] == ro["rr"] and rd["bb"] == ro["bb"]
    leaf = typ == po.LEAF
    assert_same(P.get(ca.CH_X), o.get(po.CH_X), "pressure", mask=leaf)


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("nu", [2, 3])
def test_solve_bitexact_one_launch_iterations(name, nu):
    """Red-black iterations after the first as one launch of both colours
    (cg_set_level_forms pair_sweeps = 1; the values alternate between two
    arrays): the same solve bit for bit."""
    o, P, typ, nbr, b = pair(name)
    P.set_level_forms(-1, -1, 1)
    try:
        o, P, typ, ro, rd = _solve_both(name, ca.SMOOTH_RBGS, nu)
    finally:
        P.set_level_forms(-1, -1, -1)
    assert rd["converged"] == 1
    assert (rd["iters"], rd["rr"], rd["bb"]) == (ro["iters"], ro["rr"], ro["bb"])
    assert_same(P.get(ca.CH_X), o.get(po.CH_X), "pressure", mask=typ == po.LEAF)


@pytest.mark.parametrize("use_graph,stride", [(False, 1), (True, 4), (False, 3)])
def test_solve_launch_modes_and_sync_stride(use_graph, stride):
    """Graph replay or direct launches; the residual tested every `stride`
    iterations, as GMGSolver::solve does (the count is a multiple of it)."""
    o, P, typ, ro, rd = _solve_both("band", ca.SMOOTH_RBGS, 2, use_graph=use_graph, sync_stride=stride)
    assert rd["iters"] == ro["iters"] and rd["rr"] == ro["rr"] and rd["rel_res"] == ro["rel_res"]
    assert rd["iters"] % stride == 0
    assert_same(P.get(ca.CH_X), o.get(po.CH_X), "pressure", mask=typ == po.LEAF)


def test_solve_without_convergence_test():
    """sync_stride <= 0: exactly max_iters iterations, relative residual -1."""
    o, P, typ, ro, rd = _solve_both("plane", ca.SMOOTH_RBGS, 1, sync_stride=0, max_iters=7)
    assert rd["iters"] == ro["iters"] == 7 and rd["rel_res"] == ro["rel_res"] == -1.0
    assert rd["rr"] == ro["rr"]
    assert_same(P.get(ca.CH_X), o.get(po.CH_X), "pressure", mask=typ == po.LEAF)


def test_zero_rhs_returns_immediately():
    o, P, typ, nbr, b = pair("plane")
    P.set(ca.CH_B, np.zeros_like(b))
    S = ca.GMGSolver()
    assert S.solve(P, False, 100, 1e-6, 1, 10, 1, False) == (0, 0.0)
    assert not P.get(ca.CH_X)[typ == po.LEAF].any()


def test_solve_repeatable():
    _, P, typ, _, rd1 = _solve_both("plane", ca.SMOOTH_RBGS, 2)
    x1 = P.get(ca.CH_X)
    _, P, typ, _, rd2 = _solve_both("plane", ca.SMOOTH_RBGS, 2)
    assert rd1["iters"] == rd2["iters"] and rd1["rr"] == rd2["rr"]
    assert np.array_equal(x1, P.get(ca.CH_X))


def test_pure_neumann_solve_bitexact():
    kw = dict(uniform_level=2, face_bc=[po.NEUMANN] * 6)
    o = po.Oracle(**kw)
    lev, crd, typ, nbr = o.topology()
    P = ca.Problem(ca.Tree(o.params.h0, lev, crd, typ, o.celltypes()))
    b = o.get(po.CH_B)
    o2, P2, typ2, ro, rd = _solve_both(None, ca.SMOOTH_RBGS, 1, pure=True, scene=(o, P, typ, nbr, b))
    assert rd["iters"] == ro["iters"] and rd["rr"] == ro["rr"]
    assert_same(P.get(ca.CH_X), o.get(po.CH_X), "pressure", mask=typ == po.LEAF)


def test_recenter_bitexact():
    o, P, typ, nbr, b = pair("band")
    both_set(o, P, po.CH_R, rnd((P.n, 512), 11))
    o.recenter(po.CH_R)
    ca.ReCenterLeafVoxels(P, ca.CH_R)
    assert_same(P.get(ca.CH_R), o.get(po.CH_R), "re-centred r")


def test_damped_jacobi_solve_bitexact():
    o, P, typ, nbr, b = pair("uniform32")
    both_set(o, P, po.CH_B, b)
    ro = o.damped_jacobi_solve(25, 1e-6, 1.0)
    it, rel = ca.GMGSolver().dampedJacobiSolve(P, False, 25, 1e-6, 1.0)
    assert it == ro["iters"]
    assert_same(P.get(ca.CH_X), o.get(po.CH_X), "Jacobi-solve x", mask=typ == po.LEAF)


# Pressure vs the FP64 oracle, per scene: relative L2 error over LEAF INTERIOR
# cells, and the true residual |b - A x| / |b| recomputed in FP64.
FP64_TOL = {"uniform32": 1e-4, "plain64": 1e-4, "plane": 1e-4, "band": 1e-4, "vortex2root": 1e-4}


@pytest.mark.parametrize("name", list(SCENES))
def test_solve_vs_fp64_oracle(name):
    o, P, typ, ro, rd = _solve_both(name, ca.SMOOTH_RBGS, 2)
    x32 = P.get(ca.CH_X).astype(np.float64)
    b = rd["b"].astype(np.float64)  # channel 1 now holds the residual
    o64 = po.Oracle(params=o.params, fp64=True)
    o64.set(po.CH_B, b)
    r64 = o64.solve(max_iters=500, tol=1e-13, level_iters=2, coarsest_iters=10)
    assert r64["rel_res"] < 1e-12
    x64 = o64.get(po.CH_X)
    m = (typ == po.LEAF)[:, None] & (o.celltypes() == po.INTERIOR)
    err = np.linalg.norm((x32 - x64)[m]) / np.linalg.norm(x64[m])
    assert err < FP64_TOL[name], f"rel L2 error vs FP64 {err:.3e}"
    # true residual of the FP32 pressure, in FP64
    xin = np.where(m, x32, 0.0)
    o64.set(po.CH_P, xin)
    o64.full_negative_laplacian(po.CH_P, po.CH_AP)
    ax = o64.get(po.CH_AP)
    res = np.linalg.norm((b - ax)[m]) / np.linalg.norm(b[m])
    assert res < 5e-6, f"true relative residual {res:.3e}"


@pytest.mark.parametrize("name,pure", [("band", False), ("vortex2root", False), ("plane", False)])
def test_verify_f64_equals_oracle_fp64_residual(name, pure):
    """verify_f64: the FP64 residual of the FP32 pressure, recomputed on the
    device, equals the FP64 oracle's operator and dot on the same x, bit for bit."""
    o, P, typ, nbr, b = pair(name)
    both_set(o, P, po.CH_B, b)
    S = ca.GMGSolver()
    S.solve(P, False, 200, 1e-6, 2, 10, 1, pure, verify_f64=True)
    x32 = P.get(ca.CH_X)
    o64 = po.Oracle(params=o.params, fp64=True)
    leaf = (typ == po.LEAF)[:, None]
    o64.set(po.CH_X, np.where(leaf, x32.astype(np.float64), 0.0))
    o64.full_negative_laplacian(po.CH_X, po.CH_AP)
    r = np.where(leaf, b.astype(np.float32).astype(np.float64) - o64.get(po.CH_AP), 0.0)
    o64.set(po.CH_TMP, r)
    rr = o64.dot(po.CH_TMP, po.CH_TMP, po.LEAF)
    assert S.last["rel_res_f64"] == np.sqrt(rr / S.last["bb"])
    assert S.last["rel_res_f64"] < 2e-6
