"""GPU tests of the serving loop: resident instances, continuous batching and
warm starts.

Every request's arithmetic touches only its instance's tiles and scalars and
keeps its reductions' order, so a request gives the same bits whether it runs
alone or while other requests start and finish around it; a cold request also
equals GMGSolver::solve of the same problem.
"""
import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po

pytestmark = pytest.mark.gpu

SCENES = [
    dict(kind=po.SCENE_PLANE, plane_level=[3, 2]),
    dict(kind=po.SCENE_SPHERE_BAND, solid_r=0.15, band_level=[4, 3, 2]),
    dict(kind=po.SCENE_VORTEX, nroot=[1, 1, 2], vortex_level=[4, 2], seed=3),
    dict(uniform_level=2),
]
KW = dict(level_iters=2, coarsest_iters=10)
BUDGET = 2 << 30
_cache = {}


def scenes():
    if "s" not in _cache:
        _cache["s"] = [ca.Scene(**kw) for kw in SCENES]
    return _cache["s"]


def rhs2(sc, seed, eps=0.02):
    """A correlated next right-hand side for an instance: the scene's divergence
    with a small seeded perturbation on the same cells (zero where b is zero)."""
    rng = np.random.default_rng(seed)
    b = sc.rhs.astype(np.float32)
    return (b + eps * b * rng.standard_normal(b.shape).astype(np.float32)).astype(np.float32)


def run_until_done(S, expect, max_steps=5000):
    out = {}
    for _ in range(max_steps):
        S.step()
        while True:
            r = S.next_result()
            if r is None:
                break
            out[r[0]] = r
        if len(out) >= expect:
            return out
    raise AssertionError("requests did not finish")


def alone(sc):
    P = ca.Problem(sc.tree)
    P.set(ca.CH_B, sc.rhs)
    G = ca.GMGSolver()
    G.solve(P, False, 1000, 1e-6, KW["level_iters"], KW["coarsest_iters"], 1, False)
    return G.last, P.get(ca.CH_X)


def test_cold_requests_equal_solve():
    scs = scenes()
    S = ca.Server(BUDGET, 16, **KW)
    ids = [S.admit(sc.tree) for sc in scs]
    for q, (i, sc) in enumerate(zip(ids, scs)):
        S.submit(i, sc.rhs, warm=False, tag=q)
    res = run_until_done(S, len(scs))
    for q, sc in enumerate(scs):
        tag, inst, r, x = res[q]
        ref, xr = alone(sc)
        m = sc.tree.type == po.LEAF
        assert r["converged"] == 1 and r["iters"] == ref["iters"] and r["rr"] == ref["rr"], q
        assert np.array_equal(x[m], xr[m]), q


def test_staggered_arrivals_equal_solve():
    """Requests that join while others run, and leave before them."""
    scs = scenes()
    S = ca.Server(BUDGET, 16, **KW)
    ids = [S.admit(sc.tree) for sc in scs]
    S.submit(ids[1], scs[1].rhs, tag=1)
    for _ in range(3):
        S.step()
    S.submit(ids[0], scs[0].rhs, tag=0)
    S.step()
    S.submit(ids[2], scs[2].rhs, tag=2)
    res = run_until_done(S, 3)
    for q in range(3):
        ref, xr = alone(scs[q])
        tag, inst, r, x = res[q]
        m = scs[q].tree.type == po.LEAF
        assert r["iters"] == ref["iters"] and r["rr"] == ref["rr"], q
        assert np.array_equal(x[m], xr[m]), q


def warm_sequence(S, inst, sc, others=()):
    """Cold request, then two warm-started ones; other instances' requests
    (instance, scene) are kept running alongside."""
    out = []
    tag = 100
    for k, b in enumerate([sc.rhs, rhs2(sc, 1), rhs2(sc, 2)]):
        S.submit(inst, b, warm=k > 0, tag=tag + k)
        busy = list(others)
        for j, (oi, osc) in enumerate(busy):
            try:
                S.submit(oi, rhs2(osc, 10 + k), warm=True, tag=1000 + 10 * k + j)
            except ca.CirrusError:
                pass  # still running its previous request
        got = None
        for _ in range(5000):
            S.step()
            while True:
                r = S.next_result()
                if r is None:
                    break
                if r[0] == tag + k:
                    got = r
            if got is not None:
                break
        assert got is not None
        out.append(got)
    return out


def test_warm_starts_same_alone_and_in_the_mix():
    scs = scenes()
    S1 = ca.Server(BUDGET, 16, **KW)
    a = S1.admit(scs[1].tree)
    seq_alone = warm_sequence(S1, a, scs[1])
    S2 = ca.Server(BUDGET, 16, **KW)
    ids = [S2.admit(sc.tree) for sc in scs]
    seq_mix = warm_sequence(S2, ids[1], scs[1], others=[(ids[0], scs[0]), (ids[2], scs[2]), (ids[3], scs[3])])
    m = scs[1].tree.type == po.LEAF
    for (t1, _, r1, x1), (t2, _, r2, x2) in zip(seq_alone, seq_mix):
        assert r1["iters"] == r2["iters"] and r1["rr"] == r2["rr"]
        assert np.array_equal(x1[m], x2[m])
    # a warm start from the previous pressure needs fewer iterations than a
    # cold start of the same right-hand side
    P = ca.Problem(scs[1].tree)
    P.set(ca.CH_B, rhs2(scs[1], 1))
    G = ca.GMGSolver()
    G.solve(P, False, 1000, 1e-6, KW["level_iters"], KW["coarsest_iters"], 1, False)
    assert seq_alone[1][2]["iters"] < G.last["iters"]


def test_warm_start_reaches_the_tolerance():
    """The warm-started pressure solves A x = b to the tolerance relative to |b|
    (true residual with the operator, recomputed on the device in FP32 and
    summed in FP64)."""
    scs = scenes()
    sc = scs[0]
    S = ca.Server(BUDGET, 4, **KW)
    i = S.admit(sc.tree)
    seq = warm_sequence(S, i, sc)
    b = rhs2(sc, 2)
    x = seq[-1][3]
    P = ca.Problem(sc.tree)
    P.set(ca.CH_X, x)
    ca.FullNegativeLaplacian(P, ca.CH_X, ca.CH_AP)
    ax = P.get(ca.CH_AP).astype(np.float64)
    m = (sc.tree.type == po.LEAF)[:, None] & (sc.tree.cell_type == po.INTERIOR)
    r = (b.astype(np.float64) - ax)[m]
    assert np.linalg.norm(r) / np.linalg.norm(b.astype(np.float64)[m]) < 5e-6
    assert seq[-1][2]["converged"] == 1


def test_admission_by_budget_and_release():
    scs = scenes()
    big = scs[1].tree
    per_tile = 8 * 2048 + 2 * 512 + 256 + 16 + 5 * 176
    S = ca.Server(int(2.5 * big.n * per_tile), 16, **KW)
    a = S.admit(big)
    b = S.admit(big)
    with pytest.raises(ca.CirrusBusy):
        S.admit(big)
    S.release(a)
    c = S.admit(big)
    S.submit(c, scs[1].rhs, tag=7)
    res = run_until_done(S, 1)
    ref, xr = alone(scs[1])
    assert res[7][2]["iters"] == ref["iters"]
    assert S.info()["instances"] == 2
    del b


def test_perturbation_updates_the_kept_rhs():
    """The instance keeps its right-hand side across requests: rho = 1, sigma = 0
    leaves b as it was, so the warm request solves the same b (same |b|^2) from
    its converged pressure and stops within a few iterations."""
    sc = scenes()[1]
    S = ca.Server(BUDGET, 4, **KW)
    i = S.admit(sc.tree)
    S.submit(i, sc.rhs, tag=0)
    r0 = run_until_done(S, 1)[0][2]
    S.perturb_rhs(i, 1.0, 0.0, 5)
    S.submit(i, None, warm=True, tag=1)
    r1 = run_until_done(S, 1)[1][2]
    assert r1["bb"] == r0["bb"] and r1["converged"] == 1 and r1["iters"] <= 3


def test_verified_requests_equal_verified_solve():
    """verify_f64: every reported request carries its FP64 residual
    |b - A x| / |b|, recomputed on the device from the kept b; with BF16 V-cycle
    storage a cold request still equals the verified solve alone.  (The FP32
    pressure's FP64 residual can sit just above the tolerance its recursive
    residual met; refine mode is what guarantees it.)"""
    scs = scenes()
    S = ca.Server(BUDGET, 16, storage=ca.STORE_BF16, verify_f64=1, **KW)
    ids = [S.admit(sc.tree) for sc in scs]
    for q, (i, sc) in enumerate(zip(ids, scs)):
        S.submit(i, sc.rhs, tag=q)
    res = run_until_done(S, len(scs))
    for q, sc in enumerate(scs):
        P = ca.Problem(sc.tree)
        P.set(ca.CH_B, sc.rhs)
        G = ca.GMGSolver()
        G.solve(P, False, 1000, 1e-6, KW["level_iters"], KW["coarsest_iters"], 1, False, storage=ca.STORE_BF16,
                verify_f64=True)
        r = res[q][2]
        assert 0 <= r["rel_res_f64"] < 5e-6 and r["converged"] == 1, r
        assert (r["iters"], r["rr"], r["rel_res_f64"]) == (G.last["iters"], G.last["rr"], G.last["rel_res_f64"])
    # warm requests in the mix are verified too
    for q, (i, sc) in enumerate(zip(ids, scs)):
        S.perturb_rhs(i, 0.95, 0.3 * float(np.sqrt(np.mean(sc.rhs[sc.rhs != 0] ** 2))), 100 + q)
        S.submit(i, None, warm=True, tag=10 + q)
    res = run_until_done(S, len(scs))
    for q in range(len(scs)):
        r = res[10 + q][2]
        assert 0 <= r["rel_res_f64"] < 5e-6 and r["converged"] == 1, r


def test_refined_requests_equal_pool_refinement():
    """refine: a cold request is the pool's FP64 iterative refinement with no
    inner relative tolerance (each step to half the outer tolerance), bit for bit;
    its reported residual is the FP64 one."""
    scs = scenes()
    S = ca.Server(BUDGET, 16, refine=True, storage=ca.STORE_BF16, **KW)
    ids = [S.admit(sc.tree) for sc in scs]
    for q, (i, sc) in enumerate(zip(ids, scs)):
        S.submit(i, sc.rhs, tag=q)
    res = run_until_done(S, len(scs))
    for q, sc in enumerate(scs):
        P = ca.Problem(sc.tree)
        P.set(ca.CH_B, sc.rhs)
        ref = ca.GMGSolver().solve_refined(P, 1e-6, 0.0, 5, 1000, KW["level_iters"], KW["coarsest_iters"],
                                           storage=ca.STORE_BF16)[0]
        tag, inst, r, x = res[q]
        assert r["converged"] == 1 and 0 <= r["rel_res_f64"] < 1e-6, r
        assert (r["iters"], r["rr"], r["bb"]) == (ref["iters"], ref["rr"], ref["bb"]), q
        assert r["rel_res_f64"] == ref["rel_res"]
        m = sc.tree.type == po.LEAF
        assert np.array_equal(x[m], P.get(ca.CH_X)[m]), q


def test_refined_warm_sequences_same_alone_and_in_the_mix():
    """refine: warm starts from the FP64 pressure kept on the device; each
    result is verified in FP64 and equal alone and in the mix."""
    scs = scenes()
    kw = dict(refine=True, storage=ca.STORE_BF16, **KW)
    S1 = ca.Server(BUDGET, 16, **kw)
    a = S1.admit(scs[1].tree)
    seq_alone = warm_sequence(S1, a, scs[1])
    S2 = ca.Server(BUDGET, 16, **kw)
    ids = [S2.admit(sc.tree) for sc in scs]
    seq_mix = warm_sequence(S2, ids[1], scs[1], others=[(ids[0], scs[0]), (ids[2], scs[2]), (ids[3], scs[3])])
    m = scs[1].tree.type == po.LEAF
    for (t1, _, r1, x1), (t2, _, r2, x2) in zip(seq_alone, seq_mix):
        assert r1["converged"] == 1 and r1["rel_res_f64"] < 1e-6
        assert (r1["iters"], r1["rr"]) == (r2["iters"], r2["rr"])
        assert np.array_equal(x1[m], x2[m])
    # the warm requests need fewer iterations than the cold one
    assert seq_alone[1][2]["iters"] < seq_alone[0][2]["iters"]


def test_admission_churn_results_unchanged():
    """Instances admitted, served and released in a stream (released memory and
    slots come back once the stream has passed the release, and new instances
    reuse them while other requests run; requests start and finish in groups):
    every cold refined request still equals the pool's refinement of the same
    problem alone, bit for bit."""
    scs = scenes()
    ref = {}
    for q, sc in enumerate(scs):
        P = ca.Problem(sc.tree)
        P.set(ca.CH_B, sc.rhs)
        r = ca.GMGSolver().solve_refined(P, 1e-6, 0.0, 5, 1000, KW["level_iters"], KW["coarsest_iters"],
                                         storage=ca.STORE_BF16)[0]
        ref[q] = ((r["iters"], r["rr"], r["bb"]), P.get(ca.CH_X))
    S = ca.Server(BUDGET, 3, refine=True, storage=ca.STORE_BF16, **KW)
    rng = np.random.default_rng(11)
    live = {}  # instance -> scene
    running = {}  # tag -> (instance, scene)
    done = 0
    tag = 0
    for _ in range(20000):
        while len(live) < 3 and tag < 40:
            q = int(rng.integers(len(scs)))
            i = S.admit(scs[q].tree)
            live[i] = q
            S.submit(i, scs[q].rhs, tag=tag)
            running[tag] = (i, q)
            tag += 1
        for _ in range(int(rng.integers(1, 4))):
            S.step()
        while True:
            r = S.next_result()
            if r is None:
                break
            t, inst, res, x = r
            i, q = running.pop(t)
            assert inst == i
            want, xw = ref[q]
            assert (res["iters"], res["rr"], res["bb"]) == want, (t, q, res)
            m = scs[q].tree.type == po.LEAF
            assert np.array_equal(x[m], xw[m]), (t, q)
            S.release(i)
            del live[i]
            done += 1
        if tag >= 40 and not running:
            break
    assert done == 40
    assert S.info()["instances"] == 0
