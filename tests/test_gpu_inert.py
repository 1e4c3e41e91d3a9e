"""Tiles a solve leaves untouched: a LEAF tile without ghost children and
without an INTERIOR cell, whose right-hand side is +0 in every cell, keeps the
values the iterations would write there (zeros; x unchanged), so the CG
iterations skip it once the solve's init has marked it.  The pressure, the
iteration count and |r|^2 must stay the oracle's bit for bit -- with such
tiles inert, with a right-hand side that is nonzero on their cells (they are
not inert then), and again inert on the same handle afterwards (the values a
non-inert solve left in their parents' V-cycle rhs are cleared by the init).
"""
import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po

pytestmark = pytest.mark.gpu

SCENES = {  # each holds LEAF tiles inside the solid sphere
    "uniform64_solid": dict(uniform_level=3, solid_r=0.35),
    "band": dict(kind=po.SCENE_SPHERE_BAND, solid_r=0.15, band_level=[4, 3, 2]),
    "band_big_solid": dict(kind=po.SCENE_SPHERE_BAND, solid_r=0.3, band_r=0.3, band_level=[4, 3, 2]),
}


def candidates(o):
    """LEAF tiles with no INTERIOR cell and no GHOST child (the records' inert candidates)."""
    lev, crd, typ, _ = o.topology()
    ct = o.celltypes()
    key = {(int(l), *map(int, c)) for l, c, t in zip(lev, np.asarray(crd).reshape(-1, 3), typ) if t == po.GHOST}
    out = []
    for t, (l, c) in enumerate(zip(lev, np.asarray(crd).reshape(-1, 3))):
        if typ[t] != po.LEAF or (ct[t] == po.INTERIOR).any():
            continue
        kids = [(int(l) + 1, 2 * int(c[0]) + (q & 1), 2 * int(c[1]) + ((q >> 1) & 1), 2 * int(c[2]) + ((q >> 2) & 1))
                for q in range(8)]
        if not any(k in key for k in kids):
            out.append(t)
    return np.array(out, dtype=int)


@pytest.mark.parametrize("storage", [ca.STORE_F32, ca.STORE_BF16])
@pytest.mark.parametrize("name", list(SCENES))
def test_inert_tiles_keep_the_solve_bitexact(name, storage):
    o = po.Oracle(**SCENES[name])
    lev, crd, typ, _ = o.topology()
    cand = candidates(o)
    assert len(cand) > 0, "the scene has no tile that can be inert"
    b0 = o.get(po.CH_B).astype(np.float32)
    assert not b0[cand].any()
    P = ca.Problem(ca.Tree(o.params.h0, lev, crd, typ, o.celltypes()))
    S = ca.GMGSolver()
    leaf = typ == po.LEAF
    # nonzero right-hand sides on every other candidate tile (cells that are not
    # INTERIOR): tiny (the solve still converges) and of unit size (the CG
    # residual cannot fall below it; it may end in NaN, in the oracle alike)
    noise = np.random.default_rng(3).standard_normal((len(cand[::2]), 512)).astype(np.float32)
    b_tiny, b_big = b0.copy(), b0.copy()
    b_tiny[cand[::2]] = noise * np.float32(1e-7 * np.sqrt(np.mean(b0[b0 != 0] ** 2)))
    b_big[cand[::2]] = noise

    def same(a, b):
        return a == b or (np.isnan(a) and np.isnan(b))

    for step, b in enumerate((b0, b_tiny, b0, b_big, b0)):
        o.set(po.CH_B, b.astype(np.float64))
        P.set(ca.CH_B, b)
        ro = o.solve(max_iters=300, tol=1e-6, level_iters=2, coarsest_iters=10, storage=storage)
        S.solve(P, False, 300, 1e-6, 2, 10, 1, False, storage=storage)
        got, want = (S.last["iters"], S.last["rr"], S.last["bb"]), (ro["iters"], ro["rr"], ro["bb"])
        assert all(same(a, b) for a, b in zip(got, want)), (step, got, want)
        if step < 3:
            assert S.last["converged"] == 1, step
        x, xo = P.get(ca.CH_X)[leaf], o.get(po.CH_X).astype(np.float32)[leaf]
        assert np.array_equal(x, xo, equal_nan=True), (step, int((x != xo).sum()))
