"""Pins the oracle's CoarsenStep and its adaptation loop (AdaptScene: the
refinement steps, then CoarsenStep until none deletes, then SpawnGhostTiles;
src/PoissonGrid.h:165-326, src/FluidEuler.h:68-111,820-835).

Checks, each chosen so that a dropped or loosened rule fails one of them:
  * closed form: coarsening a uniform level-L tree to a uniform target T
    leaves exactly the uniform level-T tree;
  * two algorithms agree: for a target whose value on a tile is the maximum of
    its children's (uniform; plane with the fine side below; sphere bands,
    whose shells are connected), the only tree where refinement and
    coarsening both stop is the least 2:1-balanced tree meeting the target,
    which the scene build reaches by refinement alone.  So the adaptation of
    any tree to such a target must give the scene's own tree, ghosts and cell
    types included;
  * invariants on a target without that property (plane with the fine side
    above): 2:1 face balance, the ghost rule, refinement adds and coarsening
    removes, and the coarsening's fixed point (no group of eight LEAF
    siblings that may all go is left).
"""
import itertools

import numpy as np
import pytest

import pyoracle as po

U = lambda L: dict(kind=po.SCENE_UNIFORM, uniform_level=L)
BAND = dict(kind=po.SCENE_SPHERE_BAND, band_level=[5, 4, 3])
BAND_OFF = dict(kind=po.SCENE_SPHERE_BAND, band_c=[0.3, 0.6, 0.45], band_r=0.2, band_w=[0.03, 0.1],
                band_level=[5, 3, 1])
PLANE_DOWN = dict(kind=po.SCENE_PLANE, plane_level=[4, 1], plane_z=0.3)
PLANE_UP = dict(kind=po.SCENE_PLANE, plane_level=[1, 4], plane_z=0.6)


def tiles(o):
    lev, crd, typ, _ = o.topology()
    return {(int(l), *map(int, c)): int(t) for l, c, t in zip(lev, crd, typ)}


def uniform_tiles(L):
    out = {}
    for l in range(L + 1):
        for c in itertools.product(range(1 << l), repeat=3):
            out[(l, *c)] = po.LEAF if l == L else po.NONLEAF
    return out


@pytest.mark.parametrize("L,T", [(2, 0), (2, 1), (3, 1), (3, 2)])
def test_coarsen_uniform_closed_form(L, T):
    o = po.Oracle.adapted(U(L), U(T), po.ADAPT_COARSEN)
    assert tiles(o) == uniform_tiles(T)


def test_refine_uniform_closed_form():
    assert tiles(po.Oracle.adapted(U(1), U(3), po.ADAPT_REFINE)) == uniform_tiles(3)


MONOTONE = {"plane_down": PLANE_DOWN, "uniform2": U(2), "band": BAND}
STARTS = {"uniform4": U(4), "band": BAND, "band_off": BAND_OFF, "plane_up": PLANE_UP, "uniform0": U(0)}


@pytest.mark.parametrize("target", list(MONOTONE))
@pytest.mark.parametrize("start", list(STARTS))
def test_adapt_to_a_monotone_target_gives_the_scene_tree(start, target):
    a = po.Oracle.adapted(STARTS[start], MONOTONE[target])
    b = po.Oracle(**MONOTONE[target])
    assert tiles(a) == tiles(b)
    assert np.array_equal(a.celltypes(), b.celltypes())
    assert np.array_equal(a.get(po.CH_B), b.get(po.CH_B))


def test_coarsen_only_from_a_finer_tree_gives_the_scene_tree():
    # uniform level 4 contains the plane tree, so coarsening alone reaches it
    assert tiles(po.Oracle.adapted(U(4), PLANE_DOWN, po.ADAPT_COARSEN)) == tiles(po.Oracle(**PLANE_DOWN))


def plane_target(p, level, c):
    h = p["h0"] / (1 << level) if "h0" in p else 0.125 / (1 << level)
    return p["plane_level"][0] if c[2] * 8 * h < p["plane_z"] else p["plane_level"][1]


def check_balance_and_ghosts(t, nroot=(1, 1, 1)):
    active = {k for k, v in t.items() if v != po.GHOST}

    def in_domain(l, q):
        return all(0 <= q[d] and (q[d] >> l) < nroot[d] for d in range(3))

    for (l, *c), ty in t.items():
        if ty != po.LEAF or l < 1:
            continue
        for d, s in itertools.product(range(3), (-1, 1)):
            q = list(c)
            q[d] += s
            if in_domain(l, q):
                assert (l - 1, q[0] >> 1, q[1] >> 1, q[2] >> 1) in active, ("2:1 balance", l, c, d, s)
    ghosts = set()
    for (l, *c) in active:
        if l < 1:
            continue
        for d, s in itertools.product(range(3), (-1, 1)):
            q = list(c)
            q[d] += s
            if (l, *q) in active:
                continue
            if t.get((l - 1, q[0] >> 1, q[1] >> 1, q[2] >> 1)) == po.LEAF:
                ghosts.add((l, *q))
    assert ghosts == {k for k, v in t.items() if v == po.GHOST}


@pytest.mark.parametrize("start", ["uniform4", "band", "band_off"])
def test_adapt_to_a_non_monotone_target_invariants(start):
    s = STARTS[start]
    s_tiles = tiles(po.Oracle(**s))
    s_active = {k for k, v in s_tiles.items() if v != po.GHOST}
    ref = tiles(po.Oracle.adapted(s, PLANE_UP, po.ADAPT_REFINE))
    co = tiles(po.Oracle.adapted(s, PLANE_UP, po.ADAPT_COARSEN))
    both = tiles(po.Oracle.adapted(s, PLANE_UP))
    for t in (ref, co, both):
        check_balance_and_ghosts(t)
    ref_active = {k for k, v in ref.items() if v != po.GHOST}
    co_active = {k for k, v in co.items() if v != po.GHOST}
    assert s_active <= ref_active and co_active <= s_active
    # refinement stops only where every LEAF meets its target (or the last layer)
    for (l, *c), ty in ref.items():
        if ty == po.LEAF:
            assert l >= plane_target(PLANE_UP, l, c) or l == 9
    # coarsening stops only where no group of 8 LEAF siblings may all go:
    # each such group has a member next to a same-level NONLEAF tile
    for t in (co, both):
        for (l, *c), ty in t.items():
            if ty != po.NONLEAF:
                continue
            kids = [(l + 1, 2 * c[0] + (o & 1), 2 * c[1] + ((o >> 1) & 1), 2 * c[2] + ((o >> 2) & 1)) for o in range(8)]
            if not all(t.get(k) == po.LEAF and plane_target(PLANE_UP, k[0], k[1:]) < k[0] for k in kids):
                continue
            blocked = False
            for k in kids:
                for d, sg in itertools.product(range(3), (-1, 1)):
                    q = list(k[1:])
                    q[d] += sg
                    if t.get((k[0], *q)) == po.NONLEAF:
                        blocked = True
            assert blocked, ("coarsening stopped early", l, c)
    # a tree that has nothing to coarsen is left as it is
    assert tiles(po.Oracle.adapted(PLANE_UP, PLANE_UP, po.ADAPT_COARSEN)) == tiles(po.Oracle(**PLANE_UP))


def test_adapt_rejects_another_domain():
    with pytest.raises(RuntimeError):
        po.Oracle.adapted(U(2), dict(U(2), nroot=[1, 1, 2]))
