"""cg_dot_noalloc: the dot product through the entry point that allocates
nothing (its tile list, partials and result in a fixed device array), called
through the C ABI and compared bit for bit with cg_dot (Dot,
src/PoissonGrid.cu:217-228) on the same channels and tile types.  Two calls in
a row, the second on a smaller problem, must each equal cg_dot; a problem with
more tiles than the array holds is refused before any kernel runs.
"""
import ctypes

import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po

pytestmark = pytest.mark.gpu

TYPE_SETS = [po.LEAF, po.NONLEAF, po.GHOST, po.LEAF | po.GHOST, po.LEAF | po.NONLEAF | po.GHOST]


def problem(**kw):
    o = po.Oracle(**kw)
    lev, crd, typ, _ = o.topology()
    P = ca.Problem(ca.Tree(o.params.h0, lev, crd, typ, o.celltypes()))
    rng = np.random.default_rng(P.n)
    for ch in (ca.CH_X, ca.CH_B, ca.CH_P):
        P.set(ch, rng.standard_normal((P.n, 512)).astype(np.float32))
    return P


def c_dot(P, fn, a, b, types):
    out = ctypes.c_double()
    st = getattr(ca.lib(), fn)(P.handle, a, b, types, ctypes.byref(out))
    assert st == 0, ca.lib().cg_last_error()
    return out.value


def check(P):
    for types in TYPE_SETS:
        for a, b in [(ca.CH_X, ca.CH_B), (ca.CH_P, ca.CH_P), (ca.CH_B, ca.CH_X)]:
            ref = c_dot(P, "cg_dot", a, b, types)
            got = c_dot(P, "cg_dot_noalloc", a, b, types)
            assert np.float64(got).tobytes() == np.float64(ref).tobytes(), (P.n, types, a, b, got, ref)


def test_noalloc_dot_equals_dot_then_a_smaller_problem():
    big = problem(kind=po.SCENE_SPHERE_BAND, band_level=[4, 3, 2])
    small = problem(kind=po.SCENE_PLANE, plane_level=[3, 2])
    assert small.n < big.n
    check(big)
    check(small)  # the array keeps the larger call's contents; only what this call writes is read
    check(big)


def test_noalloc_dot_of_no_tiles_is_zero():
    P = problem(kind=po.SCENE_UNIFORM, uniform_level=2)
    assert c_dot(P, "cg_dot_noalloc", ca.CH_X, ca.CH_X, po.GHOST) == 0.0
    assert c_dot(P, "cg_dot", ca.CH_X, ca.CH_X, po.GHOST) == 0.0


def test_noalloc_dot_refuses_problems_beyond_its_array():
    o = po.Oracle(kind=po.SCENE_SPHERE_BAND, band_level=[4, 3, 2])
    lev, crd, typ, _ = o.topology()
    tree = ca.Tree(o.params.h0, lev, crd, typ, o.celltypes())
    copies = ca.NOALLOC_MAX_TILES // tree.n + 1
    B = ca.Problem.batch([tree] * copies)
    assert B.n > ca.NOALLOC_MAX_TILES
    out = ctypes.c_double(7.0)
    st = ca.lib().cg_dot_noalloc(B.handle, ca.CH_X, ca.CH_X, po.LEAF, ctypes.byref(out))
    assert st != 0 and b"scratch" in ca.lib().cg_last_error()
    assert out.value == 7.0
    B.close()
