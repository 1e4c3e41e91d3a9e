"""Pins the oracle (the CPU transcription) to what the mathematics fixes.

The reference ships no tests or golden vectors, so each part of the oracle is
checked against an independent statement: a leaf-partition finite-volume
operator (fv_reference), dense-grid closed forms, invariants (symmetry,
consistency on linear fields, R = P^T, 2:1 balance, ghost rules), brute-force
counts and a sparse direct solve.
"""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import pyoracle as po
from fv_reference import LeafCells

SCENES = {
    "uniform32": dict(uniform_level=2),
    "plane": dict(kind=po.SCENE_PLANE, plane_level=[3, 2], has_solid=0),
    "band": dict(kind=po.SCENE_SPHERE_BAND, solid_r=0.15, band_level=[4, 3, 2]),
    "vortex": dict(kind=po.SCENE_VORTEX, nroot=[1, 1, 2], vortex_level=[3, 2], seed=5),
}
_cache = {}


def scene(name, fp64=True):
    key = (name, fp64)
    if key not in _cache:
        o = po.Oracle(fp64=fp64, **SCENES[name])
        lev, crd, typ, nbr = o.topology()
        ct = o.celltypes()
        lc = LeafCells(lev, crd, typ, ct, o.params.h0, list(o.params.nroot))
        _cache[key] = (o, lev, crd, typ, nbr, ct, lc)
    return _cache[key]


def leaf_interior_vector(lc, rng):
    v = rng.standard_normal(lc.n)
    v[lc.ct != po.INTERIOR] = 0
    return v


@pytest.mark.parametrize("name", list(SCENES))
def test_operator_matches_independent_fv(name):
    o, lev, crd, typ, nbr, ct, lc = scene(name)
    A = lc.matrix()
    rng = np.random.default_rng(1)
    x = leaf_interior_vector(lc, rng)
    o.set(po.CH_X, lc.to_tiles(x, o.n))
    o.full_negative_laplacian(po.CH_X, po.CH_AP)
    ax = lc.from_tiles(o.get(po.CH_AP))
    ref = A @ x
    m = lc.ct == po.INTERIOR
    err = np.abs(ax[m] - ref[m]).max() / np.abs(ref[m]).max()
    assert err < 1e-12, err
    # diagonal
    o.full_negative_laplacian(po.CH_X, po.CH_AP, True)
    d = lc.from_tiles(o.get(po.CH_AP))
    assert np.abs(d[m] - A.diagonal()[m]).max() / A.diagonal()[m].max() < 1e-12


@pytest.mark.parametrize("name", ["band", "plane"])
def test_operator_symmetric_positive(name):
    o, lev, crd, typ, nbr, ct, lc = scene(name)
    rng = np.random.default_rng(2)
    x, y = leaf_interior_vector(lc, rng), leaf_interior_vector(lc, rng)

    def A(v):
        o.set(po.CH_X, lc.to_tiles(v, o.n))
        o.full_negative_laplacian(po.CH_X, po.CH_AP)
        return lc.from_tiles(o.get(po.CH_AP))

    ax, ay = A(x), A(y)
    assert abs(ax @ y - x @ ay) <= 1e-12 * abs(ax @ y)
    assert ax @ x > 0 and ay @ y > 0


def test_operator_exact_on_linear_field_across_coarse_fine_faces():
    """A u = 0 for u = z at INTERIOR cells whose neighbours are all INTERIOR,
    including cells on both sides of the z = 1/2 coarse/fine plane: only the
    2/3 coarse/fine coefficient with the ghost = parent rule balances the
    fluxes there."""
    o, lev, crd, typ, nbr, ct, lc = scene("plane")
    zc = (lc.g[:, 2] + 0.5) * lc.h
    o.set(po.CH_X, lc.to_tiles(zc, o.n))
    o.full_negative_laplacian(po.CH_X, po.CH_AP)
    ax = lc.from_tiles(o.get(po.CH_AP))
    # cells whose every face neighbour (any level) is INTERIOR
    A = lc.matrix().tocsr()
    ok = lc.ct == po.INTERIOR
    nb_ok = np.ones(lc.n, bool)
    for r in np.nonzero(ok)[0]:
        cols = A.indices[A.indptr[r]:A.indptr[r + 1]]
        nb_ok[r] = np.all(lc.ct[cols] == po.INTERIOR) and len(cols) >= 7
    m = ok & nb_ok
    interface = m & (np.abs(zc - 0.5) < 0.02)
    assert interface.sum() > 100
    assert np.abs(ax[m]).max() < 1e-12


def test_uniform_operator_closed_form():
    """On a uniform grid A x = h * sum_faces (x_c - x_n) with x = 0 outside and
    faces touching NEUMANN cells removed: a dense-array statement."""
    o = po.Oracle(fp64=True, uniform_level=2, has_solid=1, solid_r=0.3)
    lev, crd, typ, nbr = o.topology()
    ct = o.celltypes()
    N, h = 32, 0.125 / 4
    dense_ct = np.zeros((N, N, N), np.uint8)
    tiles = np.nonzero(typ == po.LEAF)[0]
    ci = np.arange(512)
    for t in tiles:
        i, j, k = crd[t] * 8
        dense_ct[i:i + 8, j:j + 8, k:k + 8] = ct[t].reshape(8, 8, 8)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((N, N, N)) * (dense_ct == po.INTERIOR)
    xt = np.zeros((o.n, 512))
    for t in tiles:
        i, j, k = crd[t] * 8
        xt[t] = X[i:i + 8, j:j + 8, k:k + 8].ravel()
    o.set(po.CH_X, xt)
    o.full_negative_laplacian(po.CH_X, po.CH_AP)
    got = o.get(po.CH_AP)
    Xp = np.pad(X, 1)
    Cp = np.pad(dense_ct, 1, constant_values=po.DIRICHLET)
    acc = np.zeros_like(X)
    neu0 = (dense_ct & po.NEUMANN) != 0
    for a in range(3):
        for s in (-1, 1):
            sl = [slice(1, N + 1)] * 3
            sl[a] = slice(1 + s, N + 1 + s)
            xn, cn = Xp[tuple(sl)], Cp[tuple(sl)]
            cut = neu0 | ((cn & po.NEUMANN) != 0)
            acc += np.where(cut, 0.0, X - xn)
    ref = np.where(dense_ct == po.INTERIOR, h * acc, 0.0)
    for t in tiles:
        i, j, k = crd[t] * 8
        assert np.allclose(got[t], ref[i:i + 8, j:j + 8, k:k + 8].ravel(), rtol=1e-12, atol=1e-14)


def test_gauss_seidel_textbook_update():
    """One red sweep on the finest level = x_c + (b_c - (A x)_c) / A_cc on red
    INTERIOR cells, black cells unchanged."""
    o, lev, crd, typ, nbr, ct, lc = scene("uniform32")
    A = lc.matrix()
    rng = np.random.default_rng(4)
    x, b = leaf_interior_vector(lc, rng), rng.standard_normal(lc.n)
    o.set(po.CH_Z, lc.to_tiles(x, o.n))
    o.set(po.CH_TMP, lc.to_tiles(b, o.n))
    L = o.max_level
    o.gs_color(L, po.CH_Z, po.CH_TMP, 0)
    got = lc.from_tiles(o.get(po.CH_Z))
    red = (lc.g.sum(axis=1) % 2 == 0) & (lc.ct == po.INTERIOR)
    ref = x.copy()
    ref[red] = x[red] + (b[red] - (A @ x)[red]) / A.diagonal()[red]
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name", ["band", "vortex"])
def test_restriction_is_transpose_of_prolongation(name):
    o, lev, crd, typ, nbr, ct, lc = scene(name)
    rng = np.random.default_rng(5)
    for level in range(1, o.max_level + 1):
        f = rng.standard_normal((o.n, 512)) * (ct == po.INTERIOR) * ((lev == level) & (typ != po.GHOST))[:, None]
        c = rng.standard_normal((o.n, 512)) * (lev == level - 1)[:, None]
        o.set(po.CH_TMP, f)
        o.set(po.CH_Z, np.zeros((o.n, 512)))
        o.set(po.CH_AP, np.zeros((o.n, 512)))
        o.residual_restrict(level, po.CH_Z, po.CH_TMP, po.CH_AP)
        Rf = o.get(po.CH_AP)
        o.set(po.CH_AP, c)
        o.prolongate(level, po.CH_AP, po.CH_Z)
        Pc = o.get(po.CH_Z)
        lhs = np.sum(Rf[lev == level - 1] * c[lev == level - 1])
        rhs = np.sum(f[lev == level] * Pc[lev == level])
        assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs))


def test_restriction_maps_fine_cell_to_its_parent_cell():
    o, lev, crd, typ, nbr, ct, lc = scene("band")
    level = o.max_level
    t = np.nonzero((lev == level) & (typ == po.LEAF))[0][7]
    f = np.zeros((o.n, 512))
    f[t, 5 * 64 + 2 * 8 + 3] = 1.0
    o.set(po.CH_TMP, f)
    o.set(po.CH_Z, np.zeros((o.n, 512)))
    o.set(po.CH_AP, np.zeros((o.n, 512)))
    o.residual_restrict(level, po.CH_Z, po.CH_TMP, po.CH_AP)
    got = o.get(po.CH_AP)
    g = crd[t] * 8 + np.array([5, 2, 3])
    pg = g // 2
    pt = np.nonzero((lev == level - 1) & np.all(crd == pg // 8, axis=1))[0][0]
    pc = pg % 8
    assert got[pt, pc[0] * 64 + pc[1] * 8 + pc[2]] == 1.0
    assert np.count_nonzero(got) == 1


@pytest.mark.parametrize("smoother,nu", [(po.SMOOTH_RBGS, 1), (po.SMOOTH_RBGS, 2), (po.SMOOTH_JACOBI, 2)])
def test_vcycle_symmetric_positive(smoother, nu):
    """The V-cycle is a symmetric positive preconditioner (mirrored smoothing,
    R = P^T, symmetric bottom solve), as CG requires."""
    o, lev, crd, typ, nbr, ct, lc = scene("band")
    rng = np.random.default_rng(6)
    r1, r2 = leaf_interior_vector(lc, rng), leaf_interior_vector(lc, rng)

    def V(r):
        o.set(po.CH_R, lc.to_tiles(r, o.n))
        o.vcycle(po.CH_Z, po.CH_R, po.CH_AP, nu, 10, smoother, 2.0 / 3.0)
        return lc.from_tiles(o.get(po.CH_Z))

    v1, v2 = V(r1), V(r2)
    assert abs(v1 @ r2 - r1 @ v2) <= 1e-10 * abs(v1 @ r2)
    assert v1 @ r1 > 0 and v2 @ r2 > 0


@pytest.mark.parametrize("name", ["band", "vortex", "plane"])
def test_fp64_mgpcg_solves_independent_system(name):
    """The FP64 MGPCG pressure satisfies the independently assembled
    finite-volume system: ||b - A_fv x|| / ||b|| at the solver's tolerance."""
    o, lev, crd, typ, nbr, ct, lc = scene(name)
    o32 = po.Oracle(fp64=False, **SCENES[name])
    b = o32.get(po.CH_B)
    o.set(po.CH_B, b)
    r = o.solve(max_iters=300, tol=1e-12, level_iters=2, coarsest_iters=10)
    assert r["rel_res"] < 1e-12
    x = lc.from_tiles(o.get(po.CH_X))
    m = lc.ct == po.INTERIOR
    A = lc.matrix()[m][:, m]
    bl = lc.from_tiles(b)[m]
    assert np.linalg.norm(bl - A @ x[m]) / np.linalg.norm(bl) < 1e-10


def test_fp64_mgpcg_matches_direct_solve():
    """On the 32^3 uniform scene the FP64 MGPCG pressure equals a sparse direct
    solve of the independently assembled system."""
    o, lev, crd, typ, nbr, ct, lc = scene("uniform32")
    o32 = po.Oracle(fp64=False, **SCENES["uniform32"])
    b = o32.get(po.CH_B)
    o.set(po.CH_B, b)
    r = o.solve(max_iters=300, tol=1e-13, level_iters=2, coarsest_iters=10)
    assert r["rel_res"] < 1e-13
    x = lc.from_tiles(o.get(po.CH_X))
    m = lc.ct == po.INTERIOR
    A = lc.matrix()[m][:, m]
    xd = spla.spsolve(A.tocsc(), lc.from_tiles(b)[m])
    assert np.linalg.norm(x[m] - xd) / np.linalg.norm(xd) < 1e-9


def test_dot_close_to_exact_sum():
    o, lev, crd, typ, nbr, ct, lc = scene("band", fp64=False)
    rng = np.random.default_rng(7)
    a = rng.standard_normal((o.n, 512)).astype(np.float32)
    b = rng.standard_normal((o.n, 512)).astype(np.float32)
    o.set(po.CH_P, a.astype(np.float64))
    o.set(po.CH_Z, b.astype(np.float64))
    d = o.dot(po.CH_P, po.CH_Z, po.LEAF)
    leaf = typ == po.LEAF
    exact = math.fsum((a[leaf].astype(np.float64) * b[leaf].astype(np.float64)).ravel())
    assert abs(d - exact) <= 1e-13 * np.sum(np.abs(a[leaf].astype(np.float64) * b[leaf]))


def test_uniform_tree_counts_and_no_ghosts():
    o = po.Oracle()
    lev, crd, typ, nbr = o.topology()
    assert list(np.bincount(lev)) == [1, 8, 64, 512]
    assert (typ[lev == 3] == po.LEAF).all() and (typ[lev < 3] == po.NONLEAF).all()


@pytest.mark.parametrize("name", list(SCENES))
def test_topology_invariants(name):
    o, lev, crd, typ, nbr, ct, lc = scene(name)
    keys = {(int(l), *map(int, c)): i for i, (l, c) in enumerate(zip(lev, crd))}
    # neighbour table symmetric and consistent with coordinates
    for t in range(o.n):
        for f in range(6):
            n = nbr[t, f]
            q = crd[t].copy()
            q[f % 3] += -1 if f < 3 else 1
            assert n == keys.get((int(lev[t]), *map(int, q)), -1)
            if n >= 0:
                assert nbr[n, (f + 3) % 6] == t
    # every ghost has a LEAF parent; every non-root active tile a NONLEAF parent
    for t in range(o.n):
        if lev[t] == 0:
            continue
        p = keys[(int(lev[t]) - 1, *map(int, crd[t] // 2))]
        assert typ[p] == (po.LEAF if typ[t] == po.GHOST else po.NONLEAF)
    # ghosts are exactly the empty face positions of active tiles whose parent is a LEAF
    want = set()
    for t in range(o.n):
        if lev[t] == 0 or typ[t] == po.GHOST:
            continue
        for f in range(6):
            q = crd[t].copy()
            q[f % 3] += -1 if f < 3 else 1
            k = (int(lev[t]), *map(int, q))
            pk = (int(lev[t]) - 1, *map(int, q // 2))
            if (k not in keys or typ[keys[k]] == po.GHOST) and pk in keys and typ[keys[pk]] == po.LEAF:
                want.add(k)
    have = {(int(lev[t]), *map(int, crd[t])) for t in range(o.n) if typ[t] == po.GHOST}
    assert want == have
    # 2:1 face balance of the leaves
    leaf = {(int(lev[t]), *map(int, crd[t])) for t in range(o.n) if typ[t] == po.LEAF}
    for (l, x, y, z) in leaf:
        for f in range(6):
            q = np.array([x, y, z])
            q[f % 3] += -1 if f < 3 else 1
            for up in range(2, l + 1):
                assert (l - up, *map(int, q >> up)) not in leaf


def test_solid_cells_match_brute_force_count():
    o = po.Oracle()
    lev, crd, typ, nbr = o.topology()
    ct = o.celltypes()
    N, h = 64, 1.0 / 64
    idx = np.arange(N)
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    inside = ((X + 0.5) * h - 0.5) ** 2 + ((Y + 0.5) * h - 0.5) ** 2 + ((Z + 0.5) * h - 0.5) ** 2 <= 0.2 ** 2
    boundary = (X == 0) | (X == N - 1) | (Y == 0) | (Y == N - 1) | (Z == 0) | (Z == N - 1)
    n_neu = sum(int(np.sum(ct[t] == po.NEUMANN)) for t in np.nonzero(typ == po.LEAF)[0])
    n_dir = sum(int(np.sum(ct[t] == po.DIRICHLET)) for t in np.nonzero(typ == po.LEAF)[0])
    top_only = (Z == N - 1) & ~((X == 0) | (X == N - 1) | (Y == 0) | (Y == N - 1))
    assert n_dir == int(top_only.sum())
    assert n_neu == int((inside & ~boundary).sum() + (boundary & ~top_only).sum())


def test_rhs_sum_equals_boundary_flux():
    """Divergence theorem on the plain scene: the sum of b over INTERIOR cells
    equals h^2 times the net face velocity through faces between INTERIOR and
    non-INTERIOR cells (the velocities come from the oracle's channels)."""
    o = po.Oracle(fp64=True)
    lev, crd, typ, nbr = o.topology()
    ct, b = o.celltypes(), o.get(po.CH_B)
    N, h = 64, 1.0 / 64
    dct = np.zeros((N, N, N), np.uint8)
    U = np.zeros((3, N, N, N))
    B = np.zeros((N, N, N))
    for t in np.nonzero(typ == po.LEAF)[0]:
        i, j, k = crd[t] * 8
        dct[i:i + 8, j:j + 8, k:k + 8] = ct[t].reshape(8, 8, 8)
        B[i:i + 8, j:j + 8, k:k + 8] = b[t].reshape(8, 8, 8)
        for a in range(3):
            U[a, i:i + 8, j:j + 8, k:k + 8] = o.get(po.CH_U + a)[t].reshape(8, 8, 8)
    interior = dct == po.INTERIOR
    flux = 0.0
    for a in range(3):
        ua = U[a]
        # face between cell c-1 and c along axis a carries ua[c] (minus face of c)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(0, N - 1)
        hi[a] = slice(1, N)
        cm, cp, uf = interior[tuple(lo)], interior[tuple(hi)], ua[tuple(hi)]
        flux += np.sum(uf[cm & ~cp]) - np.sum(uf[~cm & cp])
    assert abs(B[interior].sum() - h * h * flux) < 1e-9 * max(1.0, h * h * np.abs(U).sum())


def test_jacobi_textbook_update():
    """One damped-Jacobi sweep on the finest level = x + w (b - A x) / A_cc on
    INTERIOR cells, with A assembled independently."""
    o, lev, crd, typ, nbr, ct, lc = scene("uniform32")
    A = lc.matrix()
    rng = np.random.default_rng(8)
    x, b = leaf_interior_vector(lc, rng), rng.standard_normal(lc.n)
    o.set(po.CH_Z, lc.to_tiles(x, o.n))
    o.set(po.CH_TMP, lc.to_tiles(b, o.n))
    w = 2.0 / 3.0
    o.jacobi_sweep(o.max_level, po.CH_Z, po.CH_TMP, w)
    got = lc.from_tiles(o.get(po.CH_Z))
    m = lc.ct == po.INTERIOR
    ref = x.copy()
    ref[m] = x[m] + w * (b[m] - (A @ x)[m]) / A.diagonal()[m]
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-14)


def _reduce_types(fine):
    """2x2x2 block reduction of a dense cell-type array: DIRICHLET if any child
    is, else INTERIOR if any child is, else NEUMANN."""
    n = fine.shape[0] // 2
    blk = fine.reshape(n, 2, n, 2, n, 2).transpose(0, 2, 4, 1, 3, 5).reshape(n, n, n, 8)
    anyd = (blk == po.DIRICHLET).any(axis=-1)
    anyi = (blk == po.INTERIOR).any(axis=-1)
    return np.where(anyd, po.DIRICHLET, np.where(anyi, po.INTERIOR, po.NEUMANN)).astype(np.uint8)


def test_nonleaf_cell_types_are_dense_block_reductions():
    """On the uniform 64^3 scene every NONLEAF level equals the 2x2x2 block
    reduction of the level below, assembled as dense arrays."""
    o = po.Oracle()
    lev, crd, typ, nbr = o.topology()
    ct = o.celltypes()

    def dense(level):
        n = 8 << level
        d = np.zeros((n, n, n), np.uint8)
        for t in np.nonzero(lev == level)[0]:
            i, j, k = crd[t] * 8
            d[i:i + 8, j:j + 8, k:k + 8] = ct[t].reshape(8, 8, 8)
        return d

    fine = dense(3)
    for level in (2, 1, 0):
        want = _reduce_types(fine)
        assert np.array_equal(dense(level), want), level
        fine = want


def test_ghost_cell_types_copy_the_parent_cell():
    o, lev, crd, typ, nbr, ct, lc = scene("band")
    keys = {(int(l), *map(int, c)): i for i, (l, c) in enumerate(zip(lev, crd))}
    ci = np.arange(512)
    loc = np.stack([ci >> 6, (ci >> 3) & 7, ci & 7], axis=1)
    for t in np.nonzero(typ == po.GHOST)[0]:
        g = crd[t] * 8 + loc
        pg = g // 2
        p = keys[(int(lev[t]) - 1, *map(int, crd[t] // 2))]
        pc = pg - crd[p] * 8
        assert np.array_equal(ct[t], ct[p][pc[:, 0] * 64 + pc[:, 1] * 8 + pc[:, 2]])


def test_recenter_removes_the_interior_mean():
    """ReCenterLeafVoxels: the sum over all LEAF cells divided by the INTERIOR
    count is subtracted from INTERIOR cells, so their sum becomes 0."""
    o, lev, crd, typ, nbr, ct, lc = scene("band")
    rng = np.random.default_rng(9)
    v = leaf_interior_vector(lc, rng) + 0.25
    v[lc.ct != po.INTERIOR] = 0
    o.set(po.CH_TMP, lc.to_tiles(v, o.n))
    o.recenter(po.CH_TMP)
    got = lc.from_tiles(o.get(po.CH_TMP))
    m = lc.ct == po.INTERIOR
    assert abs(got[m].sum()) < 1e-9 * m.sum()
    d = v[m] - got[m]
    assert np.allclose(d, v[m].sum() / m.sum(), rtol=1e-12)
    assert np.array_equal(got[~m], v[~m])


def test_damped_jacobi_solve_first_iterate_and_progress():
    """dampedJacobiSolve from x = 0: the first iterate is w b / A_cc on INTERIOR
    cells; more iterations lower the residual."""
    o, lev, crd, typ, nbr, ct, lc = scene("uniform32")
    A = lc.matrix()
    b = lc.to_tiles(leaf_interior_vector(lc, np.random.default_rng(10)), o.n)
    w = 2.0 / 3.0
    o.set(po.CH_B, b)
    o.damped_jacobi_solve(1, 1e-30, w)
    x = lc.from_tiles(o.get(po.CH_X))
    m = lc.ct == po.INTERIOR
    bl = lc.from_tiles(b)
    assert np.allclose(x[m], w * bl[m] / A.diagonal()[m], rtol=1e-12)
    res = []
    for it in (10, 40):
        o.set(po.CH_B, b)
        res.append(o.damped_jacobi_solve(it, 1e-30, w)["rel_res"])
    assert res[1] < res[0] < 1.0


ALL_NEUMANN = dict(kind=po.SCENE_SPHERE_BAND, solid_r=0.15, band_level=[4, 3, 2], face_bc=[po.NEUMANN] * 6)


def test_operator_annihilates_constants_without_dirichlet_cells():
    """With every wall NEUMANN there is no DIRICHLET cell, and A 1 = 0 on LEAF
    INTERIOR cells across coarse/fine faces too."""
    o = po.Oracle(fp64=True, **ALL_NEUMANN)
    lev, crd, typ, nbr = o.topology()
    ct = o.celltypes()
    assert not (ct[typ == po.LEAF] == po.DIRICHLET).any()
    one = (ct == po.INTERIOR).astype(np.float64) * (typ == po.LEAF)[:, None]
    o.set(po.CH_X, one)
    o.full_negative_laplacian(po.CH_X, po.CH_AP)
    got = o.get(po.CH_AP)[typ == po.LEAF]
    assert np.abs(got).max() < 1e-13


def test_pure_neumann_fp64_solve_satisfies_independent_system():
    """Pure-Neumann MGPCG with re-centring (ReCenterLeafVoxels each iteration)."""
    o = po.Oracle(fp64=True, **ALL_NEUMANN)
    lev, crd, typ, nbr = o.topology()
    ct = o.celltypes()
    lc = LeafCells(lev, crd, typ, ct, o.params.h0, list(o.params.nroot))
    b = po.Oracle(fp64=False, **ALL_NEUMANN).get(po.CH_B)
    o.set(po.CH_B, b)
    r = o.solve(max_iters=300, tol=1e-11, level_iters=2, coarsest_iters=10, pure_neumann=True)
    assert r["rel_res"] < 1e-11
    x = lc.from_tiles(o.get(po.CH_X))
    m = lc.ct == po.INTERIOR
    bl = lc.from_tiles(b)[m]
    # A is singular (A 1 = 0) and b is not exactly mean-free (coarse/fine faces
    # see ghost velocities), so the re-centred solve satisfies A x = b - mean(b)
    A = lc.matrix()[m][:, m]
    bc = bl - bl.mean()
    assert np.linalg.norm(bc - A @ x[m]) / np.linalg.norm(bl) < 1e-8


# ---------------------------------------------------------------------------
# Mixed-precision form
# ---------------------------------------------------------------------------
def _bf16_nearest(v):
    """Independent bfloat16 rounding: of the two bfloat16 neighbours of each FP32
    value (pattern truncated, and one bfloat16 step further from zero), the
    nearer in exact float64 arithmetic, ties to the even significand."""
    v = np.ascontiguousarray(v, np.float32)
    u = v.view(np.uint32)
    lo_bits = u & np.uint32(0xFFFF0000)
    hi_bits = lo_bits + np.uint32(0x10000)
    lo = lo_bits.view(np.float32).astype(np.float64)
    hi = hi_bits.view(np.float32).astype(np.float64)
    x = v.astype(np.float64)
    dlo, dhi = np.abs(x - lo), np.abs(hi - x)
    take_hi = (dhi < dlo) | ((dhi == dlo) & (((lo_bits >> 16) & 1) == 1))
    return np.where(take_hi, hi_bits, lo_bits).view(np.float32)


def test_bf16_round_every_value_and_tie():
    pats = np.arange(0, 1 << 16, dtype=np.uint32) << 16
    vals = pats.view(np.float32)
    fin = np.isfinite(vals)
    # bfloat16 values round to themselves; their midpoints (ties) go to the even neighbour
    exact = vals[fin]
    assert np.array_equal(po.bf16_round(exact).view(np.uint32), exact.view(np.uint32))
    mids = (pats[fin] | np.uint32(0x8000)).view(np.float32)
    # (the top binade's upper half rounds to infinity, outside the two-neighbour rule)
    mids = mids[(mids.view(np.uint32) & 0x7FFFFFFF) < 0x7F7F8000]
    assert np.array_equal(po.bf16_round(mids).view(np.uint32), _bf16_nearest(mids).view(np.uint32))
    rng = np.random.default_rng(0)
    r = rng.integers(0, 1 << 32, 1 << 20, dtype=np.uint64).astype(np.uint32).view(np.float32)
    r = r[(r.view(np.uint32) & 0x7FFFFFFF) < 0x7F7F8000]
    assert np.array_equal(po.bf16_round(r).view(np.uint32), _bf16_nearest(r).view(np.uint32))


def _is_bf16(a):
    return np.all((np.ascontiguousarray(a, np.float32).view(np.uint32) & 0xFFFF) == 0)


@pytest.mark.parametrize("smoother,nu", [(po.SMOOTH_RBGS, 2), (po.SMOOTH_JACOBI, 2)])
def test_bf16_vcycle_stores_only_bf16_values_and_stays_close(smoother, nu):
    """With BF16 storage every value the V-cycle leaves in its x and rhs
    channels is a bfloat16 value (a store that skipped the rounding would leave
    an FP32 value), and z stays within bfloat16 accuracy of the FP32 V-cycle."""
    o = po.Oracle(fp64=False, **SCENES["band"])
    b = o.get(po.CH_B)
    o.vcycle(po.CH_Z, po.CH_R, po.CH_AP, nu, 10, smoother, 2.0 / 3.0, storage=po.STORE_BF16)
    z16, rhs16 = o.get(po.CH_Z), o.get(po.CH_AP)
    assert _is_bf16(z16) and _is_bf16(rhs16)
    o.set(po.CH_B, b)
    o.vcycle(po.CH_Z, po.CH_R, po.CH_AP, nu, 10, smoother, 2.0 / 3.0)
    z32 = o.get(po.CH_Z)
    assert not _is_bf16(z32)
    rel = np.linalg.norm(z16 - z32) / np.linalg.norm(z32)
    assert 0 < rel < 2e-2, rel


def test_fp16_round_every_value_and_tie():
    """float16 stores: every float16 value is kept, the midpoints between
    neighbouring float16 values and random FP32 patterns round as numpy's
    float32 -> float16 conversion (nearest, ties to even, subnormals kept), and
    magnitudes past the largest finite value clamp to +-65504."""
    h = np.arange(0, 1 << 16, dtype=np.uint32).astype(np.uint16).view(np.float16)
    f = h[np.isfinite(h)].astype(np.float32)
    assert np.array_equal(po.fp16_round(f).view(np.uint32), f.view(np.uint32))
    pos = np.unique(np.abs(f.astype(np.float64)))
    mids = ((pos[:-1] + pos[1:]) / 2).astype(np.float32)  # exact in FP32
    mids = np.concatenate([mids, -mids])
    assert np.array_equal(po.fp16_round(mids), mids.astype(np.float16).astype(np.float32))
    rng = np.random.default_rng(0)
    r = rng.integers(0, 1 << 32, 1 << 20, dtype=np.uint64).astype(np.uint32).view(np.float32)
    r = r[np.isfinite(r) & (np.abs(r) <= 65504)]
    assert np.array_equal(po.fp16_round(r), r.astype(np.float16).astype(np.float32))
    big = np.array([65504.0, 65519.0, 65520.0, 1e6, 3e38, -7e4], np.float32)
    assert np.array_equal(po.fp16_round(big), np.array([65504, 65504, 65504, 65504, 65504, -65504], np.float32))


def test_fp16_scale_exponent():
    """The float16 solve's right-hand-side scale 2^k puts the mean square of
    the scaled b in [2^-8, 2^-6) (root mean square in [2^-4, 2^-3))."""
    rng = np.random.default_rng(1)
    for _ in range(2000):
        cells = float(512 * rng.integers(1, 20000))
        bb = cells * 2.0 ** rng.uniform(-90, 60)
        k = po.fp16_scale_exp(bb, cells)
        ms = (bb / cells) * 4.0 ** k
        assert 2.0 ** -8 <= ms < 2.0 ** -6, (bb, cells, k)
    assert po.fp16_scale_exp(0.0, 512.0) == 0


def _is_fp16(a):
    a = np.ascontiguousarray(a, np.float32)
    return np.array_equal(a.astype(np.float16).astype(np.float32), a)


def test_fp16_vcycle_stores_only_fp16_values_and_stays_close():
    """With float16 storage every value the V-cycle leaves in its x and rhs
    channels is a float16 value, and on a right-hand side scaled as the solve
    scales it, z stays within float16 accuracy of the FP32 V-cycle."""
    o = po.Oracle(fp64=False, **SCENES["band"])
    lev, crd, typ, nbr = o.topology()
    b = o.get(po.CH_B)
    k = po.fp16_scale_exp(o.dot(po.CH_B, po.CH_B), 512.0 * np.sum(typ == po.LEAF))
    bs = (b.astype(np.float32) * np.float32(2.0 ** k)).astype(np.float64)
    o.set(po.CH_B, bs)
    o.vcycle(po.CH_Z, po.CH_R, po.CH_AP, 2, 10, po.SMOOTH_RBGS, 2.0 / 3.0, storage=po.STORE_F16)
    z16, rhs16 = o.get(po.CH_Z), o.get(po.CH_AP)
    assert _is_fp16(z16) and _is_fp16(rhs16)
    o.set(po.CH_B, bs)
    o.vcycle(po.CH_Z, po.CH_R, po.CH_AP, 2, 10, po.SMOOTH_RBGS, 2.0 / 3.0)
    z32 = o.get(po.CH_Z)
    assert not _is_fp16(z32)
    rel = np.linalg.norm(z16 - z32) / np.linalg.norm(z32)
    assert 0 < rel < 5e-3, rel


@pytest.mark.parametrize("j", [-7, 5])
def test_fp16_solve_is_exactly_scale_equivariant(j):
    """The float16 solve of 2^j b is 2^j times the solve of b, bit for bit (the
    scale exponent absorbs j), with the same iterations; it converges."""
    o = po.Oracle(fp64=False, **SCENES["band"])
    b = o.get(po.CH_B)
    r1 = o.solve(max_iters=200, tol=1e-6, level_iters=2, coarsest_iters=10, storage=po.STORE_F16)
    x1 = o.get(po.CH_X)
    o.set(po.CH_B, b * 2.0 ** j)
    r2 = o.solve(max_iters=200, tol=1e-6, level_iters=2, coarsest_iters=10, storage=po.STORE_F16)
    x2 = o.get(po.CH_X)
    assert r1["iters"] == r2["iters"] and r1["rel_res"] == r2["rel_res"] and r1["rel_res"] < 1e-6
    assert r2["bb"] == r1["bb"] * 4.0 ** j
    assert np.array_equal(x2, x1 * 2.0 ** j)


@pytest.mark.parametrize("storage", [po.STORE_BF16, po.STORE_F16])
def test_refined_solve_meets_the_independent_fp64_system_16bit(storage):
    """Refinement with either 16-bit V-cycle storage reaches the FP64 system's
    tolerance (the bfloat16 case is checked in detail below)."""
    o64, lev, crd, typ, nbr, ct, lc = scene("band")
    o = po.Oracle(fp64=False, **SCENES["band"])
    b = o.get(po.CH_B)
    r = o.solve_refined(tol=1e-9, inner_tol=1e-4, level_iters=2, coarsest_iters=10, storage=storage)
    assert r["converged"] == 1 and r["rel_res"] < 1e-9
    x = lc.from_tiles(o.refined_x())
    m = lc.ct == po.INTERIOR
    A = lc.matrix()[m][:, m]
    bl = lc.from_tiles(b)[m]
    res = np.linalg.norm(bl - A @ x[m]) / np.linalg.norm(bl)
    assert res < 1e-9 and abs(res - r["rel_res"]) < 1e-3 * r["rel_res"], (res, r["rel_res"])


def test_refined_solve_meets_the_independent_fp64_system():
    """Iterative refinement (FP64 x and residual, BF16-storage MGPCG inside):
    the FP64 x satisfies the independently assembled system to the outer
    tolerance, the reported residual is that system's, and on 32^3 x equals a
    sparse direct solve."""
    for name in ["band", "uniform32"]:
        o64, lev, crd, typ, nbr, ct, lc = scene(name)
        o = po.Oracle(fp64=False, **SCENES[name])
        b = o.get(po.CH_B)
        r = o.solve_refined(tol=1e-9, inner_tol=1e-4, level_iters=2, coarsest_iters=10)
        assert r["converged"] == 1 and r["outer"] >= 2 and r["rel_res"] < 1e-9
        x = lc.from_tiles(o.refined_x())
        m = lc.ct == po.INTERIOR
        A = lc.matrix()[m][:, m]
        bl = lc.from_tiles(b)[m]
        res = np.linalg.norm(bl - A @ x[m]) / np.linalg.norm(bl)
        assert res < 1e-9 and abs(res - r["rel_res"]) < 1e-3 * r["rel_res"], (res, r["rel_res"])
        if name == "uniform32":
            xd = spla.spsolve(A.tocsc(), bl)
            assert np.linalg.norm(x[m] - xd) / np.linalg.norm(xd) < 1e-6
        # the FP32 channel holds the FP64 x rounded
        assert np.array_equal(lc.from_tiles(o.get(po.CH_X)), x.astype(np.float32).astype(np.float64))


def test_refined_pure_neumann_solve():
    o = po.Oracle(fp64=False, **ALL_NEUMANN)
    lev, crd, typ, nbr = o.topology()
    ct = o.celltypes()
    lc = LeafCells(lev, crd, typ, ct, o.params.h0, list(o.params.nroot))
    b = o.get(po.CH_B)
    r = o.solve_refined(tol=1e-9, inner_tol=1e-4, level_iters=2, coarsest_iters=10, pure_neumann=True)
    assert r["converged"] == 1
    x = lc.from_tiles(o.refined_x())
    m = lc.ct == po.INTERIOR
    bl = lc.from_tiles(b)[m]
    A = lc.matrix()[m][:, m]
    assert np.linalg.norm(bl - bl.mean() - A @ x[m]) / np.linalg.norm(bl) < 1e-8


# ---------------------------------------------------------------------------
# The projection around the solve
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["band", "vortex", "plane"])
def test_divergence_of_gradient_is_minus_the_operator(name):
    """From a zero velocity field, AddGradientToFaceCenters then
    VelocityVolumeDivergenceOnLeafs (no ghost re-propagation) gives -A p on the
    LEAF INTERIOR cells, A the operator pinned against the independent
    finite-volume assembly: the divergence, the gradient with its 1.5 h
    coarse/fine spacing and the ghost accumulations are consistent with it.
    Recomputing the divergence of the scene's velocities reproduces its rhs."""
    o, lev, crd, typ, nbr, ct, lc = scene(name)
    o32 = po.Oracle(fp64=True, **SCENES[name])
    b = o32.get(po.CH_B)
    o32.velocity_divergence(po.CH_TMP)
    assert np.array_equal(o32.get(po.CH_TMP)[typ == po.LEAF], b[typ == po.LEAF])
    for a in range(3):
        o32.set(po.CH_U + a, np.zeros((o32.n, 512)))
    rng = np.random.default_rng(7)
    pv = rng.standard_normal((o32.n, 512))
    pv[(typ != po.LEAF)] = 0
    pv[ct != po.INTERIOR] = 0
    o32.set(po.CH_X, pv)
    o32.set(po.CH_P, pv)
    o32.full_negative_laplacian(po.CH_P, po.CH_AP)
    ap = o32.get(po.CH_AP)
    o32.add_gradient(po.CH_X)
    o32.velocity_divergence(po.CH_TMP, propagate=False)
    d = o32.get(po.CH_TMP)
    m = (typ == po.LEAF)[:, None] & (ct == po.INTERIOR)
    assert np.abs(d[m] + ap[m]).max() < 1e-13 * np.abs(ap[m]).max()
    # the same against the independent matrix
    x = lc.from_tiles(pv)
    ref = (lc.matrix() @ x)[lc.ct == po.INTERIOR]
    assert np.abs(lc.from_tiles(d)[lc.ct == po.INTERIOR] + ref).max() < 1e-12 * np.abs(ref).max()


# ---------------------------------------------------------------------------
# AMGSolver: the stored-coefficient preconditioner
# ---------------------------------------------------------------------------
def _amg_scene(name, fp64=True):
    o = po.Oracle(fp64=fp64, **SCENES[name])
    b = o.get(po.CH_B)
    o.amg_prepare()
    return o, b


def test_amg_leaf_operator_equals_the_pinned_operator_on_a_uniform_level():
    """On a uniform LEAF level (no ghosts) the stored coefficients give the same
    rows as the GMG operator, which the independent finite-volume assembly pins
    (x = 0 off INTERIOR cells, as in every solve: a DIRICHLET neighbour enters
    the diagonal only)."""
    o, _ = _amg_scene("uniform32")
    lev, crd, typ, nbr = o.topology()
    L = lev.max()
    x = np.random.default_rng(3).standard_normal((len(lev), 512)) * (o.celltypes() == po.INTERIOR)
    o.set(po.CH_P, x)
    o.amg_apply(L, po.CH_P, po.CH_AP)
    a_amg = o.get(po.CH_AP)[lev == L]
    o.set(po.CH_P, x)
    o.full_negative_laplacian(po.CH_P, po.CH_AP)
    a_gmg = o.get(po.CH_AP)[lev == L]
    assert np.abs(a_amg - a_gmg).max() <= 1e-12 * np.abs(a_gmg).max()


@pytest.mark.parametrize("name", ["uniform32", "band"])
def test_amg_coarse_coefficients_are_the_galerkin_product(name):
    """A NONLEAF cell's stored row is R A P of its children's rows (R sums the
    8 children, P injects): diagonal = children's diagonals + twice their
    internal couplings, minus-face entry = the sum over the 4 children on that
    face.  Checked by assembling the child level's stored operator as a matrix
    and multiplying, cell by cell, in FP64."""
    o, _ = _amg_scene(name)
    lev, crd, typ, nbr = o.topology()
    C = o.amg_coeffs()
    ct = o.celltypes()
    n = len(lev)
    # global cell index -> (tile, cell) per level
    where = {}
    for t in range(n):
        for c in range(512):
            i, j, k = c >> 6, (c >> 3) & 7, c & 7
            where[(int(lev[t]), int(crd[t, 0]) * 8 + i, int(crd[t, 1]) * 8 + j, int(crd[t, 2]) * 8 + k)] = (t, c)

    def row(l, g):
        """Stored row of cell g at level l as {neighbour cell key: value}."""
        t, c = where[(l,) + g]
        r = {(l,) + g: C[3, t, c]}
        for a in range(3):
            m = list(g); m[a] -= 1
            if C[a, t, c] != 0:
                r[(l,) + tuple(m)] = r.get((l,) + tuple(m), 0.0) + C[a, t, c]
            p = list(g); p[a] += 1
            key = (l,) + tuple(p)
            if key in where:
                tp, cp = where[key]
                if C[a, tp, cp] != 0:
                    r[key] = r.get(key, 0.0) + C[a, tp, cp]
        return r

    checked = 0
    rng = np.random.default_rng(4)
    for t in rng.permutation(np.where(typ == po.NONLEAF)[0])[:6]:
        l = int(lev[t])
        for c in rng.choice(512, 24, replace=False):
            i, j, k = c >> 6, (c >> 3) & 7, c & 7
            G = (int(crd[t, 0]) * 8 + i, int(crd[t, 1]) * 8 + j, int(crd[t, 2]) * 8 + k)
            kids = [(l + 1, 2 * G[0] + a, 2 * G[1] + b, 2 * G[2] + d) for d in (0, 1) for b in (0, 1) for a in (0, 1)]
            # (R A P)[G, H] = sum over children f of G and children g of H of A[f, g]
            rap = {}
            for f in kids:
                for key, v in row(f[0], f[1:]).items():
                    parent = (l, key[1] >> 1, key[2] >> 1, key[3] >> 1)
                    rap[parent] = rap.get(parent, 0.0) + v
            if ct[t, c] != po.INTERIOR:
                assert C[:, t, c].tolist() == [0.0, 0.0, 0.0, C[3, t, c]] or ct[t, c] == po.DIRICHLET
                continue
            scale = max(abs(v) for v in rap.values()) or 1.0
            assert abs(rap.get((l,) + G, 0.0) - C[3, t, c]) <= 1e-12 * scale
            for a in range(3):
                m = list(G); m[a] -= 1
                assert abs(rap.get((l,) + tuple(m), 0.0) - C[a, t, c]) <= 1e-12 * scale, (a, rap, C[:, t, c])
            checked += 1
    assert checked > 0


@pytest.mark.parametrize("name", ["band", "vortex", "plane"])
def test_amg_solve_reaches_the_same_pressure(name):
    """The AMG-preconditioned CG solves the same system as the GMG one (the
    CG operator is FullNegativeLaplacian in both, src/AMGSolver.cu:752), so in
    FP64 both reach the same pressure; the AMG V-cycle only changes the path."""
    o, b = _amg_scene(name)
    ra = o.amg_solve(max_iters=500, tol=1e-12, level_iters=2, coarsest_iters=10)
    xa = o.get(po.CH_X)
    g = po.Oracle(fp64=True, **SCENES[name])
    rg = g.solve(max_iters=500, tol=1e-12, level_iters=2, coarsest_iters=10)
    xg = g.get(po.CH_X)
    lev, crd, typ, nbr = g.topology()
    m = (typ == po.LEAF)[:, None] & (g.celltypes() == po.INTERIOR)
    assert ra["rel_res"] < 1e-12 and rg["rel_res"] < 1e-12
    assert np.linalg.norm((xa - xg)[m]) / np.linalg.norm(xg[m]) < 1e-9
    assert ra["iters"] < 200
