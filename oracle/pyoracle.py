"""TEST INFRASTRUCTURE ONLY -- ctypes loader for the oracle library.

The oracle (``cirrus_oracle.hpp``) is a CPU restatement of the reference MGPCG
path.  Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may use it, as the checker / baseline, never as the product.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

LEAF, NONLEAF, GHOST = 1, 2, 4
DIRICHLET, INTERIOR, NEUMANN = 1, 2, 4
CH_X, CH_B, CH_R, CH_P, CH_AP, CH_Z, CH_TMP, CH_U, CH_V, CH_W = 0, 1, 1, 2, 3, 4, 5, 6, 7, 8
SCENE_UNIFORM, SCENE_SPHERE_BAND, SCENE_PLANE, SCENE_VORTEX = 0, 1, 2, 3
SMOOTH_RBGS, SMOOTH_JACOBI = 0, 1
STORE_F32, STORE_BF16, STORE_F16 = 0, 1, 2
ADAPT_REFINE, ADAPT_COARSEN = 1, 2


def fp16_round(a):
    """FP32 values as a float16 V-cycle store leaves them (clamped to +-65504,
    rounded to nearest even; float32 array)."""
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty_like(a)
    lib().orc_fp16_round(_ptr(a), _ptr(out), a.size)
    return out


def fp16_scale_exp(bb, cells):
    """Exponent of the power-of-two right-hand-side scale of a float16-storage solve."""
    return int(lib().orc_fp16_scale_exp(float(bb), float(cells)))


def bf16_round(a):
    """FP32 values rounded to bfloat16 as the V-cycle stores do (float32 array)."""
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty_like(a)
    lib().orc_bf16_round(_ptr(a), _ptr(out), a.size)
    return out


class SceneParams(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("nroot", ctypes.c_int * 3),
        ("h0", ctypes.c_double),
        ("uniform_level", ctypes.c_int),
        ("has_solid", ctypes.c_int),
        ("solid_c", ctypes.c_double * 3),
        ("solid_r", ctypes.c_double),
        ("band_c", ctypes.c_double * 3),
        ("band_r", ctypes.c_double),
        ("band_w", ctypes.c_double * 2),
        ("band_level", ctypes.c_int * 3),
        ("plane_z", ctypes.c_double),
        ("plane_level", ctypes.c_int * 2),
        ("n_tubes", ctypes.c_int),
        ("tube_r_lo", ctypes.c_double),
        ("tube_r_hi", ctypes.c_double),
        ("vortex_thresh", ctypes.c_double),
        ("vortex_level", ctypes.c_int * 2),
        ("face_bc", ctypes.c_int * 6),
        ("seed", ctypes.c_ulonglong),
        ("num_layers", ctypes.c_int),
    ]


def scene_params(**kw):
    """SceneParams with the defaults of the plain configuration (uniform 64^3,
    solid sphere r=0.2 at the centre, top face Dirichlet, other walls Neumann)."""
    p = SceneParams()
    p.kind = SCENE_UNIFORM
    p.nroot[:] = [1, 1, 1]
    p.h0 = 0.125
    p.uniform_level = 3
    p.has_solid = 1
    p.solid_c[:] = [0.5, 0.5, 0.5]
    p.solid_r = 0.2
    p.band_c[:] = [0.5, 0.5, 0.5]
    p.band_r = 0.15
    p.band_w[:] = [0.05, 0.15]
    p.band_level[:] = [5, 4, 3]
    p.plane_z = 0.5
    p.plane_level[:] = [3, 2]
    p.n_tubes = 8
    p.tube_r_lo, p.tube_r_hi = 0.02, 0.06
    p.vortex_thresh = 0.5
    p.vortex_level[:] = [6, 4]
    p.face_bc[:] = [NEUMANN] * 5 + [DIRICHLET]
    p.seed = 0
    p.num_layers = 10
    for k, v in kw.items():
        f = getattr(p, k)
        if isinstance(f, ctypes.Array):
            f[:] = list(v)
        else:
            setattr(p, k, v)
    return p


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_mgpcg.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        P = ctypes.POINTER
        L.orc_create.restype = vp
        L.orc_create.argtypes = [P(SceneParams), i]
        L.orc_create_adapted.restype = vp
        L.orc_create_adapted.argtypes = [P(SceneParams), P(SceneParams), i, i]
        L.orc_destroy.argtypes = [vp]
        for n in ("orc_num_tiles", "orc_num_leaf_tiles", "orc_max_level"):
            getattr(L, n).restype = i
            getattr(L, n).argtypes = [vp]
        L.orc_export_topology.argtypes = [vp, vp, vp, vp, vp]
        L.orc_export_celltypes.argtypes = [vp, vp]
        L.orc_set_celltypes.argtypes = [vp, vp]
        L.orc_calc_celltypes_from_leafs.argtypes = [vp]
        L.orc_get_channel.argtypes = [vp, i, vp]
        L.orc_set_channel.argtypes = [vp, i, vp]
        L.orc_full_negative_laplacian.argtypes = [vp, i, i, i]
        L.orc_negative_laplacian_same_level.argtypes = [vp, i, i, i, i, i, i]
        L.orc_gs_color.argtypes = [vp, i, i, i, i]
        L.orc_jacobi_sweep.argtypes = [vp, i, i, i, d]
        L.orc_residual_restrict.argtypes = [vp, i, i, i, i]
        L.orc_prolongate.argtypes = [vp, i, i, i]
        L.orc_vcycle.argtypes = [vp, i, i, i, i, i, i, d, i]
        L.orc_dot.restype = d
        L.orc_dot.argtypes = [vp, i, i, i]
        L.orc_recenter.argtypes = [vp, i]
        L.orc_solve.restype = i
        L.orc_solve.argtypes = [vp, i, d, i, i, i, i, d, i, i, vp]
        L.orc_solve_refined.restype = i
        L.orc_solve_refined.argtypes = [vp, d, d, i, i, i, i, i, i, d, i, i, vp]
        L.orc_get_refined_x.restype = i
        L.orc_get_refined_x.argtypes = [vp, vp]
        L.orc_bf16_round.argtypes = [vp, vp, ctypes.c_long]
        L.orc_fp16_round.argtypes = [vp, vp, ctypes.c_long]
        L.orc_amg_prepare.argtypes = [vp]
        L.orc_amg_solve.restype = i
        L.orc_amg_solve.argtypes = [vp, i, d, i, i, i, i, vp]
        L.orc_amg_vcycle.argtypes = [vp, i, i]
        L.orc_amg_apply.argtypes = [vp, i, i, i]
        L.orc_amg_coeffs.argtypes = [vp, i, vp]
        L.orc_fp16_scale_exp.restype = i
        L.orc_fp16_scale_exp.argtypes = [d, d]
        L.orc_velocity_divergence.argtypes = [vp, i, i]
        L.orc_add_gradient.argtypes = [vp, i]
        L.orc_damped_jacobi_solve.restype = i
        L.orc_damped_jacobi_solve.argtypes = [vp, i, d, d, vp]
        L.orc_set_threads.argtypes = [i]
        L.orc_get_threads.restype = i
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One scene (grid + RHS) held by the oracle, FP32 or FP64."""

    def __init__(self, params=None, fp64=False, _handle=None, **kw):
        self.params = params if params is not None else scene_params(**kw)
        self.fp64 = bool(fp64)
        self.h = _handle if _handle is not None else lib().orc_create(ctypes.byref(self.params), int(self.fp64))
        if not self.h:
            raise RuntimeError("oracle scene build failed")
        self.n = lib().orc_num_tiles(self.h)
        self.n_leaf = lib().orc_num_leaf_tiles(self.h)
        self.max_level = lib().orc_max_level(self.h)

    @classmethod
    def adapted(cls, frm, to, ops=ADAPT_REFINE | ADAPT_COARSEN, fp64=False):
        """The tree of scene `frm` (a dict of scene parameters) adapted to the
        target levels of scene `to`: refinement steps, then coarsening steps
        (AdaptScene), with scene `to`'s cell types and RHS."""
        pf, pt = scene_params(**frm), scene_params(**to)
        h = lib().orc_create_adapted(ctypes.byref(pf), ctypes.byref(pt), int(ops), int(bool(fp64)))
        if not h:
            raise RuntimeError("oracle adaptation failed")
        return cls(params=pt, fp64=fp64, _handle=h)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def topology(self):
        lev = np.zeros(self.n, np.int32)
        crd = np.zeros((self.n, 3), np.int32)
        typ = np.zeros(self.n, np.uint8)
        nbr = np.zeros((self.n, 6), np.int32)
        lib().orc_export_topology(self.h, _ptr(lev), _ptr(crd), _ptr(typ), _ptr(nbr))
        return lev, crd, typ, nbr

    def celltypes(self):
        ct = np.zeros((self.n, 512), np.uint8)
        lib().orc_export_celltypes(self.h, _ptr(ct))
        return ct

    def set_celltypes(self, ct):
        ct = np.ascontiguousarray(ct, np.uint8)
        assert ct.shape == (self.n, 512)
        lib().orc_set_celltypes(self.h, _ptr(ct))

    def calc_celltypes_from_leafs(self):
        lib().orc_calc_celltypes_from_leafs(self.h)

    def get(self, ch):
        a = np.zeros((self.n, 512), np.float64)
        lib().orc_get_channel(self.h, ch, _ptr(a))
        return a

    def set(self, ch, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.shape == (self.n, 512)
        lib().orc_set_channel(self.h, ch, _ptr(a))

    def full_negative_laplacian(self, xch, axch, diag=False):
        lib().orc_full_negative_laplacian(self.h, xch, axch, int(diag))

    def negative_laplacian_same_level(self, level, subtree_level, types, xch, axch, diag=False):
        lib().orc_negative_laplacian_same_level(self.h, level, subtree_level, types, xch, axch, int(diag))

    def gs_color(self, level, xch, rhsch, color):
        lib().orc_gs_color(self.h, level, xch, rhsch, color)

    def jacobi_sweep(self, level, xch, rhsch, omega):
        lib().orc_jacobi_sweep(self.h, level, xch, rhsch, omega)

    def residual_restrict(self, level, xch, rhsch, cch):
        lib().orc_residual_restrict(self.h, level, xch, rhsch, cch)

    def prolongate(self, level, cxch, fxch):
        lib().orc_prolongate(self.h, level, cxch, fxch)

    def vcycle(self, xch, fch, rhsch, level_iters=1, coarsest_iters=10, smoother=SMOOTH_RBGS, omega=2.0 / 3.0,
               storage=STORE_F32):
        lib().orc_vcycle(self.h, xch, fch, rhsch, level_iters, coarsest_iters, smoother, omega, int(storage))

    def dot(self, a, b, types=LEAF):
        return lib().orc_dot(self.h, a, b, types)

    def recenter(self, ch):
        lib().orc_recenter(self.h, ch)

    def solve(self, max_iters=1000, tol=1e-6, level_iters=1, coarsest_iters=10, pure_neumann=False,
              smoother=SMOOTH_RBGS, omega=2.0 / 3.0, sync_stride=1, storage=STORE_F32):
        out = np.zeros(4, np.float64)
        it = lib().orc_solve(self.h, max_iters, tol, level_iters, coarsest_iters, int(pure_neumann),
                             smoother, omega, int(sync_stride), int(storage), _ptr(out))
        return {"iters": it, "rel_res": out[0], "bb": out[1], "rr": out[2], "seconds": out[3]}

    def amg_prepare(self):
        """AMGSolver::prepareTypesAndCoeffs: stored coefficients, coarse cell types."""
        lib().orc_amg_prepare(self.h)

    def amg_solve(self, max_iters=1000, tol=1e-6, level_iters=1, coarsest_iters=10, pure_neumann=False, sync_stride=1):
        out = np.zeros(3, np.float64)
        it = lib().orc_amg_solve(self.h, max_iters, tol, level_iters, coarsest_iters, int(pure_neumann),
                                 int(sync_stride), _ptr(out))
        return {"iters": it, "rel_res": out[0], "bb": out[1], "rr": out[2]}

    def amg_vcycle(self, level_iters=1, coarsest_iters=10):
        """z = M_amg^-1 r (channels CH_Z from CH_R; CH_AP and CH_TMP are scratch)."""
        lib().orc_amg_vcycle(self.h, level_iters, coarsest_iters)

    def amg_apply(self, level, xch, axch):
        lib().orc_amg_apply(self.h, level, xch, axch)

    def amg_coeffs(self):
        """[4, tiles, 512]: minus-face off-diagonals along x, y, z and the diagonal."""
        n = len(self.topology()[0])
        out = np.zeros((4, n, 512), np.float64)
        for a in range(4):
            lib().orc_amg_coeffs(self.h, a, _ptr(out[a]))
        return out

    def solve_refined(self, tol=1e-8, inner_tol=1e-4, max_outer=20, max_iters=1000, level_iters=1, coarsest_iters=10,
                      pure_neumann=False, smoother=SMOOTH_RBGS, omega=2.0 / 3.0, sync_stride=1, storage=STORE_BF16):
        """Mixed-precision iterative refinement (FP32 handles); x64 via refined_x()."""
        assert not self.fp64
        out = np.zeros(6, np.float64)
        k = lib().orc_solve_refined(self.h, tol, inner_tol, max_outer, max_iters, level_iters, coarsest_iters,
                                    int(pure_neumann), smoother, omega, int(sync_stride), int(storage), _ptr(out))
        return {"outer": k, "iters": int(out[3]), "converged": int(out[4]), "rel_res": out[0], "bb": out[1],
                "rr": out[2], "seconds": out[5]}

    def velocity_divergence(self, div_ch=CH_B, propagate=True):
        """[Ghost velocities from the parents,] then the volume divergence of
        channels 6-8 into div_ch (VelocityVolumeDivergenceOnLeafs)."""
        lib().orc_velocity_divergence(self.h, div_ch, int(propagate))

    def add_gradient(self, p_ch=CH_X):
        """u += grad p on the face velocities (AddGradientToFaceCenters)."""
        lib().orc_add_gradient(self.h, p_ch)

    def refined_x(self):
        a = np.zeros((self.n, 512), np.float64)
        assert lib().orc_get_refined_x(self.h, _ptr(a)) == 1
        return a

    def damped_jacobi_solve(self, max_iters, tol, omega):
        out = np.zeros(3, np.float64)
        it = lib().orc_damped_jacobi_solve(self.h, max_iters, tol, omega, _ptr(out))
        return {"iters": it, "rel_res": out[0], "bb": out[1], "rr": out[2]}


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()
