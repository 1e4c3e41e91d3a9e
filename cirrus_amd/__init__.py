"""cirrus_amd -- MI355X-native MGPCG pressure projection on adaptive octrees.

Python mirror of the C ABI in ``include/cirrus_mgpcg.h`` (ctypes over
``libcirrus_mgpcg.so``), shaped like the reference's solver API:

* :class:`Problem` -- a device-resident HAGrid (all LEAF/NONLEAF/GHOST tiles,
  cell types, channels x, b/r, p, Ap, z, tmp), the analogue of
  ``HADeviceGrid<PoissonTile<float>>`` (src/HAGrid.h:205).
* :class:`GMGSolver` -- ``GMGSolver(one_over_alpha, prolong_coeff)`` with
  ``solve``, ``VCycle`` and ``dampedJacobiSolve`` (src/GMGSolver.h:17-52).
* free functions mirroring src/GMGSolver.h:7-15 and src/PoissonSolver.h:8-31.
* :class:`Scene` -- the synthetic benchmark workloads.

The library must be built (``python -m cirrus_amd.build``); nothing here falls
back to a CPU path: a missing library or a failed HIP call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcirrus_mgpcg.so")

LEAF, NONLEAF, GHOST = 1, 2, 4
DIRICHLET, INTERIOR, NEUMANN = 1, 2, 4
CH_X, CH_B, CH_R, CH_P, CH_AP, CH_Z, CH_TMP = 0, 1, 1, 2, 3, 4, 5
CH_U, CH_V, CH_W = 6, 7, 8
NUM_CHANNELS = 6
SMOOTH_RBGS, SMOOTH_JACOBI = 0, 1
NOALLOC_MAX_TILES = 262144  # CG_NOALLOC_MAX_TILES
ADAPT_REFINE, ADAPT_COARSEN = 1, 2  # CG_ADAPT_*
STORE_F32, STORE_BF16, STORE_F16 = 0, 1, 2
SCENE_UNIFORM, SCENE_SPHERE_BAND, SCENE_PLANE, SCENE_VORTEX = 0, 1, 2, 3
(KERNEL_GS_SWEEP, KERNEL_APPLY, KERNEL_RESTRICT, KERNEL_PROLONG, KERNEL_JACOBI, KERNEL_UPDATE_XR, KERNEL_COARSE,
 KERNEL_EMPTY, KERNEL_COPY, KERNEL_GS_PROLONG, KERNEL_GS_SWEEP_BF16, KERNEL_GS_PROLONG_BF16,
 KERNEL_GS_SWEEP_F16, KERNEL_AMG_GS, KERNEL_GS_PAIR, KERNEL_GS_PAIR_BF16) = range(16)

EXPORTED_SYMBOLS = (
    "cg_params_default", "cg_last_error", "cg_device_count", "cg_problem_create", "cg_problem_create_batch",
    "cg_problem_num_problems", "cg_solve_batch", "cg_problem_destroy", "cg_problem_info",
    "cg_problem_topology", "cg_set_channel", "cg_get_channel", "cg_set_channel_device", "cg_get_channel_device",
    "cg_solve", "cg_damped_jacobi_solve",
    "cg_vcycle", "cg_full_negative_laplacian", "cg_negative_laplacian_same_level", "cg_gauss_seidel_color",
    "cg_jacobi_sweep", "cg_residual_restrict", "cg_prolongate", "cg_dot", "cg_dot_noalloc", "cg_recenter_leaf_voxels",
    "cg_calc_cell_types_from_leafs", "cg_benchmark_kernel", "cg_scene_create", "cg_scene_destroy", "cg_scene_tree", "cg_scene_velocity",
    "cg_server_create", "cg_server_destroy", "cg_server_admit", "cg_server_release", "cg_server_submit", "cg_server_step",
    "cg_server_next_result", "cg_server_info", "cg_server_perturb_rhs",
    "cg_refine_params_default", "cg_solve_refined", "cg_get_solution_f64",
    "cg_velocity_divergence", "cg_add_gradient", "cg_project", "cg_set_schedule", "cg_set_level_forms", "cg_autotune", "cg_autotune_refined",
    "cg_scene_tree_device", "cg_tree_adapt_device", "cg_tiles_save", "cg_tiles_load", "cg_tiles_import_holder", "cg_tileset_tree", "cg_tileset_channels",
    "cg_tileset_destroy", "cg_amg_prepare", "cg_amg_solve", "cg_amg_coefficients",
)


class TreeDesc(ctypes.Structure):
    _fields_ = [("h0", ctypes.c_double), ("num_tiles", ctypes.c_int32),
                ("level", ctypes.c_void_p), ("coord", ctypes.c_void_p),
                ("type", ctypes.c_void_p), ("cell_type", ctypes.c_void_p)]


class Params(ctypes.Structure):
    _fields_ = [("max_iters", ctypes.c_int32), ("rel_tol", ctypes.c_double), ("level_iters", ctypes.c_int32),
                ("coarsest_iters", ctypes.c_int32), ("smoother", ctypes.c_int32), ("omega", ctypes.c_float),
                ("sync_stride", ctypes.c_int32), ("pure_neumann", ctypes.c_int32),
                ("one_over_alpha", ctypes.c_float), ("prolong_coeff", ctypes.c_float), ("use_graph", ctypes.c_int32),
                ("storage", ctypes.c_int32), ("verify_f64", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int32), ("converged", ctypes.c_int32), ("rel_res", ctypes.c_double),
                ("bb", ctypes.c_double), ("rr", ctypes.c_double), ("seconds", ctypes.c_double),
                ("rel_res_f64", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RefineParams(ctypes.Structure):
    _fields_ = [("rel_tol", ctypes.c_double), ("inner_rel_tol", ctypes.c_double), ("max_outer", ctypes.c_int32)]


class RefineResult(ctypes.Structure):
    _fields_ = [("outer", ctypes.c_int32), ("iters", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("rel_res", ctypes.c_double), ("bb", ctypes.c_double), ("rr", ctypes.c_double),
                ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneReport(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_int32), ("base_seconds", ctypes.c_double), ("best_seconds", ctypes.c_double),
                ("coarse_shift", ctypes.c_int32), ("swizzle", ctypes.c_int32), ("tile_block", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SceneParams(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("nroot", ctypes.c_int32 * 3), ("h0", ctypes.c_double),
        ("uniform_level", ctypes.c_int32), ("has_solid", ctypes.c_int32), ("solid_c", ctypes.c_double * 3),
        ("solid_r", ctypes.c_double), ("band_c", ctypes.c_double * 3), ("band_r", ctypes.c_double),
        ("band_w", ctypes.c_double * 2), ("band_level", ctypes.c_int32 * 3), ("plane_z", ctypes.c_double),
        ("plane_level", ctypes.c_int32 * 2), ("n_tubes", ctypes.c_int32), ("tube_r_lo", ctypes.c_double),
        ("tube_r_hi", ctypes.c_double), ("vortex_thresh", ctypes.c_double), ("vortex_level", ctypes.c_int32 * 2),
        ("face_bc", ctypes.c_int32 * 6), ("seed", ctypes.c_uint64), ("num_layers", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_int64), ("max_problems", ctypes.c_int32), ("params", Params),
                ("refine", ctypes.c_int32)]


class CirrusError(RuntimeError):
    pass


class CirrusBusy(CirrusError):
    """No capacity now (HBM budget or problem slots); retry later."""


_lib = None


def lib():
    """Load libcirrus_mgpcg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CirrusError(f"{LIB_PATH} not built: run `python -m cirrus_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, f32, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_double
        P = ctypes.POINTER
        sig = {
            "cg_params_default": (None, [P(Params)]),
            "cg_last_error": (ctypes.c_char_p, []),
            "cg_device_count": (i32, []),
            "cg_problem_create": (i32, [P(TreeDesc), P(vp)]),
            "cg_problem_create_batch": (i32, [P(TreeDesc), i32, P(vp)]),
            "cg_problem_num_problems": (i32, [vp, P(i32)]),
            "cg_solve_batch": (i32, [vp, P(Params), P(Result)]),
            "cg_problem_destroy": (i32, [vp]),
            "cg_problem_info": (i32, [vp, P(i32), P(i32), P(i32)]),
            "cg_problem_topology": (i32, [vp, vp, vp]),
            "cg_set_channel": (i32, [vp, i32, vp]),
            "cg_get_channel": (i32, [vp, i32, vp]),
            "cg_set_channel_device": (i32, [vp, i32, vp, vp]),
            "cg_get_channel_device": (i32, [vp, i32, vp, vp]),
            "cg_solve": (i32, [vp, P(Params), P(Result)]),
            "cg_damped_jacobi_solve": (i32, [vp, i32, f64, f32, P(Result)]),
            "cg_vcycle": (i32, [vp, i32, i32, i32, i32, i32, f32]),
            "cg_full_negative_laplacian": (i32, [vp, i32, i32, i32]),
            "cg_negative_laplacian_same_level": (i32, [vp, i32, i32, i32, i32, i32, i32]),
            "cg_gauss_seidel_color": (i32, [vp, i32, i32, i32, i32]),
            "cg_jacobi_sweep": (i32, [vp, i32, i32, i32, f32]),
            "cg_residual_restrict": (i32, [vp, i32, i32, i32, i32, f32]),
            "cg_prolongate": (i32, [vp, i32, i32, i32, f32]),
            "cg_dot": (i32, [vp, i32, i32, i32, P(f64)]),
            "cg_dot_noalloc": (i32, [vp, i32, i32, i32, P(f64)]),
            "cg_recenter_leaf_voxels": (i32, [vp, i32]),
            "cg_calc_cell_types_from_leafs": (i32, [P(TreeDesc), vp]),
            "cg_benchmark_kernel": (i32, [vp, i32, i32, i32, P(f64), P(ctypes.c_int64)]),
            "cg_server_create": (i32, [P(ServerConfig), P(vp)]),
            "cg_server_destroy": (i32, [vp]),
            "cg_server_admit": (i32, [vp, P(TreeDesc), P(i32)]),
            "cg_server_release": (i32, [vp, i32]),
            "cg_server_submit": (i32, [vp, i32, vp, i32, ctypes.c_int64]),
            "cg_server_perturb_rhs": (i32, [vp, i32, f32, f32, ctypes.c_uint64]),
            "cg_server_step": (i32, [vp, P(i32)]),
            "cg_server_next_result": (i32, [vp, i32, P(i32), P(ctypes.c_int64), P(i32), P(Result), vp]),
            "cg_server_info": (i32, [vp, P(ctypes.c_int64), P(ctypes.c_int64), P(i32)]),
            "cg_refine_params_default": (None, [P(RefineParams)]),
            "cg_velocity_divergence": (i32, [vp, i32, i32]),
            "cg_add_gradient": (i32, [vp, i32]),
            "cg_project": (i32, [vp, P(Params), P(Result)]),
            "cg_set_schedule": (i32, [vp, i32, i32, i32]),
            "cg_set_level_forms": (i32, [vp, i32, i32, i32]),
            "cg_autotune": (i32, [vp, P(Params), i32, P(Params), P(TuneReport)]),
            "cg_autotune_refined": (i32, [vp, P(Params), P(RefineParams), i32, P(Params), P(TuneReport)]),
            "cg_solve_refined": (i32, [vp, P(Params), P(RefineParams), P(RefineResult)]),
            "cg_get_solution_f64": (i32, [vp, vp]),
            "cg_scene_create": (i32, [P(SceneParams), P(vp)]),
            "cg_scene_tree_device": (i32, [P(SceneParams), ctypes.c_int64, P(vp)]),
            "cg_tree_adapt_device": (i32, [P(TreeDesc), P(SceneParams), i32, ctypes.c_int64, P(vp)]),
            "cg_scene_destroy": (i32, [vp]),
            "cg_scene_tree": (i32, [vp, P(TreeDesc), P(vp)]),
            "cg_scene_velocity": (i32, [vp, P(vp)]),
            "cg_tiles_save": (i32, [ctypes.c_char_p, P(TreeDesc), i32, vp, vp]),
            "cg_tiles_load": (i32, [ctypes.c_char_p, P(vp)]),
            "cg_tiles_import_holder": (i32, [ctypes.c_char_p, i32, vp, P(vp)]),
            "cg_tileset_tree": (i32, [vp, P(TreeDesc)]),
            "cg_tileset_channels": (i32, [vp, P(i32), P(vp), P(vp)]),
            "cg_tileset_destroy": (i32, [vp]),
            "cg_amg_prepare": (i32, [vp, f32]),
            "cg_amg_solve": (i32, [vp, P(Params), P(Result)]),
            "cg_amg_coefficients": (i32, [vp, i32, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def device_count():
    """Number of visible HIP devices (0 without a GPU)."""
    return int(lib().cg_device_count())


CG_ERR_BUSY = 5


def _check(st, what):
    if st != 0:
        msg = lib().cg_last_error()
        cls = CirrusBusy if st == CG_ERR_BUSY else CirrusError
        raise cls(f"{what} failed (status {st}): {msg.decode() if msg else ''}")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def scene_params(**kw):
    """SceneParams with the plain configuration's defaults: uniform 64^3 (level 3),
    solid sphere r=0.2 at the centre (NEUMANN), top face DIRICHLET, other walls
    NEUMANN, seed 0."""
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


class Tree:
    """A host-side HAGrid description: tile levels, coordinates, types and cell types."""

    def __init__(self, h0, level, coord, type_, cell_type):
        self.h0 = float(h0)
        self.level = np.ascontiguousarray(level, np.int32)
        self.coord = np.ascontiguousarray(coord, np.int32).reshape(-1, 3)
        self.type = np.ascontiguousarray(type_, np.uint8)
        self.cell_type = np.ascontiguousarray(cell_type, np.uint8).reshape(-1, 512)
        self.n = len(self.level)

    def desc(self):
        d = TreeDesc()
        d.h0 = self.h0
        d.num_tiles = self.n
        d.level = self.level.ctypes.data
        d.coord = self.coord.ctypes.data
        d.type = self.type.ctypes.data
        d.cell_type = self.cell_type.ctypes.data
        return d


def calc_cell_types_from_leafs(tree):
    """CalcCellTypesFromLeafs (src/PoissonSolver.cu:7-61) on a host tree (in place)."""
    d = tree.desc()
    ct = np.ascontiguousarray(tree.cell_type)
    _check(lib().cg_calc_cell_types_from_leafs(ctypes.byref(d), _ptr(ct)), "cg_calc_cell_types_from_leafs")
    tree.cell_type = ct
    return ct


class Scene:
    """A synthetic workload: the octree, its cell types, the face velocities and
    the divergence right-hand side, built on the host by the library."""

    def __init__(self, params=None, **kw):
        self.params = params if params is not None else scene_params(**kw)
        h = ctypes.c_void_p()
        _check(lib().cg_scene_create(ctypes.byref(self.params), ctypes.byref(h)), "cg_scene_create")
        self._h = h
        d = TreeDesc()
        rhs = ctypes.c_void_p()
        _check(lib().cg_scene_tree(h, ctypes.byref(d), ctypes.byref(rhs)), "cg_scene_tree")
        n = d.num_tiles

        def arr(ptr, ctype, count):
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(count,)).copy()

        self.tree = Tree(d.h0, arr(d.level, ctypes.c_int32, n), arr(d.coord, ctypes.c_int32, 3 * n),
                         arr(d.type, ctypes.c_uint8, n), arr(d.cell_type, ctypes.c_uint8, 512 * n))
        self.rhs = arr(rhs, ctypes.c_float, 512 * n).reshape(n, 512)
        uvw = ctypes.c_void_p()
        _check(lib().cg_scene_velocity(h, ctypes.byref(uvw)), "cg_scene_velocity")
        self.uvw = arr(uvw, ctypes.c_float, 3 * 512 * n).reshape(n, 3, 512)
        lib().cg_scene_destroy(h)
        self._h = None


def save_tiles(path, tree, channels=None):
    """Write a tree and channels ({channel number: [tiles, 512] float32}) to a
    portable tile file (cg_tiles_save)."""
    channels = channels or {}
    ids = np.array(sorted(channels), np.int32)
    arrs = [np.ascontiguousarray(channels[int(c)], np.float32).reshape(tree.n, 512) for c in ids]
    ptrs = (ctypes.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
    d = tree.desc()
    _check(lib().cg_tiles_save(os.fsencode(path), ctypes.byref(d), len(ids), _ptr(ids), ptrs), "cg_tiles_save")


def _tileset(h):
    try:
        d = TreeDesc()
        _check(lib().cg_tileset_tree(h, ctypes.byref(d)), "cg_tileset_tree")
        n = d.num_tiles

        def arr(ptr, ctype, count):
            if count == 0:
                return np.zeros(0, np.dtype(ctype))
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(count,)).copy()

        tree = Tree(d.h0, arr(d.level, ctypes.c_int32, n), arr(d.coord, ctypes.c_int32, 3 * n),
                    arr(d.type, ctypes.c_uint8, n), arr(d.cell_type, ctypes.c_uint8, 512 * n))
        m = ctypes.c_int32()
        ids, data = ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().cg_tileset_channels(h, ctypes.byref(m), ctypes.byref(ids), ctypes.byref(data)), "cg_tileset_channels")
        idv = arr(ids, ctypes.c_int32, m.value)
        dv = arr(data, ctypes.c_float, m.value * 512 * n).reshape(m.value, n, 512)
        return tree, {int(c): dv[q] for q, c in enumerate(idv)}
    finally:
        lib().cg_tileset_destroy(h)


def scene_tree_device(max_tiles=0, **kw):
    """The scene's octree and cell types built on the device (uniform, sphere
    band and plane scenes; cg_scene_tree_device): a Tree, tiles in the
    device build's order."""
    p = scene_params(**kw)
    h = ctypes.c_void_p()
    _check(lib().cg_scene_tree_device(ctypes.byref(p), int(max_tiles), ctypes.byref(h)), "cg_scene_tree_device")
    return _tileset(h)[0]


def tree_adapt_device(tree, ops=ADAPT_REFINE | ADAPT_COARSEN, max_tiles=0, **kw):
    """A Tree adapted on the device to the target levels of the scene given by
    the keywords (cg_tree_adapt_device): refinement rounds, then CoarsenStep
    until none deletes, then ghosts and the scene's cell types."""
    p = scene_params(**kw)
    d = tree.desc()
    h = ctypes.c_void_p()
    _check(lib().cg_tree_adapt_device(ctypes.byref(d), ctypes.byref(p), int(ops), int(max_tiles), ctypes.byref(h)),
           "cg_tree_adapt_device")
    return _tileset(h)[0]


def load_tiles(path):
    """Read a portable tile file: (Tree, {channel number: [tiles, 512] float32})."""
    h = ctypes.c_void_p()
    _check(lib().cg_tiles_load(os.fsencode(path), ctypes.byref(h)), "cg_tiles_load")
    return _tileset(h)


def import_holder(path, channels=(0, 1, 6, 7, 8)):
    """Read the reference's HAHostTileHolder dump (src/PoissonIOFunc.cu:38-124):
    (Tree, {channel number: [tiles, 512] float32}) for the given channels."""
    ids = np.array(channels, np.int32)
    h = ctypes.c_void_p()
    _check(lib().cg_tiles_import_holder(os.fsencode(path), len(ids), _ptr(ids), ctypes.byref(h)),
           "cg_tiles_import_holder")
    return _tileset(h)


class Problem:
    """A device-resident hierarchy (cg_problem); arrays are in the tree's tile order.

    ``Problem.batch(trees)`` holds several independent trees in one handle
    (solved together by :meth:`GMGSolver.solve_batch`); its arrays concatenate
    the trees in order, and :meth:`split` cuts such an array per tree."""

    def __init__(self, tree, _batch=None):
        trees = [tree] if _batch is None else list(_batch)
        self.tree = trees[0]
        self.trees = trees
        h = ctypes.c_void_p()
        descs = (TreeDesc * len(trees))(*[t.desc() for t in trees])
        _check(lib().cg_problem_create_batch(descs, len(trees), ctypes.byref(h)), "cg_problem_create")
        self._h = h
        n, nl, ml = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        lib().cg_problem_info(h, ctypes.byref(n), ctypes.byref(nl), ctypes.byref(ml))
        self.n, self.n_leaf, self.max_level = n.value, nl.value, ml.value
        self.num_problems = len(trees)
        self.offsets = np.cumsum([0] + [t.n for t in trees])

    @classmethod
    def batch(cls, trees):
        return cls(None, _batch=trees)

    def split(self, a):
        """Per-tree views of a [tiles, ...] array of this handle."""
        return [a[self.offsets[q]:self.offsets[q + 1]] for q in range(self.num_problems)]

    def close(self):
        if getattr(self, "_h", None):
            lib().cg_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def topology(self):
        nbr = np.zeros((self.n, 6), np.int32)
        par = np.zeros(self.n, np.int32)
        _check(lib().cg_problem_topology(self._h, _ptr(nbr), _ptr(par)), "cg_problem_topology")
        return nbr, par

    def set(self, ch, a):
        a = np.ascontiguousarray(a, np.float32).reshape(self.n, 512)
        _check(lib().cg_set_channel(self._h, ch, _ptr(a)), "cg_set_channel")

    def get(self, ch):
        a = np.zeros((self.n, 512), np.float32)
        _check(lib().cg_get_channel(self._h, ch, _ptr(a)), "cg_get_channel")
        return a

    def set_device(self, ch, device_ptr, stream=None):
        """Channel from a device buffer (512 float32 per tile, this handle's
        tile order; e.g. a torch CUDA tensor's data_ptr()), ordered on the
        caller's stream (a hipStream_t as an int, None: the null stream)."""
        _check(lib().cg_set_channel_device(self._h, ch, ctypes.c_void_p(int(device_ptr)),
                                           ctypes.c_void_p(int(stream) if stream else None)), "cg_set_channel_device")

    def get_device(self, ch, device_ptr, stream=None):
        """Channel into a device buffer (see :meth:`set_device`)."""
        _check(lib().cg_get_channel_device(self._h, ch, ctypes.c_void_p(int(device_ptr)),
                                           ctypes.c_void_p(int(stream) if stream else None)), "cg_get_channel_device")

    def set_schedule(self, coarse_shift=0, swizzle=True, tile_block=1):
        """Kernel coverage only, never a result: LDS-resident coarse levels given
        back to grid-wide launches, XCD-aware tile order, tiles per wavefront of
        the colour sweeps (cg_set_schedule)."""
        _check(lib().cg_set_schedule(self._h, int(coarse_shift), int(bool(swizzle)), int(tile_block)),
               "cg_set_schedule")

    def set_level_forms(self, eager_max_tiles=-1, fused16_max_tiles=-1, pair_sweeps=-1):
        """Level-size limits of the eager prolongation/restriction forms and of
        the fused first 16-bit sweep, and whether red-black iterations after
        the first run as one launch of both colours (cg_set_level_forms;
        results unchanged; 0 never, negative the default)."""
        _check(lib().cg_set_level_forms(self._h, int(eager_max_tiles), int(fused16_max_tiles), int(pair_sweeps)),
               "cg_set_level_forms")

    def get_f64(self):
        """The FP64 x of the last refined solve, [tiles, 512] float64."""
        a = np.zeros((self.n, 512), np.float64)
        _check(lib().cg_get_solution_f64(self._h, _ptr(a)), "cg_get_solution_f64")
        return a

    def benchmark_kernel(self, kernel, level=0, reps=50):
        """(average launch duration in us, tiles per launch) of one kernel of
        the solve, timed with hipEvents on the problem's stream."""
        us = ctypes.c_double()
        nt = ctypes.c_int64()
        _check(lib().cg_benchmark_kernel(self._h, kernel, level, reps, ctypes.byref(us), ctypes.byref(nt)),
               "cg_benchmark_kernel")
        return us.value, nt.value

    def dot(self, a, b, types=LEAF):
        out = ctypes.c_double()
        _check(lib().cg_dot(self._h, a, b, types, ctypes.byref(out)), "cg_dot")
        return out.value

    def dot_noalloc(self, a, b, types=LEAF):
        """dot() through the entry point that allocates nothing (at most
        NOALLOC_MAX_TILES tiles; calls overlapping on one GPU race)."""
        out = ctypes.c_double()
        _check(lib().cg_dot_noalloc(self._h, a, b, types, ctypes.byref(out)), "cg_dot_noalloc")
        return out.value


# --- operator entry points (src/GMGSolver.h:7-15, src/GMGSolver.cu) -----------
def FullNegativeLaplacian(prob, x_channel, Ax_channel, calc_diag=False):
    _check(lib().cg_full_negative_laplacian(prob.handle, x_channel, Ax_channel, int(calc_diag)), "FullNegativeLaplacian")


def VelocityVolumeDivergenceOnLeafs(prob, div_channel=CH_B, propagate_ghosts=True):
    """Divergence of the face velocities (channels 6-8) into div_channel
    (src/PoissonSolver.cu:66-142), ghost velocities first taken from the parents."""
    _check(lib().cg_velocity_divergence(prob.handle, div_channel, int(propagate_ghosts)), "VelocityVolumeDivergenceOnLeafs")


def AddGradientToFaceCenters(prob, p_channel=CH_X):
    """u += grad p on the face velocities (src/PoissonSolver.cu:144-209)."""
    _check(lib().cg_add_gradient(prob.handle, p_channel), "AddGradientToFaceCenters")


def NegativeLaplacianSameLevel(prob, level, subtree_level, tile_types, x_channel, Ax_channel, calc_diag=False):
    _check(lib().cg_negative_laplacian_same_level(prob.handle, level, subtree_level, tile_types, x_channel,
                                                  Ax_channel, int(calc_diag)), "NegativeLaplacianSameLevel")


def GaussSeidelColor(prob, level, x_channel, rhs_channel, color):
    _check(lib().cg_gauss_seidel_color(prob.handle, level, x_channel, rhs_channel, color), "GaussSeidelColor")


def JacobiSweep(prob, level, x_channel, rhs_channel, omega):
    _check(lib().cg_jacobi_sweep(prob.handle, level, x_channel, rhs_channel, omega), "JacobiSweep")


def ResidualAndRestrict(prob, level, x_channel, rhs_channel, coarse_channel, one_over_alpha=1.0):
    _check(lib().cg_residual_restrict(prob.handle, level, x_channel, rhs_channel, coarse_channel, one_over_alpha),
           "ResidualAndRestrict")


def ProlongateAndUpdate(prob, level, coarse_channel, fine_channel, prolong_coeff=1.0):
    _check(lib().cg_prolongate(prob.handle, level, coarse_channel, fine_channel, prolong_coeff), "ProlongateAndUpdate")


def ReCenterLeafVoxels(prob, channel):
    _check(lib().cg_recenter_leaf_voxels(prob.handle, channel), "ReCenterLeafVoxels")


def _graph_mode(use_graph):
    """cg_params.use_graph from a Python value: None -> -1 (chosen by the handle)."""
    return -1 if use_graph is None else int(bool(use_graph))


def default_params(**kw):
    p = Params()
    lib().cg_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class AMGSolver:
    """AMGSolver (src/AMGSolver.h:17-53) over a :class:`Problem`: the MGPCG with
    the stored-coefficient V-cycle.  ``prepareTypesAndCoeffs`` first (cg_amg_prepare),
    then ``solve`` (cg_amg_solve; one result per problem kept in ``self.last_batch``)."""

    def __init__(self, one_over_alpha=1.0, prolong_coeff=1.0):
        self.one_over_alpha = float(one_over_alpha)
        self.prolong_coeff = float(prolong_coeff)

    def prepareTypesAndCoeffs(self, prob):
        _check(lib().cg_amg_prepare(prob.handle, self.one_over_alpha), "AMGSolver.prepareTypesAndCoeffs")

    def solve(self, prob, verbose=False, max_iters=1000, relative_tolerance=1e-6, level_iters=1, coarsest_iters=10,
              sync_stride=1, is_pure_neumann=False, use_graph=None):
        """Returns (iters, relative residual) of the first problem like the reference."""
        p = default_params(max_iters=max_iters, rel_tol=relative_tolerance, level_iters=level_iters,
                           coarsest_iters=coarsest_iters, sync_stride=int(sync_stride),
                           pure_neumann=int(is_pure_neumann), one_over_alpha=self.one_over_alpha,
                           prolong_coeff=self.prolong_coeff, use_graph=_graph_mode(use_graph))
        res = (Result * prob.num_problems)()
        _check(lib().cg_amg_solve(prob.handle, ctypes.byref(p), res), "AMGSolver.solve")
        self.last_batch = [r.as_dict() for r in res]
        if verbose:
            print(f"AMG-PCG {res[0].iters} iters, rel residual {res[0].rel_res:.3e}")
        return res[0].iters, res[0].rel_res

    @staticmethod
    def coefficients(prob):
        """[5, tiles, 512] float32: minus-face entries along x, y, z, the
        diagonal, the AMG cell types (cg_amg_coefficients)."""
        out = np.zeros((5, prob.n, 512), np.float32)
        for w in range(5):
            _check(lib().cg_amg_coefficients(prob.handle, w, _ptr(out[w])), "cg_amg_coefficients")
        return out


class GMGSolver:
    """GMGSolver (src/GMGSolver.h:17-52) over a :class:`Problem`."""

    def __init__(self, one_over_alpha=1.0, prolong_coeff=1.0):
        self.one_over_alpha = float(one_over_alpha)
        self.prolong_coeff = float(prolong_coeff)

    def solve(self, prob, verbose=False, max_iters=1000, relative_tolerance=1e-6, level_iters=1, coarsest_iters=10,
              sync_stride=1, is_pure_neumann=False, smoother=SMOOTH_RBGS, omega=2.0 / 3.0, use_graph=None,
              storage=STORE_F32, verify_f64=False):
        """Returns (iters, relative residual) like the reference; the full result
        is kept in ``self.last`` (with verify_f64, ``rel_res_f64`` is the FP64
        recomputed |b - A x| / |b|)."""
        p = default_params(max_iters=max_iters, rel_tol=relative_tolerance, level_iters=level_iters,
                           coarsest_iters=coarsest_iters, smoother=smoother, omega=omega,
                           sync_stride=int(sync_stride), pure_neumann=int(is_pure_neumann),
                           one_over_alpha=self.one_over_alpha, prolong_coeff=self.prolong_coeff,
                           use_graph=_graph_mode(use_graph), storage=int(storage), verify_f64=int(verify_f64))
        r = Result()
        _check(lib().cg_solve(prob.handle, ctypes.byref(p), ctypes.byref(r)), "GMGSolver.solve")
        self.last = r.as_dict()
        if verbose:
            print(f"MGPCG {r.iters} iters, rel residual {r.rel_res:.3e}, {r.seconds * 1e3:.3f} ms")
        return r.iters, r.rel_res

    def VCycle(self, prob, x_channel, f_channel, level_iters=1, coarsest_iters=10, smoother=SMOOTH_RBGS,
               omega=2.0 / 3.0):
        _check(lib().cg_vcycle(prob.handle, x_channel, f_channel, level_iters, coarsest_iters, smoother, omega),
               "GMGSolver.VCycle")

    def solve_batch(self, prob, max_iters=1000, relative_tolerance=1e-6, level_iters=1, coarsest_iters=10,
                    sync_stride=1, is_pure_neumann=False, smoother=SMOOTH_RBGS, omega=2.0 / 3.0, use_graph=None,
                    storage=STORE_F32, verify_f64=False):
        """All problems of a batched :class:`Problem`; returns one result dict
        per problem (iters, converged, rel_res, bb, rr, seconds, rel_res_f64)."""
        p = default_params(max_iters=max_iters, rel_tol=relative_tolerance, level_iters=level_iters,
                           coarsest_iters=coarsest_iters, smoother=smoother, omega=omega,
                           sync_stride=int(sync_stride), pure_neumann=int(is_pure_neumann),
                           one_over_alpha=self.one_over_alpha, prolong_coeff=self.prolong_coeff,
                           use_graph=_graph_mode(use_graph), storage=int(storage), verify_f64=int(verify_f64))
        res = (Result * prob.num_problems)()
        _check(lib().cg_solve_batch(prob.handle, ctypes.byref(p), res), "GMGSolver.solve_batch")
        self.last_batch = [r.as_dict() for r in res]
        return self.last_batch

    def solve_refined(self, prob, relative_tolerance=1e-8, inner_tolerance=1e-4, max_outer=20, max_iters=1000,
                      level_iters=1, coarsest_iters=10, sync_stride=1, is_pure_neumann=False, smoother=SMOOTH_RBGS,
                      omega=2.0 / 3.0, use_graph=None, storage=STORE_BF16):
        """Mixed-precision form: iterative refinement with FP64 x and residual
        around the MGPCG (cg_solve_refined); one result dict per problem
        (outer, iters, converged, rel_res, bb, rr, seconds).  The FP64 x is
        :meth:`Problem.get_f64`; CH_X holds it rounded to FP32."""
        p = default_params(max_iters=max_iters, rel_tol=relative_tolerance, level_iters=level_iters,
                           coarsest_iters=coarsest_iters, smoother=smoother, omega=omega,
                           sync_stride=int(sync_stride), pure_neumann=int(is_pure_neumann),
                           one_over_alpha=self.one_over_alpha, prolong_coeff=self.prolong_coeff,
                           use_graph=_graph_mode(use_graph), storage=int(storage))
        rp = RefineParams()
        lib().cg_refine_params_default(ctypes.byref(rp))
        rp.rel_tol, rp.inner_rel_tol, rp.max_outer = float(relative_tolerance), float(inner_tolerance), int(max_outer)
        res = (RefineResult * prob.num_problems)()
        _check(lib().cg_solve_refined(prob.handle, ctypes.byref(p), ctypes.byref(rp), res), "GMGSolver.solve_refined")
        self.last_refined = [r.as_dict() for r in res]
        return self.last_refined

    def project(self, prob, max_iters=1000, relative_tolerance=1e-6, level_iters=1, coarsest_iters=10,
                is_pure_neumann=False, smoother=SMOOTH_RBGS, storage=STORE_F32):
        """FluidEuler::project without the velocity BCs (cg_project): divergence of
        the face velocities, the MGPCG solve, the gradient added back; one result
        dict per problem."""
        p = default_params(max_iters=max_iters, rel_tol=relative_tolerance, level_iters=level_iters,
                           coarsest_iters=coarsest_iters, smoother=smoother, pure_neumann=int(is_pure_neumann),
                           one_over_alpha=self.one_over_alpha, prolong_coeff=self.prolong_coeff, storage=int(storage))
        res = (Result * prob.num_problems)()
        _check(lib().cg_project(prob.handle, ctypes.byref(p), res), "GMGSolver.project")
        return [r.as_dict() for r in res]

    def autotune(self, prob, max_iters=1000, relative_tolerance=1e-6, level_iters=1, coarsest_iters=10,
                 smoother=SMOOTH_RBGS, storage=STORE_F32, allow_bf16=False):
        """cg_autotune on the problem's current right-hand sides: returns (the
        chosen solver keyword arguments, the report); the handle keeps the
        chosen schedule."""
        p = default_params(max_iters=max_iters, rel_tol=relative_tolerance, level_iters=level_iters,
                           coarsest_iters=coarsest_iters, smoother=smoother, storage=int(storage),
                           one_over_alpha=self.one_over_alpha, prolong_coeff=self.prolong_coeff)
        best = Params()
        rep = TuneReport()
        _check(lib().cg_autotune(prob.handle, ctypes.byref(p), int(bool(allow_bf16)), ctypes.byref(best),
                                 ctypes.byref(rep)), "GMGSolver.autotune")
        kw = dict(level_iters=best.level_iters, smoother=best.smoother, storage=best.storage)
        return kw, rep.as_dict()

    def autotune_refined(self, prob, relative_tolerance=1e-6, inner_tolerance=5e-7, max_outer=20, max_iters=1000,
                         level_iters=1, coarsest_iters=10, smoother=SMOOTH_RBGS, storage=STORE_BF16, allow_16bit=True):
        """cg_autotune_refined: the same search, every candidate timed as the
        FP64-refined solve (:meth:`solve_refined` with these tolerances);
        returns (the chosen solver keyword arguments, the report)."""
        p = default_params(max_iters=max_iters, rel_tol=relative_tolerance, level_iters=level_iters,
                           coarsest_iters=coarsest_iters, smoother=smoother, storage=int(storage),
                           one_over_alpha=self.one_over_alpha, prolong_coeff=self.prolong_coeff)
        rp = RefineParams()
        lib().cg_refine_params_default(ctypes.byref(rp))
        rp.rel_tol, rp.inner_rel_tol, rp.max_outer = float(relative_tolerance), float(inner_tolerance), int(max_outer)
        best = Params()
        rep = TuneReport()
        _check(lib().cg_autotune_refined(prob.handle, ctypes.byref(p), ctypes.byref(rp), int(bool(allow_16bit)),
                                         ctypes.byref(best), ctypes.byref(rep)), "GMGSolver.autotune_refined")
        kw = dict(level_iters=best.level_iters, smoother=best.smoother, storage=best.storage)
        return kw, rep.as_dict()

    def dampedJacobiSolve(self, prob, verbose=False, max_iters=100, relative_tolerance=1e-6, omega=2.0 / 3.0):
        r = Result()
        _check(lib().cg_damped_jacobi_solve(prob.handle, max_iters, relative_tolerance, omega, ctypes.byref(r)),
               "GMGSolver.dampedJacobiSolve")
        self.last = r.as_dict()
        return r.iters, r.rel_res


class Server:
    """The serving loop (cg_server_*): resident instances, requests with warm
    starts, one CG iteration over all running requests per :meth:`step`."""

    def __init__(self, budget_bytes, max_problems, refine=False, **params):
        """params: cg_params fields (level_iters, storage, verify_f64, ...);
        refine: every request runs as an FP64 iterative refinement."""
        cfg = ServerConfig()
        cfg.budget_bytes = int(budget_bytes)
        cfg.max_problems = int(max_problems)
        cfg.refine = int(bool(refine))
        lib().cg_params_default(ctypes.byref(cfg.params))
        for k, v in params.items():
            setattr(cfg.params, k, v)
        h = ctypes.c_void_p()
        _check(lib().cg_server_create(ctypes.byref(cfg), ctypes.byref(h)), "cg_server_create")
        self._h = h
        self.ntiles = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().cg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, tree):
        """Make a tree resident; returns its instance id (raises CirrusBusy when it does not fit)."""
        inst = ctypes.c_int32()
        d = tree.desc()
        _check(lib().cg_server_admit(self._h, ctypes.byref(d), ctypes.byref(inst)), "cg_server_admit")
        self.ntiles[inst.value] = tree.n
        return inst.value

    def release(self, inst):
        _check(lib().cg_server_release(self._h, inst), "cg_server_release")
        self.ntiles.pop(inst, None)

    def submit(self, inst, rhs, warm=False, tag=0, download=True):
        """Queue a request; rhs None uses the right-hand side already on the device."""
        flags = (1 if warm else 0) | (0 if download else 2)
        if rhs is None:
            _check(lib().cg_server_submit(self._h, inst, None, flags, int(tag)), "cg_server_submit")
            return
        rhs = np.ascontiguousarray(rhs, np.float32).reshape(self.ntiles[inst], 512)
        _check(lib().cg_server_submit(self._h, inst, _ptr(rhs), flags, int(tag)), "cg_server_submit")

    def perturb_rhs(self, inst, rho, sigma, seed):
        _check(lib().cg_server_perturb_rhs(self._h, inst, float(rho), float(sigma), int(seed)), "cg_server_perturb_rhs")

    def step(self):
        n = ctypes.c_int32()
        _check(lib().cg_server_step(self._h, ctypes.byref(n)), "cg_server_step")
        return n.value

    def next_result(self, wait=False, with_x=True):
        """(tag, instance, result dict, pressure or None) of the oldest finished
        request, or None if none is ready."""
        nmax = max(self.ntiles.values()) if self.ntiles else 1
        x = np.zeros((max(nmax, 1), 512), np.float32) if with_x else None
        got, tag, inst = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
        res = Result()
        _check(lib().cg_server_next_result(self._h, int(wait), ctypes.byref(got), ctypes.byref(tag), ctypes.byref(inst),
                                           ctypes.byref(res), _ptr(x) if with_x else None), "cg_server_next_result")
        if not got.value:
            return None
        xs = x[:self.ntiles[inst.value]].copy() if with_x else None
        return tag.value, inst.value, res.as_dict(), xs

    def info(self):
        cap, free, ninst = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        _check(lib().cg_server_info(self._h, ctypes.byref(cap), ctypes.byref(free), ctypes.byref(ninst)), "cg_server_info")
        return {"capacity_tiles": cap.value, "free_tiles": free.value, "instances": ninst.value}
