"""Device-pointer channel moves (cg_set_channel_device / cg_get_channel_device):
the drop-in of the reference's in-place device tiles (GMGSolver::solve on
HADeviceGrid, src/GMGSolver.cu:531-666; FluidEuler::project,
src/FluidEuler.h:581-612) with no host staging.

The caller's device buffers and stream come from the HIP runtime the library
itself uses (libamdhip64.so.7, through ctypes), as a C++ caller's would.  A
solve fed and read through device pointers must equal the host-staged solve
and the oracle bit for bit, in the caller's tile order (here a shuffled one, so
the permutation to the canonical order is exercised); a host pointer or a
too-small device buffer is refused with an error before any kernel runs.  A
last test runs the same exchange with torch tensors in a child process that
imports torch first, so that torch and the library share one HIP runtime.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    ca.lib()  # the library's runtime is loaded first; the soname resolves to it
    h = ctypes.CDLL("libamdhip64.so.7")
    for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize"):
        getattr(h, name).restype = ctypes.c_int
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


class DevBuf:
    def __init__(self, h, nbytes):
        self.h, self.p = h, ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(nbytes)) == 0

    def put(self, a):
        assert self.h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def take(self, a):
        assert self.h.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.h.hipFree(self.p)


def shuffled_problem(seed=0):
    """The plane scene of the oracle with its tiles handed over in a random
    order: (oracle, problem, permutation, b in the oracle's order, types)."""
    o = po.Oracle(kind=po.SCENE_PLANE, plane_level=[3, 2])
    lev, crd, typ, _ = o.topology()
    ct = o.celltypes()
    perm = np.random.default_rng(seed).permutation(len(lev))
    P = ca.Problem(ca.Tree(o.params.h0, lev[perm], crd[perm], typ[perm], ct[perm]))
    b = o.get(po.CH_B).astype(np.float32)
    return o, P, perm, b, typ


def test_device_moves_match_host_moves_and_oracle():
    o, P, perm, b, typ = shuffled_problem()
    S = ca.GMGSolver()
    P.set(ca.CH_B, b[perm])
    it_h, _ = S.solve(P, False, 200, 1e-6, 2, 10, 1, False)
    x_host = P.get(ca.CH_X)
    h = hip()
    st = ctypes.c_void_p()
    assert h.hipStreamCreate(ctypes.byref(st)) == 0
    nb = P.n * 512 * 4
    bd, xd = DevBuf(h, nb), DevBuf(h, nb)
    try:
        bd.put(np.ascontiguousarray(b[perm]))
        xd.put(np.full((P.n, 512), 7.0, np.float32))
        P.set(ca.CH_X, np.full((P.n, 512), -3.0, np.float32))
        P.set_device(ca.CH_B, bd.p.value, st.value)
        S.solve(P, False, 200, 1e-6, 2, 10, 1, False)
        it_d = S.last["iters"]
        P.get_device(ca.CH_X, xd.p.value, st.value)
        assert h.hipStreamSynchronize(st) == 0
        x_dev = xd.take(np.empty((P.n, 512), np.float32))
    finally:
        bd.free()
        xd.free()
        h.hipStreamDestroy(st)
    assert it_d == it_h
    # the solve fills x with 0 on LEAF tiles only (GMGSolver.cu:537); the
    # other tiles keep the -3 set before it
    leaf_c = typ[perm] == po.LEAF
    assert np.array_equal(x_dev[leaf_c], x_host[leaf_c])
    assert np.all(x_dev[~leaf_c] == -3.0)
    ro = o.solve(max_iters=200, tol=1e-6, level_iters=2, coarsest_iters=10)
    assert it_d == ro["iters"]
    x_ora = o.get(po.CH_X).astype(np.float32)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    leaf = typ == po.LEAF
    assert np.array_equal(x_dev[inv][leaf], x_ora[leaf])


def test_device_moves_null_stream_and_velocity_channels():
    o, P, perm, b, typ = shuffled_problem(seed=1)
    u = np.random.default_rng(5).standard_normal((P.n, 512)).astype(np.float32)
    h = hip()
    ud, back = DevBuf(h, u.nbytes), DevBuf(h, u.nbytes)
    try:
        ud.put(u)
        P.set_device(ca.CH_V, ud.p.value, None)  # allocates the channel on first use
        P.get_device(ca.CH_V, back.p.value, None)
        got = back.take(np.empty_like(u))  # hipMemcpy on the null stream follows the move
    finally:
        ud.free()
        back.free()
    assert np.array_equal(got, u)
    assert np.array_equal(P.get(ca.CH_V), u)


def test_device_moves_reject_bad_buffers():
    o, P, perm, b, typ = shuffled_problem(seed=2)
    host = np.zeros((P.n, 512), np.float32)
    with pytest.raises(ca.CirrusError, match="not memory known to HIP|not device-accessible"):
        P.set_device(ca.CH_B, host.ctypes.data, None)
    h = hip()
    one = DevBuf(h, 512 * 4)  # one tile only
    try:
        with pytest.raises(ca.CirrusError, match="smaller than 512 floats per tile"):
            P.get_device(ca.CH_X, one.p.value, None)
    finally:
        one.free()
    P.set(ca.CH_B, b[perm])
    it, rel = ca.GMGSolver().solve(P, False, 200, 1e-6, 2, 10, 1, False)
    assert rel < 1e-6


TORCH_CHILD = r"""
import sys
import numpy as np
import torch  # first: the library then binds to torch's HIP runtime
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + "/oracle")
import cirrus_amd as ca
import pyoracle as po
maps = open("/proc/self/maps").read()
assert len({l.split()[-1] for l in maps.splitlines() if "libamdhip64" in l}) == 1, "two HIP runtimes"
o = po.Oracle(kind=po.SCENE_PLANE, plane_level=[3, 2])
lev, crd, typ, _ = o.topology()
P = ca.Problem(ca.Tree(o.params.h0, lev, crd, typ, o.celltypes()))
b = o.get(po.CH_B).astype(np.float32)
side = torch.cuda.Stream()
with torch.cuda.stream(side):
    bt = torch.from_numpy(b).cuda()
    xt = torch.empty_like(bt)
    P.set_device(ca.CH_B, bt.data_ptr(), side.cuda_stream)
it, rel = ca.GMGSolver().solve(P, False, 200, 1e-6, 2, 10, 1, False)
with torch.cuda.stream(side):
    P.get_device(ca.CH_X, xt.data_ptr(), side.cuda_stream)
    x = xt.cpu().numpy()
ro = o.solve(max_iters=200, tol=1e-6, level_iters=2, coarsest_iters=10)
leaf = typ == po.LEAF
assert it == ro["iters"] and np.array_equal(x[leaf], o.get(po.CH_X).astype(np.float32)[leaf])
print("torch interop ok", it)
"""


@pytest.mark.timeout(600)
def test_device_moves_with_torch_tensors():
    pytest.importorskip("torch")
    r = subprocess.run([sys.executable, "-c", TORCH_CHILD, ROOT], capture_output=True, text=True, timeout=560)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "torch interop ok" in r.stdout
