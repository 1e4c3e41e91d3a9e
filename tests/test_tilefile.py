"""Tile files on the CPU: the portable format round-trips a tree and its
channels exactly and rejects damaged files; the reader of the reference's raw
HAHostTileHolder dump (src/PoissonIOFunc.cu:38-124) takes the struct layout
that the C ABI gives the reference's members (src/PoissonTile.h:50-56,
src/HATileInfo.h:69-72), which ctypes computes independently here; the
committed fixture still reads and holds the oracle's pressure."""
import ctypes
import os
import struct

import numpy as np
import pytest

import cirrus_amd as ca
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "uniform16_solution.ctf")


def scene_channels(sc):
    return {1: sc.rhs, 6: sc.uvw[:, 0, :], 7: sc.uvw[:, 1, :], 8: sc.uvw[:, 2, :]}


def test_round_trip(tmp_path):
    sc = ca.Scene(kind=po.SCENE_SPHERE_BAND, solid_r=0.15, band_level=[4, 3, 2])
    path = str(tmp_path / "band.ctf")
    ca.save_tiles(path, sc.tree, scene_channels(sc))
    tree, ch = ca.load_tiles(path)
    assert tree.h0 == sc.tree.h0
    for a in ("level", "coord", "type", "cell_type"):
        assert np.array_equal(getattr(tree, a), getattr(sc.tree, a)), a
    assert sorted(ch) == [1, 6, 7, 8]
    for k, v in scene_channels(sc).items():
        assert np.array_equal(ch[k], v.astype(np.float32)), k
    # the header says what the file holds
    raw = open(path, "rb").read()
    assert raw[:8] == b"CIRTILES"
    ver, m, n, h0 = struct.unpack_from("<IIqd", raw, 8)
    assert (ver, m, n, h0) == (1, 4, sc.tree.n, sc.tree.h0)
    assert len(raw) == 64 + n * (4 + 12 + 1 + 512) + m * (8 + 512 * 4 * n)


def test_tree_without_channels(tmp_path):
    sc = ca.Scene(uniform_level=1)
    path = str(tmp_path / "t.ctf")
    ca.save_tiles(path, sc.tree)
    tree, ch = ca.load_tiles(path)
    assert ch == {} and np.array_equal(tree.cell_type, sc.tree.cell_type)


def test_damaged_files_are_rejected(tmp_path):
    sc = ca.Scene(uniform_level=1)
    path = str(tmp_path / "t.ctf")
    ca.save_tiles(path, sc.tree, {1: sc.rhs})
    raw = bytearray(open(path, "rb").read())
    cases = {}
    flipped = bytearray(raw)
    flipped[64 + 4 * sc.tree.n + 5] ^= 1  # a coordinate bit
    cases["checksum"] = flipped
    cases["size"] = raw[:-4]
    bad_magic = bytearray(raw)
    bad_magic[0:8] = b"NOTTILES"
    cases["magic"] = bad_magic
    bad_ver = bytearray(raw)
    bad_ver[8] = 9
    cases["version"] = bad_ver
    for name, data in cases.items():
        p = str(tmp_path / f"{name}.ctf")
        open(p, "wb").write(bytes(data))
        with pytest.raises(ca.CirrusError, match=name):
            ca.load_tiles(p)
    with pytest.raises(ca.CirrusError):
        ca.load_tiles(str(tmp_path / "missing.ctf"))


def test_invalid_types_are_not_written(tmp_path):
    sc = ca.Scene(uniform_level=1)
    bad = ca.Tree(sc.tree.h0, sc.tree.level, sc.tree.coord, sc.tree.type, sc.tree.cell_type.copy())
    bad.cell_type[3] = 3
    with pytest.raises(ca.CirrusError, match="cell type"):
        ca.save_tiles(str(tmp_path / "bad.ctf"), bad)
    with pytest.raises(ca.CirrusError, match="channel id"):
        ca.save_tiles(str(tmp_path / "bad2.ctf"), sc.tree, {11: sc.rhs})


# The reference's structs with the members in declaration order; ctypes lays
# them out by the platform C ABI, the layout the reference's host compiler
# gives these standard-layout types.
class _Coord(ctypes.Structure):
    _fields_ = [("mVec", ctypes.c_int32 * 3)]


class _TileInfo(ctypes.Structure):
    _fields_ = [("mTileCoord", _Coord), ("mTilePtr", ctypes.c_void_p), ("mLevel", ctypes.c_int),
                ("mType", ctypes.c_uint8)]


class _PoissonTile(ctypes.Structure):
    _fields_ = [("mData", (ctypes.c_float * 736) * 11), ("mNeighbors", _TileInfo * 6),
                ("mCellType", ctypes.c_uint8 * 512), ("mIsInterestArea", ctypes.c_bool),
                ("mIsLockedRefine", ctypes.c_bool), ("mStatus", ctypes.c_int), ("mSerialIdx", ctypes.c_int)]


def write_holder(path, h0, layers, max_level, tiles, levels):
    """WriteHAHostTileHolderToFile's byte sequence (src/PoissonIOFunc.cu:38-75):
    tiles = list of (cell types, {channel: 512 floats}); levels = per level a
    list of (coord, level, type, tile index or -1)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<fii", h0, layers, max_level))
        f.write(struct.pack("<Q", len(tiles)))
        for ctype, chans in tiles:
            t = _PoissonTile()
            for c, v in chans.items():
                t.mData[c][:512] = list(map(float, v))
            t.mCellType[:] = list(map(int, ctype))
            t.mStatus, t.mSerialIdx = 7, 9
            f.write(bytes(t))
        f.write(struct.pack("<Q", len(levels)))
        for lst in levels:
            f.write(struct.pack("<Q", len(lst)))
            for coord, lev, typ, idx in lst:
                f.write(struct.pack("<3iiBq", *coord, lev, typ, idx))


def test_reference_struct_layout():
    assert ctypes.sizeof(_TileInfo) == 32
    # 11*736*4 data + 6*32 neighbours + 512 cell types + 2 bools, 2 padding,
    # 2 ints, rounded up to the 8-byte alignment of the pointers
    assert ctypes.sizeof(_PoissonTile) == 33104
    assert _PoissonTile.mNeighbors.offset == 11 * 736 * 4
    assert _PoissonTile.mCellType.offset == 32576


def test_import_holder(tmp_path):
    sc = ca.Scene(kind=po.SCENE_PLANE, plane_level=[2, 1])
    tr = sc.tree
    rng = np.random.default_rng(0)
    x = rng.standard_normal((tr.n, 512)).astype(np.float32)
    chans = {0: x, 1: sc.rhs, 6: sc.uvw[:, 0, :], 7: sc.uvw[:, 1, :], 8: sc.uvw[:, 2, :]}
    # tiles stored in reverse order; the level lists point at them, plus one empty entry
    order = list(range(tr.n))[::-1]
    slot = {t: q for q, t in enumerate(order)}
    ct = tr.cell_type.reshape(tr.n, 512)
    tiles = [(ct[t], {c: v[t] for c, v in chans.items()}) for t in order]
    L = int(tr.level.max()) + 1
    levels = [[] for _ in range(L)]
    for t in range(tr.n):
        levels[tr.level[t]].append((tuple(tr.coord[t]), int(tr.level[t]), int(tr.type[t]), slot[t]))
    levels[0].append(((5, 5, 5), 0, 0, -1))
    path = str(tmp_path / "holder.bin")
    write_holder(path, tr.h0, 10, L - 1, tiles, levels)
    tree, ch = ca.import_holder(path, channels=(0, 1, 6, 7, 8))
    assert tree.n == tr.n and abs(tree.h0 - tr.h0) < 1e-7
    # tiles in level-list order, which here is the scene's own order
    for a in ("level", "coord", "type", "cell_type"):
        assert np.array_equal(getattr(tree, a), getattr(tr, a)), a
    for c, v in chans.items():
        assert np.array_equal(ch[c], np.asarray(v, np.float32)), c
    # truncated or padded files are refused
    raw = open(path, "rb").read()
    for name, data in (("short", raw[:-3]), ("long", raw + b"\0")):
        p = str(tmp_path / f"{name}.bin")
        open(p, "wb").write(data)
        with pytest.raises(ca.CirrusError):
            ca.import_holder(p)


def test_committed_fixture_holds_the_oracle_pressure():
    tree, ch = ca.load_tiles(FIXTURE)
    o = po.Oracle(uniform_level=1)
    lev, crd, typ, nbr = o.topology()
    assert np.array_equal(tree.level, lev) and np.array_equal(tree.type, typ)
    assert np.array_equal(tree.cell_type, o.celltypes())
    assert np.array_equal(ch[1], o.get(po.CH_B).astype(np.float32))
    o.solve(max_iters=200, tol=1e-6, level_iters=1, coarsest_iters=10)
    assert np.array_equal(ch[0], o.get(po.CH_X).astype(np.float32))
