"""An independent finite-volume statement of the pressure operator, for pinning
the oracle.

It works on the leaf partition of the octree (global cell coordinates per
level), not on tiles, ghosts or halos: for every face of an INTERIOR leaf cell
it finds what lies across (a same-level leaf, a coarser leaf, four finer leaves,
or the outside of the domain) and adds the volume-weighted flux

    same level or outside:   h   * (x_c - x_n)          (outside: x_n = 0)
    coarse/fine face:        (2/3) h_f * (x_c - x_n)     per fine face, h_f the fine size

(the flux area h_f^2 over the centre distance 1.5 h_f), skipping faces where
either cell is NEUMANN.  Vectorised with numpy; used on small trees.
"""
import numpy as np
import scipy.sparse as sp

LEAF, NONLEAF, GHOST = 1, 2, 4
DIRICHLET, INTERIOR, NEUMANN = 1, 2, 4


def _key(l, g):
    g = np.asarray(g, np.int64)
    return (np.asarray(l, np.int64) << 60) | (g[..., 0] << 40) | (g[..., 1] << 20) | g[..., 2]


class LeafCells:
    def __init__(self, level, coord, ttype, ctype, h0, nroot):
        leaf_tiles = np.nonzero(ttype == LEAF)[0]
        ci = np.arange(512)
        loc = np.stack([ci >> 6, (ci >> 3) & 7, ci & 7], axis=1)
        self.tile = np.repeat(leaf_tiles, 512)
        self.cell = np.tile(ci, len(leaf_tiles))
        self.level = level[self.tile].astype(np.int64)
        self.g = coord[self.tile].astype(np.int64) * 8 + loc[self.cell]
        self.ct = ctype[self.tile, self.cell]
        self.h = h0 / (2.0 ** self.level)
        self.nroot = np.asarray(nroot, np.int64)
        keys = _key(self.level, self.g)
        self.order = np.argsort(keys)
        self.skeys = keys[self.order]
        self.n = len(keys)

    def lookup(self, l, g):
        """Index of the leaf cell (l, g), or -1."""
        k = _key(l, np.maximum(g, 0))
        pos = np.searchsorted(self.skeys, k)
        pos = np.minimum(pos, self.n - 1)
        ok = (self.skeys[pos] == k) & np.all(g >= 0, axis=-1)
        return np.where(ok, self.order[pos], -1)

    def outside(self, l, g):
        ext = self.nroot * 8 * (np.int64(1) << l)[..., None]
        return np.any((g < 0) | (g >= ext), axis=-1)

    def faces(self):
        """Yield (rows, cols, coeff) of every non-NEUMANN face of an INTERIOR
        leaf cell; cols = -1 means the outside of the domain (value 0)."""
        rows_i = np.nonzero(self.ct == INTERIOR)[0]
        l = self.level[rows_i]
        g = self.g[rows_i]
        h = self.h[rows_i]
        for a in range(3):
            for s in (-1, 1):
                n = g.copy()
                n[:, a] += s
                same = self.lookup(l, n)
                coarse = np.where(same < 0, self.lookup(l - 1, n >> 1), -1)
                out = (same < 0) & (coarse < 0) & self.outside(l, n)
                # same level
                m = same >= 0
                yield rows_i[m], same[m], h[m]
                # outside: Dirichlet 0 with the same-level coefficient
                yield rows_i[out], np.full(out.sum(), -1), h[out]
                # coarser neighbour
                m = coarse >= 0
                yield rows_i[m], coarse[m], (2.0 / 3.0) * h[m]
                # finer neighbours (4 cells across the face)
                fin = (same < 0) & (coarse < 0) & ~out
                if fin.any():
                    rf, lf, nf, hf = rows_i[fin], l[fin], n[fin], h[fin] / 2
                    for oj in (0, 1):
                        for ok in (0, 1):
                            c = 2 * nf
                            c[:, a] += 1 if s < 0 else 0
                            c[:, (a + 1) % 3] += oj
                            c[:, (a + 2) % 3] += ok
                            f = self.lookup(lf + 1, c)
                            assert (f >= 0).all(), "neighbour region neither same, coarser, finer nor outside"
                            yield rf, f, (2.0 / 3.0) * hf

    def matrix(self):
        """Sparse A over all leaf cells (rows of non-INTERIOR cells empty)."""
        R, C, V = [], [], []
        diag = np.zeros(self.n)
        for rows, cols, coef in self.faces():
            nn = (self.ct[rows] & NEUMANN) == 0
            inside = cols >= 0
            nn &= np.where(inside, (self.ct[np.maximum(cols, 0)] & NEUMANN) == 0, True)
            rows, cols, coef = rows[nn], cols[nn], coef[nn]
            np.add.at(diag, rows, coef)
            m = cols >= 0
            R.append(rows[m]); C.append(cols[m]); V.append(-coef[m])
        R.append(np.arange(self.n)); C.append(np.arange(self.n)); V.append(diag)
        return sp.csr_matrix((np.concatenate(V), (np.concatenate(R), np.concatenate(C))), shape=(self.n, self.n))

    def to_tiles(self, v, ntiles):
        out = np.zeros((ntiles, 512))
        out[self.tile, self.cell] = v
        return out

    def from_tiles(self, a):
        return a[self.tile, self.cell]
