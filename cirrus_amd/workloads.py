"""The synthetic workloads of the benchmark configurations, as scene keyword
arguments of :class:`cirrus_amd.Scene` (the oracle's scene builder takes the
same keywords).  One definition serves the benchmark and the tests.

* ``ADAPTIVE256`` -- adaptive octree, 256^3 effective: leaf tiles at h=1/256
  where the tile box meets |d|<0.05 of a sphere of radius 0.15 at the domain
  centre, h=1/128 where it meets |d|<0.15, h=1/64 elsewhere, 2:1 face
  balanced; the sphere is solid (NEUMANN), the top face Dirichlet, the other
  walls Neumann; the right-hand side is the divergence of a seeded N(0,1)
  face-velocity field.
* ``PLAIN64`` -- uniform 64^3 (level 3), solid sphere r=0.2, same walls.
* ``VORTEX512`` -- two root tiles ([0,1]x[0,1]x[0,2]), leaf level 6 (h=1/512)
  where the vorticity of 8 seeded Gaussian tubes (radius U[0.02, 0.06]) exceeds
  0.5 over the tile, level 4 elsewhere: 512x512x1024 effective.
  ``VORTEX256`` is the same field one level coarser.
* ``batch_scene_params(q, seed)`` -- problem q of the batched workload:
  adaptive problems around a solid sphere (centre U[0.3,0.7]^3, radius
  U[0.1,0.25]) with a refinement band, leaf-tile counts drawn from
  LogNormal(ln 2048, 1) clipped to [64, 16384] and met approximately by the
  band levels and width; every domain face Neumann or Dirichlet at random, at
  least one Dirichlet.
"""
import numpy as np

ADAPTIVE256 = dict(kind=1, solid_r=0.15, band_r=0.15, band_w=[0.05, 0.15], band_level=[5, 4, 3])
PLAIN64 = dict(kind=0, uniform_level=3, solid_r=0.2)
VORTEX512 = dict(kind=3, nroot=[1, 1, 2], vortex_level=[6, 4], vortex_thresh=0.5, seed=0)
VORTEX256 = dict(kind=3, nroot=[1, 1, 2], vortex_level=[5, 3], vortex_thresh=0.5, seed=0)

# Band configurations (finest level, band half-width) and their leaf-tile
# counts at radius 0.175 (counts scale roughly with the sphere's area).
BAND_TABLE = [(2, 0.0, 64), (3, 0.0, 512), (3, 0.1, 288), (4, 0.005, 680), (4, 0.1, 1128), (5, 0.005, 1968),
              (5, 0.02, 2528), (5, 0.05, 3760), (5, 0.1, 7400), (6, 0.005, 11096), (6, 0.02, 14736)]


def batch_scene_params(q, seed=0):
    """(scene keyword arguments, drawn leaf-tile target) of problem q of the
    batched workload with the given seed."""
    rng = np.random.default_rng([seed, q])
    target = float(np.clip(rng.lognormal(np.log(2048.0), 1.0), 64, 16384))
    r = float(rng.uniform(0.1, 0.25))
    centre = list(rng.uniform(0.3, 0.7, 3))
    bc = [4 if rng.random() < 0.5 else 1 for _ in range(6)]  # NEUMANN = 4, DIRICHLET = 1
    if 1 not in bc:
        bc[int(rng.integers(6))] = 1
    scale = (r / 0.175) ** 2
    best = min(BAND_TABLE, key=lambda e: abs(np.log(e[2] * (scale if e[1] > 0 else 1.0)) - np.log(target)))
    lf, w, _ = best
    kw = dict(solid_c=centre, solid_r=r, face_bc=bc, seed=1000 + q)
    if w == 0.0:
        kw.update(kind=0, uniform_level=lf)
    else:
        kw.update(kind=1, band_c=centre, band_r=r, band_w=[w, 3 * w], band_level=[lf, lf - 1, max(lf - 2, 2)])
    return kw, target


def build_scenes(ca, kws, workers=16):
    """Build the scenes of the given keyword sets on the host, in parallel (the
    library's builder is re-entrant and releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max(1, min(workers, len(kws)))) as ex:
        return list(ex.map(lambda kw: ca.Scene(**kw), kws))
