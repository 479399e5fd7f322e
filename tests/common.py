"""Shared inputs of the parity tests (seeded, regenerated on any machine)."""
import functools

import numpy as np

import rasterjoin as rj


@functools.lru_cache(maxsize=None)
def polys(name):
    if name == "voronoi260":
        cp = rj.POLY_SETS["voronoi260"]
        return rj.gen_voronoi(cp["nsites"], seed=cp["seed"], spacing=cp["spacing"], jitter=cp["jitter"])
    if name == "tracts2200":
        cp = rj.POLY_SETS["tracts2200"]
        return rj.gen_voronoi(cp["nsites"], seed=cp["seed"], spacing=cp["spacing"], jitter=cp["jitter"])
    if name == "tracts400":
        return rj.gen_voronoi(400, seed=17, spacing=(2.2, 140.0), jitter=30.0)
    if name == "voronoi20":
        return rj.gen_voronoi(20, seed=11, spacing=(500.0, 500.0), jitter=30.0)
    if name == "stars":
        return star_polygons(40, seed=3, box=50000.0, rmin=500.0, rmax=9000.0, vmin=3, vmax=80)
    raise KeyError(name)


@functools.lru_cache(maxsize=None)
def points(n, seed=7):
    return rj.gen_points(n, params=rj.point_params(seed=seed))


def star_polygons(n, seed, box=1000.0, rmin=20.0, rmax=200.0, vmin=3, vmax=40):
    """Random star-shaped polygons; they overlap each other."""
    rng = np.random.default_rng(seed)
    xy, off = [], [0]
    for _ in range(n):
        k = int(rng.integers(vmin, vmax + 1))
        cx, cy = rng.uniform(rmax, box - rmax, 2)
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        r = rng.uniform(rmin, rmax, k)
        pts = np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], 1).astype(np.float32)
        keep = [0] + [i for i in range(1, k) if not np.array_equal(pts[i], pts[i - 1])]
        pts = pts[keep]
        xy.append(pts.ravel())
        off.append(off[-1] + len(pts))
    return np.concatenate(xy), np.array(off, np.int32)


COLS_TYPES = [rj.UINT, rj.FLOAT, rj.UINT8]   # time, fare, passengers
T0 = 1230768000
WINDOW_30 = [(0, rj.GTE, T0 + 5_000_000), (0, rj.LTE, T0 + 5_000_000 + int(0.3 * 31536000))]
PASS_LE2 = [(2, rj.LTE, 2)]
