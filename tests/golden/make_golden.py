"""Writes the golden fixtures of tests/golden/ from the reference's own
CPUJoin/GridIndexF.cpp (built into oracle/_ref by `make -C oracle ref`).

Fixtures hold inputs and the reference's outputs (candidate lists and
per-polygon counts of GridIndexF::getRegion as CPUJoin::execute counts them):
  toy_squares.npz    the 5-point / 2-square sample of RasterJoin/test/baseline.hpp:139-217
  squares_edges.npz  axis-aligned squares with points on vertices and edges
  voronoi20.npz      20 Voronoi cells (shared jittered edges), 6000 points, grid 64
  index_1M_voronoi260.npz        counts for the 1M-point / 260-polygon / grid-256 workload,
                     with checksums of the regenerated inputs (inputs are not stored)
Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "raster-join_amd"))

from oracle import pyoracle as O  # noqa: E402
import rasterjoin as rj  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def ref_outputs(xy, off, pts, res):
    tmp = tempfile.mkdtemp()
    g = O.RefGrid(xy, off, res, tmp)
    cell_off, cand = g.export()
    counts = g.join_count(pts)
    return cell_off, cand, counts


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def toy():
    # baseline.hpp:139-217: two squares, 2 points at (5.15, 5.15), 3 at (5, 5)
    xy = np.array([5.1, 5.1, 5.2, 5.1, 5.2, 5.2, 5.1, 5.2,
                   4.9, 4.9, 5.1, 4.9, 5.1, 5.1, 4.9, 5.1], np.float32)
    off = np.array([0, 4, 8], np.int32)
    pts = np.array([5.15, 5.15, 5.15, 5.15, 5, 5, 5, 5, 5, 5], np.float32)
    cell_off, cand, counts = ref_outputs(xy, off, pts, 4)
    np.savez_compressed(os.path.join(OUT, "toy_squares.npz"), xy=xy, off=off, pts=pts, res=4,
                        cell_off=cell_off, cand=cand, counts=counts)


def squares_edges():
    # 3x3 unit squares tiling [0,3]^2 and a diamond; points on vertices, on
    # edges, at centers and on the outer max edges.
    polys = []
    for j in range(3):
        for i in range(3):
            polys.append([i, j, i + 1, j, i + 1, j + 1, i, j + 1])
    polys.append([1.5, 0.25, 2.75, 1.5, 1.5, 2.75, 0.25, 1.5])
    xy = np.array([v for p in polys for v in p], np.float32)
    off = np.array([0] + list(np.cumsum([len(p) // 2 for p in polys])), np.int32)
    g = np.arange(0.0, 3.01, 0.25)
    px, py = np.meshgrid(g, g)
    pts = np.stack([px.ravel(), py.ravel()], 1).astype(np.float32).ravel()
    cell_off, cand, counts = ref_outputs(xy, off, pts, 8)
    np.savez_compressed(os.path.join(OUT, "squares_edges.npz"), xy=xy, off=off, pts=pts, res=8,
                        cell_off=cell_off, cand=cand, counts=counts)


def voronoi20():
    xy, off = rj.gen_voronoi(20, seed=11, spacing=(500.0, 500.0), jitter=30.0)
    pts, _, _, _ = rj.gen_points(6000, params=rj.point_params(seed=21), columns=False)
    cell_off, cand, counts = ref_outputs(xy, off, pts, 64)
    np.savez_compressed(os.path.join(OUT, "voronoi20.npz"), xy=xy, off=off, pts=pts, res=64,
                        cell_off=cell_off, cand=cand, counts=counts)


def index_1M_voronoi260():
    cp = rj.POLY_SETS["voronoi260"]
    xy, off = rj.gen_voronoi(cp["nsites"], seed=cp["seed"], spacing=cp["spacing"], jitter=cp["jitter"])
    pts, _, _, _ = rj.gen_points(1_000_000, params=rj.point_params(seed=7), columns=False)
    cell_off, cand, counts = ref_outputs(xy, off, pts, 256)
    np.savez_compressed(os.path.join(OUT, "index_1M_voronoi260.npz"), res=256, counts=counts,
                        ncand=len(cand), cand_sha=sha(cell_off, cand),
                        poly_sha=sha(xy, off), pts_sha=sha(pts), npts=1_000_000, point_seed=7)


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` (needs the reference tree)")
    toy()
    squares_edges()
    voronoi20()
    index_1M_voronoi260()
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)), "bytes")
