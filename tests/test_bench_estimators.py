"""bench.py on the CPU: its count of the reference's edge tests
(cell_edge_work: the vertices of GridIndexF's candidate lists, cell by cell)
against the oracle's grid (the CPUJoin restatement), and its one-GPU
contract."""
import os
import sys

import numpy as np

from oracle import pyoracle as O
from common import points, polys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def exact_cell_work(xy, off, res):
    cell_off, cand = O.Grid(xy, off, res).export()
    nv = np.diff(off).astype(np.int64)
    per = np.zeros(res * res, np.int64)
    for c in range(res * res):
        per[c] = nv[cand[cell_off[c]:cell_off[c + 1]]].sum()
    return per.reshape(res, res)


def test_cell_edge_work_matches_the_grid_lists():
    for name, res in (("voronoi20", 32), ("voronoi260", 64), ("voronoi260", 256)):
        xy, off = polys(name)
        est, _ = bench.cell_edge_work(xy, off, res)
        ref = exact_cell_work(xy, off, res)
        # equal but for cells a polygon only touches at a float boundary
        assert (est != ref).mean() < 0.005, name
        assert abs(est.sum() / ref.sum() - 1) < 0.005, name


def test_edge_tests_counts_points_in_cells():
    xy, off = polys("voronoi20")
    cw = bench.cell_edge_work(xy, off, 32)
    work = cw[0]
    pts = points(20_000, 5)[0]
    x, y = pts[0::2], pts[1::2]
    minx, miny, w, h = cw[1]
    tot = 0
    for k in range(len(x)):
        cx, cy = np.floor((x[k] - minx) / w), np.floor((y[k] - miny) / h)
        if 0 <= cx < 32 and 0 <= cy < 32:
            tot += int(work[int(cy), int(cx)])
    assert bench.edge_tests(cw, pts) == tot
    # points outside the grid, NaN and infinities test nothing
    odd = np.array([np.nan, 1.0, -1e9, 5.0, np.inf, 3.0, 1e9, 1e9], np.float32)
    assert bench.edge_tests(cw, odd) == 0


def test_bench_measures_one_gpu_only():
    """bench.py --gpus N>1 exits with a message instead of a figure for
    several GPUs (this build has no multi-GPU path)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
