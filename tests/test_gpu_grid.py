"""The CPUJoin grid index built on the device (K4: GridIndexF::buildGrid,
CPUJoin/GridIndexF.cpp:60-155; the reference's GPU index is
GLFunction::createPolyIndex, GLFunction.cpp:65-156): its cell offsets and
candidate lists equal, byte for byte, the lists the reference's own
GridIndexF.cpp produced (golden fixtures) and the oracle's restatement on the
bench's collections, up to the 2200 tracts at grid 1024."""
import hashlib
import os

import numpy as np
import pytest

from oracle import pyoracle as O
import rasterjoin as rj
from common import polys

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("name", ["toy_squares", "squares_edges", "voronoi20"])
def test_device_grid_equals_reference_lists(handler, name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    ps = rj.PolySet(handler, d["xy"], d["off"])
    cell_off, cand, _ = ps.grid(int(d["res"]))
    assert np.array_equal(cell_off, d["cell_off"])
    assert np.array_equal(cand, d["cand"])


def test_device_grid_1M_voronoi260_reference_hash(handler):
    d = np.load(os.path.join(GOLD, "index_1M_voronoi260.npz"))
    xy, off = polys("voronoi260")
    assert sha(xy, off) == str(d["poly_sha"])
    ps = rj.PolySet(handler, xy, off)
    cell_off, cand, _ = ps.grid(int(d["res"]))
    assert len(cand) == int(d["ncand"])
    assert sha(cell_off, cand) == str(d["cand_sha"])


@pytest.mark.parametrize("name,res", [("tracts2200", 1024), ("tracts2200", 256), ("voronoi260", 1024),
                                      ("stars", 64), ("voronoi5000", 512), ("voronoi10", 2048)])
def test_device_grid_equals_oracle(handler, name, res):
    xy, off = rj.poly_set(name) if name in rj.POLY_SETS else polys(name)
    ps = rj.PolySet(handler, xy, off)
    cell_off, cand, defi = ps.grid(res)
    ocell, ocand = O.Grid(xy, off, res).export()
    assert np.array_equal(cell_off, ocell)
    assert np.array_equal(cand, ocand)
    # deficient cells: sorted, distinct, inside the grid, and each lies in the
    # bounding box of some polygon missing from its list
    assert np.all(np.diff(defi) > 0) and (len(defi) == 0 or (defi.min() >= 0 and defi.max() < res * res))
    v = xy.reshape(-1, 2)
    bx = np.array([[v[off[p]:off[p + 1], 0].min(), v[off[p]:off[p + 1], 1].min(),
                    v[off[p]:off[p + 1], 0].max(), v[off[p]:off[p + 1], 1].max()] for p in range(len(off) - 1)])
    minx, miny = v[:, 0].min(), v[:, 1].min()
    w, h = (v[:, 0].max() - minx) / res, (v[:, 1].max() - miny) / res
    for c in defi[:200]:
        x, y = c % res, c // res
        lx, ly, hx, hy = minx + x * w, miny + y * h, minx + (x + 1) * w, miny + (y + 1) * h
        near = np.nonzero((bx[:, 0] <= hx + w) & (bx[:, 2] >= lx - w) & (bx[:, 1] <= hy + h) & (bx[:, 3] >= ly - h))[0]
        assert len(set(near.tolist()) - set(cand[cell_off[c]:cell_off[c + 1]].tolist())) > 0, c


@pytest.mark.parametrize("name,res,shift", [("tracts2200", 1024, 0.0), ("voronoi260", 2048, 0.0), ("stars", 64, 0.0),
                                            ("voronoi10", 1024, 0.0), ("voronoi260", 256, 1.0e6),
                                            ("voronoi20", 1024, 3.0e7)])
def test_device_grid_marks_equal_per_cell_test(handler, name, res, shift):
    """The grid build's per-vertex and per-edge marks (default) give the same
    candidate lists and deficient cells as testing every cell of a polygon's
    box against all its edges (RJ_OPT_GRID_DIRECT), also with coordinates far
    from the origin: at 3e7 the float grid is too fine for the marks and the
    items fall back to the per-cell test."""
    xy, off = rj.poly_set(name) if name in rj.POLY_SETS else polys(name)
    xy = (xy.astype(np.float64) + shift).astype(np.float32)
    out = []
    for direct in (0, 1):
        handler.set_option(rj.OPT_GRID_DIRECT, direct)
        ps = rj.PolySet(handler, xy, off)
        out.append(ps.grid(res))
    handler.set_option(rj.OPT_GRID_DIRECT, 0)
    (c0, k0, d0), (c1, k1, d1) = out
    assert np.array_equal(c0, c1) and np.array_equal(k0, k1) and np.array_equal(d0, d1)
    if shift == 0.0:
        ocell, ocand = O.Grid(xy, off, res).export()
        assert np.array_equal(c0, ocell) and np.array_equal(k0, ocand)
