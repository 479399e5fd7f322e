"""Parity of the four HIP joins with the oracle: bit-exact counts and
fixed-point sums (index and hybrid against the CPUJoin restatement and the
reference's own GridIndexF fixtures; raster and errorbounds against the CPU
emulation of the same pixel-center canvas rule)."""
import os

import numpy as np
import pytest

from oracle import pyoracle as O
import rasterjoin as rj
from common import COLS_TYPES, PASS_LE2, WINDOW_30, points, polys

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def oracle_exact(xy, off, pts, cols=(), cons=(), agg=rj.COUNT, agg_col=0, res=256):
    g = O.Grid(xy, off, res)
    types = COLS_TYPES[:len(cols)]
    return O.join_exact(g, O.PointSet(pts, cols, types, cons, agg_type=agg, agg_col=agg_col))


@pytest.fixture(params=[0, 1], ids=["index-via-labels", "index-grid-pip"])
def ih(request, handler):
    """The handler with each index-join path: the label-accelerated exact
    classification (default) and the per-point grid PIP of IndexJoin."""
    handler.set_option(rj.OPT_INDEX_PATH, request.param)
    return handler


def run(h, fn, xy, off, pts, cols=(), cons=(), agg=rj.COUNT, agg_col=0, eps=None, res=None):
    h.set_polygons(xy, off)
    h.set_points(pts, cols, COLS_TYPES[:len(cols)])
    h.set_query_constraints(cons)
    h.set_aggregation(agg, agg_col if agg != rj.COUNT else -1)
    if eps is not None:
        h.set_accuracy_distance(eps)
    if res is not None:
        h.set_poly_index_resolution(res)
    counts, sums = h.execute_function(fn)
    hi, lo = h.sums_fixed()
    return counts, sums, hi, lo


# ---------------------------------------------------------------------------
# index join (K4 grid + K5 PIP), CPUJoin semantics
# ---------------------------------------------------------------------------

def test_index_toy(ih):
    d = np.load(os.path.join(GOLD, "toy_squares.npz"))
    c, _, _, _ = run(ih, rj.INDEX, d["xy"], d["off"], d["pts"], res=4)
    assert c.tolist() == [2, 3]


@pytest.mark.parametrize("name", ["squares_edges", "voronoi20"])
def test_index_golden_reference_fixture(ih, name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    c, _, _, _ = run(ih, rj.INDEX, d["xy"], d["off"], d["pts"], res=int(d["res"]))
    assert np.array_equal(c, d["counts"])


def test_index_1M_voronoi260_matches_reference_counts(ih):
    """1M points, 260 polygons, grid 256 (the CPUJoin-comparable workload)."""
    d = np.load(os.path.join(GOLD, "index_1M_voronoi260.npz"))
    xy, off = polys("voronoi260")
    pts = points(1_000_000, 7)[0]
    c, _, _, _ = run(ih, rj.INDEX, xy, off, pts, res=256)
    assert np.array_equal(c, d["counts"])


def test_index_filters_and_avg_exact(ih):
    xy, off = polys("voronoi260")
    pts, t, fare, pas = points(600_000, 8)
    cons = WINDOW_30 + PASS_LE2
    c, s, hi, lo = run(ih, rj.INDEX, xy, off, pts, (t, fare, pas), cons, rj.AVG, 1, res=256)
    oc, ohi, olo, of64 = oracle_exact(xy, off, pts, (t, fare, pas), cons, rj.AVG, 1, res=256)
    assert np.array_equal(c, oc)
    assert np.array_equal(hi, ohi) and np.array_equal(lo, olo)
    assert np.allclose(s, of64, rtol=1e-9, atol=0)


def test_index_tracts2200_grid1024_filters_and_avg(ih):
    """The 2200-tract collection (30-2000 jittered vertices per polygon) at
    grid 1024 with a 30 % time window, passengers <= 2 and avg(fare): counts
    and exact sums equal the oracle's CPUJoin grid (GridIndexF.cpp:60-155)."""
    xy, off = polys("tracts2200")
    pts, t, fare, pas = points(1_000_000, 31)
    cons = WINDOW_30 + PASS_LE2
    c, s, hi, lo = run(ih, rj.INDEX, xy, off, pts, (t, fare, pas), cons, rj.AVG, 1, eps=20.0, res=1024)
    oc, ohi, olo, of64 = oracle_exact(xy, off, pts, (t, fare, pas), cons, rj.AVG, 1, res=1024)
    assert np.array_equal(c, oc)
    assert np.array_equal(hi, ohi) and np.array_equal(lo, olo)
    assert np.allclose(s, of64, rtol=1e-9, atol=0)
    assert c.sum() > 20_000


def test_index_overlapping_polygons(ih):
    xy, off = polys("stars")
    pts = points(400_000, 9)[0]
    c, _, _, _ = run(ih, rj.INDEX, xy, off, pts, res=64)
    oc, _, _, _ = oracle_exact(xy, off, pts, res=64)
    assert np.array_equal(c, oc)
    assert c.sum() > 400_000 // 4


def test_index_edge_cases(ih):
    xy, off = polys("voronoi20")
    # empty input
    c, _, _, _ = run(ih, rj.INDEX, xy, off, np.zeros(0, np.float32), res=32)
    assert c.sum() == 0
    # outside the bbox, NaN, infinities, max edges
    pts = np.array([-1, 5, 5, -1, 50000, 10, 10, 50000, np.nan, 3, 3, np.nan, np.inf, 4,
                    -np.inf, 4, 0, 0, 49999.996, 49999.996], np.float32)
    c, _, _, _ = run(ih, rj.INDEX, xy, off, pts, res=32)
    oc, _, _, _ = oracle_exact(xy, off, pts, res=32)
    assert np.array_equal(c, oc)


# ---------------------------------------------------------------------------
# raster join (K1 splat + K2 scan, or label map + gather)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name,eps,window", [("voronoi260", 10.0, False), ("voronoi260", 50.0, False),
                                             ("tracts2200", 20.0, True)])
def test_raster_paths_match_emulator(handler, name, eps, window):
    xy, off = polys(name)
    pts, t, fare, pas = points(2_000_000, 10)
    cons = [(2, rj.LTE, 4)] + (WINDOW_30 if window else [])
    p = O.plan(O.poly_bbox(xy).astype(np.float64), eps)
    oc, ohi, olo = O.raster_join(p, xy, off, O.PointSet(pts, (t, fare, pas), COLS_TYPES, cons,
                                                          agg_type=rj.AVG, agg_col=1))
    for path in (1, 2, 0):
        handler.set_option(rj.OPT_RASTER_PATH, path)
        c, s, hi, lo = run(handler, rj.RASTER, xy, off, pts, (t, fare, pas), cons, rj.AVG, 1, eps=eps)
        assert np.array_equal(c, oc), path
        assert np.array_equal(hi, ohi) and np.array_equal(lo, olo), path
    handler.set_option(rj.OPT_RASTER_PATH, 0)


def test_raster_tiled_canvas_eps5(handler):
    xy, off = polys("voronoi260")
    pts = points(1_000_000, 11)[0]
    p = O.plan(O.poly_bbox(xy).astype(np.float64), 5.0)
    assert p.splitx == 2
    oc, _, _ = O.raster_join(p, xy, off, O.PointSet(pts))
    c, _, _, _ = run(handler, rj.RASTER, xy, off, pts, eps=5.0)
    assert handler.canvas_size() == (p.W, p.H)
    assert np.array_equal(c, oc)


def test_raster_overlapping_polygons_use_the_canvas(handler):
    xy, off = polys("stars")
    pts = points(500_000, 12)[0]
    p = O.plan(O.poly_bbox(xy).astype(np.float64), 40.0)
    oc, _, _ = O.raster_join(p, xy, off, O.PointSet(pts))
    c, _, _, _ = run(handler, rj.RASTER, xy, off, pts, eps=40.0)
    assert np.array_equal(c, oc)
    handler.set_option(rj.OPT_RASTER_PATH, 2)
    with pytest.raises(rj.RJError):
        handler.execute_function(rj.RASTER)
    handler.set_option(rj.OPT_RASTER_PATH, 0)


def test_raster_pixel_aligned_squares_are_exact(handler):
    bbox = np.array([0.0, 0.0, 1000.0, 1000.0])
    squares = [np.array([i0, j0, i1, j0, i1, j1, i0, j1], np.float32) * 10
               for (i0, j0, i1, j1) in [(0, 0, 10, 10), (10, 0, 35, 10), (5, 20, 45, 70), (50, 50, 100, 100)]]
    xy = np.concatenate(squares)
    off = np.arange(0, 17, 4, dtype=np.int32)
    rng = np.random.default_rng(1)
    pts = rng.uniform(0, 1000, (300_000, 2)).astype(np.float32)
    pts = pts[~np.any(np.abs(pts / 10 - np.round(pts / 10)) < 1e-3, axis=1)].ravel()
    handler.set_polygons(xy, off, bbox)
    handler.set_points(pts)
    handler.set_query_constraints([])
    handler.set_aggregation(rj.COUNT)
    handler.set_accuracy_distance(10.0 * np.sqrt(2))
    c, _ = handler.execute_function(rj.RASTER)
    exact, _, _, _ = O.join_exact(O.Grid(xy, off, 16, bbox=[0, 0, 1000, 1000]), O.PointSet(pts))
    assert np.array_equal(c, exact)


# ---------------------------------------------------------------------------
# hybrid join (K6 mask + K7 point pass with exact border test)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name,eps,res", [("voronoi260", 20.0, 256), ("voronoi260", 5.0, 1024),
                                          ("tracts400", 20.0, 256), ("stars", 100.0, 64),
                                          ("tracts2200", 20.0, 1024)])
def test_hybrid_equals_cpujoin(handler, name, eps, res):
    xy, off = polys(name)
    pts, t, fare, pas = points(1_000_000, 14)
    cons = WINDOW_30 + PASS_LE2
    c, s, hi, lo = run(handler, rj.HYBRID, xy, off, pts, (t, fare, pas), cons, rj.AVG, 1, eps=eps, res=res)
    oc, ohi, olo, of64 = oracle_exact(xy, off, pts, (t, fare, pas), cons, rj.AVG, 1, res=res)
    assert np.array_equal(c, oc)
    assert np.array_equal(hi, ohi) and np.array_equal(lo, olo)
    assert np.allclose(s, of64, rtol=1e-9, atol=0)


@pytest.mark.parametrize("name", ["voronoi260", "tracts400", "stars"])
def test_hybrid_and_index_points_near_edges(ih, name):
    """Points on edges, on vertices and at 1e-4 m .. 7 m from edges (the
    float-error zone the hybrid border must send to the exact test). Every
    point carries its own id as the aggregate, so the per-polygon id sums
    catch any single misassigned point."""
    xy, off = polys(name)
    rng = np.random.default_rng(21)
    v = xy.reshape(-1, 2).astype(np.float64)
    prev = np.arange(len(v)) - 1
    prev[off[:-1]] = off[1:] - 1
    e = rng.integers(0, len(v), 200_000)
    a, b = v[prev[e]], v[e]
    mid = a + rng.uniform(0, 1, len(e))[:, None] * (b - a)
    nrm = np.stack([a[:, 1] - b[:, 1], b[:, 0] - a[:, 0]], 1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True) + 1e-30
    d = np.array([0, 1e-4, 1e-3, 1e-2, 0.05, 0.2, 1, 4, 7])
    offs = rng.choice(np.concatenate([d, -d]), len(e))
    p = np.concatenate([mid + nrm * offs[:, None], v[rng.integers(0, len(v), 20_000)]])
    pts = p.astype(np.float32).ravel()
    ids = np.arange(len(p), dtype=np.uint32)
    oc, ohi, _, _ = oracle_exact(xy, off, pts, (ids,), (), rj.AVG, 0, res=256)
    for fn in (rj.HYBRID, rj.INDEX):
        c, _, hi, _ = run(ih, fn, xy, off, pts, (ids,), (), rj.AVG, 0, eps=10.0, res=256)
        assert np.array_equal(c, oc), fn
        assert np.array_equal(hi, ohi), fn


def test_hybrid_count_1M_voronoi260_reference(handler):
    d = np.load(os.path.join(GOLD, "index_1M_voronoi260.npz"))
    xy, off = polys("voronoi260")
    pts = points(1_000_000, 7)[0]
    c, _, _, _ = run(handler, rj.HYBRID, xy, off, pts, eps=20.0, res=256)
    assert np.array_equal(c, d["counts"])


# ---------------------------------------------------------------------------
# errorbounds (K1 count splat + K3 boundary bands)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name,eps", [("voronoi260", 20.0), ("voronoi260", 100.0), ("tracts400", 10.0),
                                      ("stars", 40.0)])
def test_errorbounds_match_emulator_and_contain_exact(handler, name, eps):
    """Bounds at several accuracies and collections, overlapping polygons
    (stars) included: equal to the CPU emulation, and the exact count of
    every polygon lies within them."""
    xy, off = polys(name)
    pts, t, fare, pas = points(1_000_000, 15)
    cons = PASS_LE2
    ps = O.PointSet(pts, (t, fare, pas), COLS_TYPES, cons)
    p = O.plan(O.poly_bbox(xy).astype(np.float64), eps)
    oc, _, _, canvas = O.raster_join(p, xy, off, ps, want_canvas=True)
    ob = O.errorbounds(p, xy, off, canvas)
    c, _, _, _ = run(handler, rj.ERRORBOUNDS, xy, off, pts, (t, fare, pas), cons, eps=eps)
    b = handler.get_error_bounds(rj.ERRORBOUNDS)
    assert np.array_equal(c, oc)
    assert np.array_equal(b, ob)
    np_ = len(off) - 1
    exact, _, _, _ = oracle_exact(xy, off, pts, (t, fare, pas), cons)
    assert np.all(c - b[3 * np_:] <= exact) and np.all(exact <= c + b[2 * np_:3 * np_])


@pytest.mark.parametrize("W", [4000, 9000])
def test_errorbounds_wide_bands(handler, W):
    """Polygons spanning most of a wide canvas: error-bounds bands of three
    rows (W=4000) and of one row each (W=9000: all 256 threads of a band on
    one row's area prefix sums), equal to the CPU emulation."""
    from common import star_polygons
    xy, off = star_polygons(3, seed=77, box=1000.0, rmin=300, rmax=480, vmin=20, vmax=300)
    rng = np.random.default_rng(78)
    pts = rng.uniform(0, 1000, 2 * 2_000_000).astype(np.float32)
    bbox = O.poly_bbox(xy).astype(np.float64)
    eps = float(max(bbox[2] - bbox[0], bbox[3] - bbox[1]) / W * np.sqrt(2))
    p = O.plan(bbox, eps)
    assert W - 1 <= max(p.W, p.H) <= W + 1
    oc, _, _, canvas = O.raster_join(p, xy, off, O.PointSet(pts), want_canvas=True)
    ob = O.errorbounds(p, xy, off, canvas)
    c, _, _, _ = run(handler, rj.ERRORBOUNDS, xy, off, pts, eps=eps)
    b = handler.get_error_bounds(rj.ERRORBOUNDS)
    assert np.array_equal(c, oc)
    assert np.array_equal(b, ob)
    assert (ob[:2 * (len(off) - 1)] > 0).all()


def test_errorbounds_band_with_many_edges(handler):
    """A comb whose 800 teeth edges all cross the same few pixel rows: the
    error-bounds bands there hold more edges than one pass of a workgroup
    lists (255), so their (edge, row) items come in several rounds; equal to
    the CPU emulation."""
    teeth = 400
    xs = np.linspace(1000.0, 0.0, 2 * teeth + 1)
    top = np.stack([xs, 500.0 + 4.0 * (np.arange(2 * teeth + 1) % 2)], axis=1)
    comb = np.concatenate([np.array([[0.0, 0.0], [1000.0, 0.0]]), top]).astype(np.float32).ravel()
    square = np.array([200, 600, 800, 600, 800, 900, 200, 900], np.float32)
    xy = np.concatenate([comb, square])
    off = np.array([0, len(comb) // 2, len(comb) // 2 + 4], np.int32)
    rng = np.random.default_rng(79)
    pts = rng.uniform(0, 1000, 2 * 1_000_000).astype(np.float32)
    bbox = O.poly_bbox(xy).astype(np.float64)
    eps = float(max(bbox[2] - bbox[0], bbox[3] - bbox[1]) / 1000 * np.sqrt(2))
    p = O.plan(bbox, eps)
    oc, _, _, canvas = O.raster_join(p, xy, off, O.PointSet(pts), want_canvas=True)
    ob = O.errorbounds(p, xy, off, canvas)
    c, _, _, _ = run(handler, rj.ERRORBOUNDS, xy, off, pts, eps=eps)
    b = handler.get_error_bounds(rj.ERRORBOUNDS)
    assert np.array_equal(c, oc)
    assert np.array_equal(b, ob)
    assert ob[0] > 0 and ob[2] > 0  # the comb's teeth make boundary pixels on both sides


# ---------------------------------------------------------------------------
# API behaviour
# ---------------------------------------------------------------------------

def test_legacy_result_layout_and_time_stats(handler):
    xy, off = polys("voronoi20")
    pts, t, fare, pas = points(100_000, 16)
    run(handler, rj.RASTER, xy, off, pts, (t, fare, pas), (), rj.AVG, 1, eps=100.0)
    counts, sums, leg = handler.execute_function(rj.RASTER, legacy=True)
    np_ = len(off) - 1
    assert np.array_equal(leg[:np_], counts)
    approx = leg[np_:2 * np_] * 100.0 + leg[2 * np_:] / 10.0   # main.cpp:206-209 decode
    assert np.all(np.abs(approx - sums) <= 0.1 + 1e-9 * np.abs(sums))
    # GLHandler.cpp:228-240: N, np, fn, passes, nConstraints, 9 times, accuracy
    line = handler.print_time_stats(rj.RASTER).split()
    assert line[0] == "100000" and line[1] == str(np_) and line[2] == "0"
    assert len(line) == 15 and float(line[-1]) == 100.0


def test_errors_are_status_codes(handler):
    with pytest.raises(rj.RJError) as e:
        handler.execute_function(rj.RASTER)
    assert e.value.code == -2
    xy, off = polys("voronoi20")
    handler.set_polygons(xy, off)
    pts, t, fare, pas = points(1000, 1)
    handler.set_points(pts, (t, fare, pas), COLS_TYPES)
    with pytest.raises(rj.RJError) as e:
        handler.set_query_constraints([(0, rj.GT, 1)] * 11)
    assert e.value.code == -5
    handler.set_query_constraints([(7, rj.GT, 1)])
    with pytest.raises(rj.RJError) as e:
        handler.execute_function(rj.INDEX)
    assert e.value.code == -1
    with pytest.raises(rj.RJError) as e:
        handler.execute_function(9)
    assert e.value.code == -6


def test_memory_budget_rejects_oversized_polygons():
    h = rj.Handler(gpu_mem_mb=101)   # 1 MB after the reserve
    a = np.linspace(0, 2 * np.pi, 200_000, endpoint=False)
    xy = np.stack([1000 * np.cos(a), 1000 * np.sin(a)], 1).astype(np.float32).ravel()
    off = np.array([0, 200_000], np.int32)   # 1.6 MB of vertices
    with pytest.raises(rj.RJError) as e:
        h.set_polygons(xy, off)
    assert e.value.code == -3
    h.close()


def test_table_order_does_not_change_results(handler):
    xy, off = polys("voronoi260")
    pts, t, fare, pas = points(500_000, 18)
    res = []
    for org in (0, 1):
        handler.set_option(rj.OPT_ORGANIZE, org)
        res.append(run(handler, rj.HYBRID, xy, off, pts, (t, fare, pas), (), rj.AVG, 1, eps=20.0, res=256))
    handler.set_option(rj.OPT_ORGANIZE, 1)
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


def jittered_grid_polygons(nx, ny, box=50000.0, seed=4):
    """nx * ny quadrilaterals tiling the box: a grid whose interior nodes are
    jittered, each cell one polygon sharing its edges with its neighbours."""
    rng = np.random.default_rng(seed)
    gx = np.linspace(0.0, box, nx + 1)
    gy = np.linspace(0.0, box, ny + 1)
    X, Y = np.meshgrid(gx, gy, indexing="ij")
    jx, jy = 0.3 * box / nx, 0.3 * box / ny
    X[1:-1, 1:-1] += rng.uniform(-jx, jx, (nx - 1, ny - 1))
    Y[1:-1, 1:-1] += rng.uniform(-jy, jy, (nx - 1, ny - 1))
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    i, j = i.ravel(), j.ravel()
    cx = np.stack([X[i, j], X[i + 1, j], X[i + 1, j + 1], X[i, j + 1]], 1)
    cy = np.stack([Y[i, j], Y[i + 1, j], Y[i + 1, j + 1], Y[i, j + 1]], 1)
    xy = np.stack([cx, cy], 2).astype(np.float32).ravel()
    off = (np.arange(len(i) + 1) * 4).astype(np.int32)
    return xy, off


def test_many_polygons_32bit_label_maps(handler):
    """75,000 polygons: more than a 16-bit label map holds, so the raster and
    hybrid maps are 32-bit; raster against the pixel-center emulation,
    hybrid and index against CPUJoin."""
    xy, off = jittered_grid_polygons(300, 250)
    pts, t, fare, pas = points(500_000, 27)
    cols = (t, fare, pas)
    p = O.plan(O.poly_bbox(xy).astype(np.float64), 20.0)
    oc, ohi, olo = O.raster_join(p, xy, off, O.PointSet(pts, cols, COLS_TYPES, (), agg_type=rj.AVG, agg_col=1))
    c, _, hi, lo = run(handler, rj.RASTER, xy, off, pts, cols, (), rj.AVG, 1, eps=20.0)
    assert np.array_equal(c, oc) and np.array_equal(hi, ohi) and np.array_equal(lo, olo)
    ec, ehi, elo, _ = oracle_exact(xy, off, pts, cols, (), rj.AVG, 1, res=256)
    for fn in (rj.HYBRID, rj.INDEX):
        c, _, hi, lo = run(handler, fn, xy, off, pts, cols, (), rj.AVG, 1, eps=20.0, res=256)
        assert np.array_equal(c, ec), fn
        assert np.array_equal(hi, ehi) and np.array_equal(lo, elo), fn
    assert ec.sum() == len(pts) // 2  # the cells tile the box: every point is in one
