"""The two-stage point pass of label and hybrid queries with constraints:
each lane tests its rows' constraints and queues the rows that pass, and the
queue goes through the label maps a wave's worth at a time. Every constraint
set of the filter tests (each column type and comparison, NaN, signed zeros,
infinities, integer extremes, empty and contradictory ranges, no row or
every row passing) through the hybrid join (exact: CPUJoin's counts and
fixed-point sums) and the label-map raster join (the oracle's raster
emulator); row ranges that start and end inside a workgroup's rows, on a
table whose size is not a multiple of a row group; time windows on a table
ordered by time buckets, whose whole buckets run without the filter."""
import numpy as np
import pytest

from oracle import pyoracle as O
import rasterjoin as rj
from common import COLS_TYPES, PASS_LE2, WINDOW_30, points, polys
from test_gpu_filters import TYPES, columns, constraint_sets

pytestmark = pytest.mark.gpu
N = 200_000


@pytest.fixture(scope="module")
def data():
    return points(N, 21)[0], columns(5)


def test_hybrid_every_constraint_set(handler, data):
    pts, cols = data
    xy, off = polys("voronoi20")
    g = O.Grid(xy, off, 64)
    h = handler
    h.set_polygons(xy, off)
    h.set_points(pts, cols, TYPES)
    h.set_poly_index_resolution(64)
    h.set_accuracy_distance(20.0)
    for k, cons in enumerate(constraint_sets()):
        # avg over the float column on every third set (its sums take the
        # aggregate's value through the queue)
        agg, col = (rj.AVG, 1) if k % 3 == 0 else (rj.COUNT, -1)
        h.set_query_constraints(cons)
        h.set_aggregation(agg, col)
        c, hi, lo, _ = O.join_exact(g, O.PointSet(pts, cols, TYPES, cons, agg_type=agg,
                                                  agg_col=max(col, 0)))
        got, _ = h.execute_function(rj.HYBRID)
        assert np.array_equal(got, c), cons
        if agg == rj.AVG:
            ghi, glo = h.sums_fixed()
            assert np.array_equal(ghi, hi) and np.array_equal(glo, lo), cons


def test_label_raster_every_constraint_set(handler, data):
    pts, cols = data
    xy, off = polys("voronoi20")
    eps = 20.0
    p = O.plan(O.poly_bbox(xy).astype(np.float64), eps)
    h = handler
    h.set_polygons(xy, off)
    h.set_points(pts, cols, TYPES)
    h.set_accuracy_distance(eps)
    h.set_option(rj.OPT_RASTER_PATH, 2)
    try:
        for k, cons in enumerate(constraint_sets()[::2]):
            agg, col = (rj.AVG, 1) if k % 2 == 0 else (rj.COUNT, -1)
            h.set_query_constraints(cons)
            h.set_aggregation(agg, col)
            oc, ohi, olo = O.raster_join(p, xy, off, O.PointSet(pts, cols, TYPES, cons, agg_type=agg,
                                                                  agg_col=max(col, 0)))
            got, _ = h.execute_function(rj.RASTER)
            assert np.array_equal(got, oc), cons
            if agg == rj.AVG:
                ghi, glo = h.sums_fixed()
                assert np.array_equal(ghi, ohi) and np.array_equal(glo, olo), cons
    finally:
        h.set_option(rj.OPT_RASTER_PATH, 0)


@pytest.mark.parametrize("rows", [(1234, 777_777), (5, 6), (0, 1_000_003), (999_000, 1_000_003),
                                  (2047, 4097), (300_001, 300_001)])
def test_row_ranges_inside_row_groups(handler, rows):
    """Queries over rows [a, b) of a table in input order: the first and last
    row groups of the pass are partial; equal to the oracle on the slice."""
    n = 1_000_003
    pts, t, fare, pas = points(n, 33)
    xy, off = polys("voronoi260")
    h = handler
    table = rj.Table(h, pts, (t, fare, pas), COLS_TYPES, organize=False)
    ps = rj.PolySet(h, xy, off)
    a, b = rows
    cons = PASS_LE2 + WINDOW_30
    sl = O.PointSet(pts[2 * a:2 * b], (t[a:b], fare[a:b], pas[a:b]), COLS_TYPES, cons, agg_type=rj.AVG, agg_col=1)
    c, hi, lo, _ = O.join_exact(O.Grid(xy, off, 256), sl)
    q = rj.make_query(ps, rj.HYBRID, rj.AVG, 1, cons, eps=20.0, index_res=256, rows=rows)
    r = h.execute_query(table, q, ps.np)
    assert np.array_equal(r.counts, c)
    assert np.array_equal(r.hi, hi) and np.array_equal(r.lo, lo)
    eps = 10.0
    p = O.plan(O.poly_bbox(xy).astype(np.float64), eps)
    oc, ohi, olo = O.raster_join(p, xy, off, sl)
    h.set_option(rj.OPT_RASTER_PATH, 2)
    try:
        q = rj.make_query(ps, rj.RASTER, rj.AVG, 1, cons, eps=eps, rows=rows)
        r = h.execute_query(table, q, ps.np)
    finally:
        h.set_option(rj.OPT_RASTER_PATH, 0)
    assert np.array_equal(r.counts, oc)
    assert np.array_equal(r.hi, ohi) and np.array_equal(r.lo, olo)
    table.close()


DAY = 86400
T0B = 1230768000


def bucket_windows():
    d = DAY
    return [
        [(0, rj.GTE, T0B + 5 * d), (0, rj.LTE, T0B + 9 * d - 1)],       # whole buckets
        [(0, rj.GTE, T0B + 5 * d + 1), (0, rj.LTE, T0B + 9 * d)],       # one second off each end
        [(0, rj.GT, T0B + 5 * d - 1), (0, rj.LT, T0B + 9 * d)],
        [(0, rj.LTE, T0B + 2 * d - 1)],                                  # from time 0 (bucket 0 holds t < t0)
        [(0, rj.GTE, T0B - 100)],                                        # to the end (the last bucket)
        [(0, rj.GTE, 0)],
        [(0, rj.EQ, T0B + 7 * d)],
        [(0, rj.LT, T0B)],
        [(0, rj.GTE, T0B + 30 * d), (0, rj.LTE, 0xFFFFFFFF)],
        [(0, rj.GTE, T0B + 3 * d), (0, rj.LTE, T0B + 20 * d - 1), (2, rj.LTE, 3)],  # and another column
        [(0, rj.GT, 0xFFFFFFFF)],
        [],
    ]


def test_time_bucket_windows(handler):
    """A table ordered by day buckets (times below the first bucket's start
    and on bucket edges included): rows in whole buckets inside a window run
    without the filter; single queries of three joins and one fused batch
    equal the oracle on every window."""
    n = 300_001
    pts, _, fare, pas = points(n, 41)
    rng = np.random.default_rng(9)
    t = (T0B + rng.integers(-3 * DAY, 40 * DAY, n)).astype(np.uint32)
    edge = rng.random(n) < 0.05
    t[edge] = (T0B + rng.integers(-2, 40, edge.sum()) * DAY - rng.integers(0, 2, edge.sum())).astype(np.uint32)
    cols = (t, fare, pas)
    xy, off = polys("voronoi260")
    h = handler
    table = rj.Table(h, pts, cols, COLS_TYPES, time_col=0, time0=T0B, bucket_seconds=DAY)
    ps = rj.PolySet(h, xy, off)
    g = O.Grid(xy, off, 256)
    p = O.plan(O.poly_bbox(xy).astype(np.float64), 10.0)
    h.set_option(rj.OPT_RASTER_PATH, 2)
    try:
        qs, want = [], []
        for k, cons in enumerate(bucket_windows()):
            agg, col = (rj.AVG, 1) if k % 2 == 0 else (rj.COUNT, -1)
            ops = O.PointSet(pts, cols, COLS_TYPES, cons, agg_type=agg, agg_col=max(col, 0))
            exact = O.join_exact(g, ops)[:3]
            raster = O.raster_join(p, xy, off, ops)
            for fn, ref in ((rj.RASTER, raster), (rj.HYBRID, exact), (rj.INDEX, exact)):
                q = rj.make_query(ps, fn, agg, col, cons, eps=10.0, index_res=256)
                r = h.execute_query(table, q, ps.np)
                assert np.array_equal(r.counts, ref[0]), (fn, cons)
                assert np.array_equal(r.hi, ref[1]) and np.array_equal(r.lo, ref[2]), (fn, cons)
                qs.append(q)
                want.append(ref)
        for r, ref in zip(h.execute_batch(table, qs, [ps.np] * len(qs)), want):
            assert r.status == 0
            assert np.array_equal(r.counts, ref[0])
            assert np.array_equal(r.hi, ref[1]) and np.array_equal(r.lo, ref[2])
    finally:
        h.set_option(rj.OPT_RASTER_PATH, 0)
    table.close()
