"""Attribute filters (QueryConstraint, GLFunction.cpp:250-323) on every
column type and comparison, with the values where typed comparisons differ
from bit comparisons: NaN, signed zeros, infinities, denormals, the integer
extremes. Index-join counts (exact membership, so the filter alone decides)
against the oracle's CPUJoin restatement, which compares as
index.glsl:184-208 does; single queries and one fused batch."""
import numpy as np
import pytest

from oracle import pyoracle as O
import rasterjoin as rj
from common import points, polys

pytestmark = pytest.mark.gpu
N = 200_000
TYPES = (rj.UINT, rj.FLOAT, rj.INT, rj.UINT8)
F32_SPECIAL = [np.nan, -np.nan, np.inf, -np.inf, 0.0, -0.0, 1e-45, -1e-45, 1.1754944e-38,
               -1.1754944e-38, 3.4028235e38, -3.4028235e38, 1.0, -1.0, 2.5, 7.25]
U32_SPECIAL = [0, 1, 2, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFE, 0xFFFFFFFF, 1234567]
I32_SPECIAL = [0, 1, -1, -2147483648, -2147483647, 2147483647, 2147483646, -77]
U8_SPECIAL = [0, 1, 2, 127, 128, 254, 255, 256, 1000]


def columns(seed):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, 1 << 32, N, dtype=np.uint64).astype(np.uint32)
    pick = rng.random(N) < 0.3
    u[pick] = rng.choice(np.array(U32_SPECIAL, np.uint32), pick.sum())
    f = rng.normal(0, 3, N).astype(np.float32)
    pick = rng.random(N) < 0.4
    f[pick] = rng.choice(np.array(F32_SPECIAL, np.float32), pick.sum())
    i = rng.integers(-(1 << 31), 1 << 31, N, dtype=np.int64).astype(np.int32)
    pick = rng.random(N) < 0.3
    i[pick] = rng.choice(np.array(I32_SPECIAL, np.int32), pick.sum())
    b = rng.integers(0, 256, N).astype(np.uint8)
    return u, f, i, b


def constraint_sets():
    ops = (rj.EQ, rj.LT, rj.LTE, rj.GT, rj.GTE)
    out = []
    for op in ops:
        out += [[(0, op, v)] for v in U32_SPECIAL]
        out += [[(1, op, float(v))] for v in F32_SPECIAL]
        out += [[(2, op, v)] for v in I32_SPECIAL]
        out += [[(3, op, v)] for v in U8_SPECIAL]
    # several constraints on one column (a range, an empty range, contradictory
    # equalities) and on several columns
    out += [
        [(1, rj.GT, -1.0), (1, rj.LTE, 2.5)],
        [(1, rj.GTE, 0.0), (1, rj.LTE, -0.0)],
        [(1, rj.GT, 0.0), (1, rj.LT, 1e-45)],
        [(1, rj.GTE, -np.inf), (1, rj.LTE, np.inf)],
        [(2, rj.GT, -2147483648), (2, rj.LT, 2147483647), (2, rj.EQ, -77)],
        [(2, rj.EQ, 1), (2, rj.EQ, -1)],
        [(0, rj.GTE, 0x80000000), (0, rj.LT, 0xFFFFFFFF)],
        [(3, rj.GTE, 2), (3, rj.LTE, 127), (0, rj.GT, 1), (1, rj.LT, 7.25), (2, rj.GTE, -1)],
        [(3, rj.GT, 255)],
        [(0, rj.LT, 0)],
    ]
    return out


@pytest.fixture(scope="module")
def data():
    pts = points(N, 21)[0]
    return pts, columns(5)


def oracle_counts(g, pts, cols, cons):
    return O.join_exact(g, O.PointSet(pts, cols, TYPES, cons))[0]


def test_filters_single_queries(handler, data):
    pts, cols = data
    xy, off = polys("voronoi20")
    g = O.Grid(xy, off, 64)
    h = handler
    h.set_polygons(xy, off)
    h.set_points(pts, cols, TYPES)
    h.set_aggregation(rj.COUNT, -1)
    h.set_poly_index_resolution(64)
    for path in (0, 1):
        h.set_option(rj.OPT_INDEX_PATH, path)
        for cons in constraint_sets():
            h.set_query_constraints(cons)
            c, _ = h.execute_function(rj.INDEX)
            assert np.array_equal(c, oracle_counts(g, pts, cols, cons)), (path, cons)
    h.set_option(rj.OPT_INDEX_PATH, 0)


def test_filters_fused_batch(handler, data):
    pts, cols = data
    xy, off = polys("voronoi20")
    g = O.Grid(xy, off, 64)
    h = handler
    table = rj.Table(h, pts, cols, TYPES)
    ps = rj.PolySet(h, xy, off)
    sets = constraint_sets()[::3]
    qs = [rj.make_query(ps, rj.INDEX, rj.COUNT, -1, cons, index_res=64) for cons in sets]
    for k in range(0, len(qs), 48):
        res = h.execute_batch(table, qs[k:k + 48], [ps.np] * len(qs[k:k + 48]))
        for cons, r in zip(sets[k:k + 48], res):
            assert r.status == 0
            assert np.array_equal(r.counts, oracle_counts(g, pts, cols, cons)), cons


def test_float_sums_of_special_values(handler, data):
    """avg over the float column that holds NaN, infinities, signed zeros,
    denormals and tiny negatives: the exact fixed-point sums (hi, lo) equal
    the oracle's on both index paths, the hybrid join and a fused batch.
    A float in [-2^-25, 0) counts as 0; non-finite values count nothing."""
    pts, cols = data
    xy, off = polys("voronoi20")
    g = O.Grid(xy, off, 64)
    h = handler
    h.set_polygons(xy, off)
    h.set_points(pts, cols, TYPES)
    h.set_aggregation(rj.AVG, 1)
    h.set_poly_index_resolution(64)
    h.set_accuracy_distance(20.0)
    for cons in ([], [(1, rj.LTE, 0.0)], [(1, rj.GT, -1.0), (1, rj.LT, 1.0)]):
        c, hi, lo, _ = O.join_exact(g, O.PointSet(pts, cols, TYPES, cons, agg_type=1, agg_col=1))
        h.set_query_constraints(cons)
        for fn, path in ((rj.INDEX, 0), (rj.INDEX, 1), (rj.HYBRID, 0)):
            h.set_option(rj.OPT_INDEX_PATH, path)
            got, _ = h.execute_function(fn)
            ghi, glo = h.sums_fixed()
            assert np.array_equal(got, c), (fn, path, cons)
            assert np.array_equal(ghi, hi) and np.array_equal(glo, lo), (fn, path, cons)
    h.set_option(rj.OPT_INDEX_PATH, 0)
    table = rj.Table(h, pts, cols, TYPES)
    ps = rj.PolySet(h, xy, off)
    qs = [rj.make_query(ps, fn, rj.AVG, 1, [], eps=20.0, index_res=64) for fn in (rj.INDEX, rj.HYBRID)]
    c, hi, lo, _ = O.join_exact(g, O.PointSet(pts, cols, TYPES, [], agg_type=1, agg_col=1))
    for r in h.execute_batch(table, qs, [ps.np] * len(qs)):
        assert r.status == 0
        assert np.array_equal(r.counts, c)
        assert np.array_equal(r.hi, hi) and np.array_equal(r.lo, lo)


@pytest.mark.parametrize("col", [0, 2, 3])
def test_integer_sums_of_every_type(handler, data, col):
    """sum over the Uint, Int and u8 columns (the integer extremes among the
    values): the exact sums equal the oracle's on both index paths, the hybrid
    join, the raster join and a fused batch of all four, with and without a
    filter on the summed column. (u8 sums were garbage before: the device
    compiler dropped the integer branch of the aggregate decode for type 4.)"""
    pts, cols = data
    xy, off = polys("voronoi20")
    g = O.Grid(xy, off, 64)
    p = O.plan(O.poly_bbox(xy), 20.0)
    h = handler
    h.set_polygons(xy, off)
    h.set_points(pts, cols, TYPES)
    h.set_aggregation(rj.SUM, col)
    h.set_poly_index_resolution(64)
    h.set_accuracy_distance(20.0)
    filt = {0: (0, rj.LT, 0x80000000), 2: (2, rj.GTE, -77), 3: (3, rj.LTE, 200)}[col]
    table = rj.Table(h, pts, cols, TYPES)
    ps = rj.PolySet(h, xy, off)
    for cons in ([], [filt]):
        pset = O.PointSet(pts, cols, TYPES, cons, agg_type=rj.SUM, agg_col=col)
        c, hi, lo, _ = O.join_exact(g, pset)
        rc, rhi, rlo = O.raster_join(p, xy, off, pset)
        h.set_query_constraints(cons)
        for fn, path in ((rj.INDEX, 0), (rj.INDEX, 1), (rj.HYBRID, 0), (rj.RASTER, 0)):
            h.set_option(rj.OPT_INDEX_PATH, path)
            got, _ = h.execute_function(fn)
            ghi, glo = h.sums_fixed()
            want = (rc, rhi, rlo) if fn == rj.RASTER else (c, hi, lo)
            assert np.array_equal(got, want[0]), (fn, path, cons)
            assert np.array_equal(ghi, want[1]) and np.array_equal(glo, want[2]), (fn, path, cons)
        h.set_option(rj.OPT_INDEX_PATH, 0)
        fns = (rj.INDEX, rj.HYBRID, rj.RASTER, rj.RASTER)
        qs = [rj.make_query(ps, fn, rj.SUM, col, cons, eps=20.0, index_res=64) for fn in fns]
        for fn, r in zip(fns, h.execute_batch(table, qs, [ps.np] * len(qs))):
            want = (rc, rhi, rlo) if fn == rj.RASTER else (c, hi, lo)
            assert r.status == 0
            assert np.array_equal(r.counts, want[0]), (fn, cons)
            assert np.array_equal(r.hi, want[1]) and np.array_equal(r.lo, want[2]), (fn, cons)
        # the same queries on a host (streamed) table and through the entry
        # point without allocations
        host = rj.Table(h, pts, cols, TYPES, host=True)
        for fn, q in zip(fns[:3], qs[:3]):
            want = (rc, rhi, rlo) if fn == rj.RASTER else (c, hi, lo)
            alone = h.execute_query(table, q, ps.np)  # also prepares the structures
            for r in (alone, h.execute_query(host, q, ps.np), h.execute_query_noalloc(table, q, ps.np)):
                assert r.status == 0
                assert np.array_equal(r.counts, want[0]), (fn, cons)
                assert np.array_equal(r.hi, want[1]) and np.array_equal(r.lo, want[2]), (fn, cons)
        host.close()


def test_float_sums_of_a_column_in_range(handler, data):
    """A float column whose every value is finite and in [-2^31, 2^31) takes
    the decode without range tests. Its edge values: tiny negatives that
    count as 0 (-1e-45, -2^-26, -2^-25), signed zeros, denormals, -2^31
    itself and the largest float below 2^31, fractions. Sums (hi, lo) equal
    the oracle's on both index paths, the hybrid and raster joins and a fused
    batch, and equal the same column's sums with one NaN added (the decode
    with range tests)."""
    pts, cols = data
    rng = np.random.default_rng(9)
    edge = np.array([-1e-45, -2.0 ** -26, -2.0 ** -25, -2.0 ** -24, 0.0, -0.0, 1e-45, 1.1754944e-38,
                     -1.1754944e-38, -2147483648.0, 2147483520.0, -2147483520.0, 0.5, -0.5, 7.25,
                     -7.75, 1e9 + 0.5, -1e9], np.float32)
    f = rng.normal(0, 1e4, N).astype(np.float32)
    pick = rng.random(N) < 0.4
    f[pick] = rng.choice(edge, pick.sum())
    with_nan = f.copy()
    with_nan[int(np.flatnonzero(f == 0.0)[0])] = np.nan  # a row that adds 0 either way
    xy, off = polys("voronoi20")
    g = O.Grid(xy, off, 64)
    p = O.plan(O.poly_bbox(xy), 20.0)
    h = handler
    ps = rj.PolySet(h, xy, off)
    types = (rj.FLOAT, rj.UINT8)
    got = {}
    for name, col in (("in_range", f), ("nan", with_nan)):
        tcols = (col, cols[3])
        pset = O.PointSet(pts, tcols, types, [], agg_type=rj.AVG, agg_col=0)
        c, hi, lo, _ = O.join_exact(g, pset)
        rc, rhi, rlo = O.raster_join(p, xy, off, pset)
        h.set_polygons(xy, off)
        h.set_points(pts, tcols, types)
        h.set_aggregation(rj.AVG, 0)
        h.set_poly_index_resolution(64)
        h.set_accuracy_distance(20.0)
        h.set_query_constraints([])
        for fn, path in ((rj.INDEX, 0), (rj.INDEX, 1), (rj.HYBRID, 0), (rj.RASTER, 0)):
            h.set_option(rj.OPT_INDEX_PATH, path)
            cnt, _ = h.execute_function(fn)
            ghi, glo = h.sums_fixed()
            want = (rc, rhi, rlo) if fn == rj.RASTER else (c, hi, lo)
            assert np.array_equal(cnt, want[0]), (name, fn, path)
            assert np.array_equal(ghi, want[1]) and np.array_equal(glo, want[2]), (name, fn, path)
            got[(name, fn, path)] = (ghi, glo)
        h.set_option(rj.OPT_INDEX_PATH, 0)
        table = rj.Table(h, pts, tcols, types)
        fns = (rj.INDEX, rj.HYBRID, rj.RASTER)
        qs = [rj.make_query(ps, fn, rj.AVG, 0, [(1, rj.LTE, 200)] if k else [], eps=20.0, index_res=64)
              for fn in fns for k in (0, 1)]
        for q, r in zip(qs, h.execute_batch(table, qs, [ps.np] * len(qs))):
            assert r.status == 0
            assert np.array_equal(r.hi, h.execute_query(table, q, ps.np).hi)
        table.close()
    # the NaN replaced a zero and adds nothing: both decodes give the same sums
    for fn, path in ((rj.INDEX, 0), (rj.HYBRID, 0), (rj.RASTER, 0)):
        a, b = got[("in_range", fn, path)], got[("nan", fn, path)]
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (fn, path)
