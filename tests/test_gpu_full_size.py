"""The joins at the workloads' full sizes, checked through properties that
hold at any size (the oracle pins the same kernels at smaller sizes in the
other GPU tests):

* 100M points, raster eps = 10 m, count + avg(fare), 260 polygons: the
  label-map and canvas paths agree bit for bit; results are additive over
  row ranges; the exact count lies inside the error bounds of every polygon.
* 500M points, hybrid eps = 20 m, 2 filters, 2200 tract polygons, count +
  avg(fare): the hybrid join equals the index join's per-point grid test (two
  independent CPUJoin implementations); additive over row ranges; filtered
  counts never exceed unfiltered ones.
* 1B points in day buckets: fused batches equal the queries run alone.
"""
import gc

import numpy as np
import pytest

import rasterjoin as rj
from common import COLS_TYPES, PASS_LE2, T0, WINDOW_30, polys

pytestmark = pytest.mark.gpu


def gen(n, seed):
    # not cached: these tables are large
    return rj.gen_points(n, params=rj.point_params(seed=seed))


def fixed(h, fn):
    c, _ = h.execute_function(fn)
    hi, lo = h.sums_fixed()
    return c, hi, lo


def value(r):
    """Per polygon: the count and the exact sum hi + lo * 2^-32 as the integer
    hi * 2^32 + lo (Python ints, no rounding)."""
    return [(int(c), int(h) * (1 << 32) + int(l)) for c, h, l in zip(*r)]


def additive(parts, full):
    vp = [value(p) for p in parts]
    tot = [(sum(v[i][0] for v in vp), sum(v[i][1] for v in vp)) for i in range(len(full[0]))]
    return tot == value(full)


def test_raster_100M_eps10_paths_additivity_bounds():
    n = 100_000_000
    pts, t, fare, pas = gen(n, 7)
    xy, off = polys("voronoi260")
    np_ = len(off) - 1
    h = rj.Handler()
    table = rj.Table(h, pts, (t, fare, pas), COLS_TYPES, organize=False)
    del t, pas
    h.set_polygons(xy, off)
    h.bind_table(table)
    h.set_query_constraints(())
    h.set_aggregation(rj.AVG, 1)
    h.set_accuracy_distance(10.0)
    out = {}
    for path in (2, 1):  # label map, then canvas
        h.set_option(rj.OPT_RASTER_PATH, path)
        out[path] = fixed(h, rj.RASTER)
    h.set_option(rj.OPT_RASTER_PATH, 0)
    for a, b in zip(out[1], out[2]):
        assert np.array_equal(a, b)
    full = out[2]
    assert full[0].sum() <= n and full[0].sum() > 0.99 * n
    # additive over row ranges (an odd split point)
    parts = []
    for rb, re in ((0, 37_777_777), (37_777_777, n)):
        h.bind_table(table, rb, re)
        parts.append(fixed(h, rj.RASTER))
    assert additive(parts, full)
    # error bounds contain the exact (hybrid) count; errorbounds counts = raster counts
    h.bind_table(table)
    h.set_aggregation(rj.COUNT, -1)
    c, _ = h.execute_function(rj.ERRORBOUNDS)
    b = h.get_error_bounds()
    assert np.array_equal(c, full[0])
    exact, _ = h.execute_function(rj.HYBRID)
    assert np.all(c - b[3 * np_:] <= exact) and np.all(exact <= c + b[2 * np_:3 * np_])
    h.close()
    del pts, fare, table
    gc.collect()


def test_hybrid_500M_filters_equals_grid_index():
    n = 500_000_000
    pts, t, fare, pas = gen(n, 11)
    xy, off = polys("tracts2200")
    h = rj.Handler()
    table = rj.Table(h, pts, (t, fare, pas), COLS_TYPES, organize=False)
    del pts, t, fare, pas
    gc.collect()
    h.set_polygons(xy, off)
    h.bind_table(table)
    cons = WINDOW_30 + PASS_LE2
    h.set_query_constraints(cons)
    h.set_aggregation(rj.AVG, 1)
    h.set_accuracy_distance(20.0)
    h.set_poly_index_resolution(1024)
    hyb = fixed(h, rj.HYBRID)
    h.set_option(rj.OPT_INDEX_PATH, 1)  # per-point test over the grid cell's candidates
    idx = fixed(h, rj.INDEX)
    h.set_option(rj.OPT_INDEX_PATH, 0)
    for a, b in zip(hyb, idx):
        assert np.array_equal(a, b)
    # about 10 % of the rows pass the two filters
    assert 0.05 * n < hyb[0].sum() < 0.15 * n
    parts = []
    for rb, re in ((0, 123_456_789), (123_456_789, n)):
        h.bind_table(table, rb, re)
        parts.append(fixed(h, rj.HYBRID))
    assert additive(parts, hyb)
    h.bind_table(table)
    h.set_query_constraints(WINDOW_30)
    h.set_aggregation(rj.COUNT, -1)
    c1, _ = h.execute_function(rj.HYBRID)
    assert np.all(hyb[0] <= c1)
    h.close()
    del table
    gc.collect()


def test_batch_1B_equals_single():
    n = 1_000_000_000
    pts, t, fare, pas = gen(n, 7)
    h = rj.Handler()
    table = rj.Table(h, pts, (t, fare, pas), COLS_TYPES, time_col=0, time0=T0, bucket_seconds=86400)
    del pts, t, fare, pas
    gc.collect()
    sets = [rj.PolySet(h, *polys(k)) for k in ("voronoi260", "tracts400", "voronoi20", "tracts2200")]
    rng = np.random.default_rng(77)
    qs, nps = [], []
    for i in range(16):
        ps = sets[i % len(sets)]
        length = int(np.exp(rng.uniform(np.log(3600), np.log(31536000))))
        lo = T0 + int(rng.uniform(0, 31536000 - length))
        cons = [(0, rj.GTE, lo), (0, rj.LTE, lo + length)]
        fn = (rj.RASTER, rj.HYBRID, rj.INDEX)[i % 3]
        agg = rj.AVG if i % 2 else rj.COUNT
        qs.append(rj.make_query(ps, fn, agg, 1 if agg == rj.AVG else -1, cons,
                                eps=float((5, 10, 50, 100)[i % 4]), index_res=256))
        nps.append(ps.np)
    batch = h.execute_batch(table, qs, nps)
    for q, k, b in zip(qs, nps, batch):
        a = h.execute_query(table, q, k)
        assert b.status == 0
        assert np.array_equal(a.counts, b.counts) and np.array_equal(a.hi, b.hi)
        assert np.array_equal(a.lo, b.lo)
    assert sum(b.path == rj.PATH_FUSED for b in batch) >= 12
    h.close()
    del table
    gc.collect()
