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
* 1B points in day buckets (x, y, time, fare, passengers):
  - 64 fused raster + hybrid queries after the autotuner (polygon sets of
    10-2200, eps in {5, 10, 50, 100}, windows selecting 1M-1B rows) equal
    the queries run alone;
  - the serving loop under Poisson arrivals (60/30/10 raster/hybrid/index,
    10-5000 polygons, windows 1 hour - 1 year, passengers filters) returns
    every query's alone result;
  - in both, queries restricted to a few million rows of the 1B table (rows
    straddling a window's start, deep in the table) run in the same fused
    passes or served mix and equal the oracle on those rows.
"""
import gc
import threading
import time

import numpy as np
import pytest

from oracle import pyoracle as O
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
    # error bounds contain the exact (hybrid) count; errorbounds counts = raster
    # counts; the label-map and canvas forms of errorbounds agree bit for bit
    h.bind_table(table)
    h.set_aggregation(rj.COUNT, -1)
    h.set_option(rj.OPT_RASTER_PATH, 1)
    c1, _ = h.execute_function(rj.ERRORBOUNDS)
    b1 = h.get_error_bounds()
    h.set_option(rj.OPT_RASTER_PATH, 0)
    c, _ = h.execute_function(rj.ERRORBOUNDS)
    b = h.get_error_bounds()
    assert np.array_equal(c, c1) and np.array_equal(b, b1)
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


YEAR = 31536000


@pytest.fixture(scope="module")
def big():
    """The resident 1B-row table in (day bucket, Morton) order and the mixes'
    collections (10 to 5000 polygons)."""
    n = 1_000_000_000
    pts, t, fare, pas = gen(n, 7)
    h = rj.Handler()
    table = rj.Table(h, pts, (t, fare, pas), COLS_TYPES, time_col=0, time0=T0, bucket_seconds=86400)
    del pts, t, fare, pas
    gc.collect()
    names = ("voronoi10", "voronoi40", "voronoi260", "voronoi900", "tracts2200", "voronoi5000")
    lib = {k: rj.PolySet(h, *rj.poly_set(k)) for k in names}
    yield h, table, lib
    for ps in lib.values():
        ps.close()
    table.close()
    h.close()
    gc.collect()


def same(a, b):
    return (a.status == 0 and b.status == 0 and np.array_equal(a.counts, b.counts)
            and np.array_equal(a.hi, b.hi) and np.array_equal(a.lo, b.lo))


def window_frac(rng, lo_frac, hi_frac):
    f = float(np.exp(rng.uniform(np.log(lo_frac), np.log(hi_frac))))
    length = max(1, int(f * YEAR))
    start = T0 + int(rng.uniform(0, max(0.0, 1.0 - f)) * YEAR)
    return [(0, rj.GTE, start), (0, rj.LTE, start + length - 1)]


def oracle_checked(table, lib, specs):
    """Queries over 3M rows of the table that straddle the start of a 30 %
    time window (the rows before it fail the time filter), with their
    oracle results on the rows read back: raster against the pixel-center
    emulation, hybrid and index against the CPUJoin restatement."""
    lo, hi = T0 + 200 * 86400, T0 + 200 * 86400 + int(0.3 * YEAR)
    b0, _ = table.time_rows(lo, hi)
    rb, re = b0 - 1_000_000, b0 + 2_000_000
    xy, cols = table.read(rb, re - rb)
    out = []
    for name, fn, eps, agg, extra in specs:
        ps = lib[name]
        cons = [(0, rj.GTE, lo), (0, rj.LTE, hi)] + extra
        q = rj.make_query(ps, fn, agg, 1 if agg == rj.AVG else -1, cons, eps=eps, index_res=256, rows=(rb, re))
        pxy, poff = rj.poly_set(name)
        pset = O.PointSet(xy, cols, COLS_TYPES, cons, agg_type=agg, agg_col=1)
        if fn == rj.RASTER:
            want = O.raster_join(O.plan(O.poly_bbox(pxy).astype(np.float64), eps), pxy, poff, pset)
        else:
            want = O.join_exact(O.Grid(pxy, poff, 256), pset)[:3]
        out.append((q, ps.np, want, agg))
    return out


def check_oracle(r, want, agg):
    assert r.status == 0
    assert np.array_equal(r.counts, want[0])
    if agg == rj.AVG:
        assert np.array_equal(r.hi, want[1]) and np.array_equal(r.lo, want[2])
    assert r.counts.sum() > 0


def test_batch_1B_64_queries_autotuned(big):
    """64 raster (70 %) and hybrid (30 %) queries fused over the 1B-row
    table after rj_autotune, each equal to its single-query result, plus six
    row-range queries in the same passes equal to the oracle."""
    h, table, lib = big
    rng = np.random.default_rng(41)
    names = ("voronoi10", "voronoi40", "voronoi260", "voronoi900", "tracts2200")
    qs, nps = [], []
    for _ in range(64):
        ps = lib[names[int(rng.integers(len(names)))]]
        fn = rj.RASTER if rng.random() < 0.7 else rj.HYBRID
        eps = float((5.0, 10.0, 50.0, 100.0)[int(rng.integers(4))])
        cons = window_frac(rng, 1e6 / table.n, 1.0)
        agg = rj.AVG if rng.random() < 0.5 else rj.COUNT
        qs.append(rj.make_query(ps, fn, agg, 1 if agg == rj.AVG else -1, cons, eps=eps, index_res=256))
        nps.append(ps.np)
    rows = [table.time_rows(q.cons[0].val, q.cons[1].val) for q in qs]
    sel = [e - b for b, e in rows]
    assert min(sel) < 20_000_000 and max(sel) > 500_000_000
    chosen = h.autotune(table, qs)
    assert chosen["tile_shift"] in (3, 4, 5)
    checked = oracle_checked(table, lib, [("voronoi260", rj.RASTER, 10.0, rj.AVG, []),
                                          ("tracts2200", rj.RASTER, 50.0, rj.COUNT, []),
                                          ("voronoi40", rj.RASTER, 100.0, rj.AVG, [(2, rj.LTE, 2)]),
                                          ("tracts2200", rj.HYBRID, 20.0, rj.AVG, []),
                                          ("voronoi900", rj.HYBRID, 10.0, rj.COUNT, [(2, rj.GTE, 3)]),
                                          ("voronoi10", rj.HYBRID, 5.0, rj.AVG, [])])
    allq = qs + [c[0] for c in checked]
    alln = nps + [c[1] for c in checked]
    batch = h.execute_batch(table, allq, alln)
    for i, (q, k, b) in enumerate(zip(allq, alln, batch)):
        assert b.status == 0 and b.path == rj.PATH_FUSED, i
        assert same(h.execute_query(table, q, k), b), i
    for (q, k, want, agg), b in zip(checked, batch[64:]):
        check_oracle(b, want, agg)


def test_serving_1B_mix_with_5000_polygons(big):
    """The serving loop on the 1B-row table: Poisson arrivals at 250 qps of
    the serving mix (60/30/10 raster/hybrid/index, 10-5000 polygons, windows
    log-uniform from 1 hour to 1 year, 30 % with a passengers filter), with
    six row-range queries interleaved: every served result equals the query
    run alone, the row-range ones equal the oracle."""
    h, table, lib = big
    rng = np.random.default_rng(43)
    names = list(lib)
    qs, nps = [], []
    for _ in range(300):
        ps = lib[names[int(rng.integers(len(names)))]]
        u = rng.random()
        fn = rj.RASTER if u < 0.6 else (rj.HYBRID if u < 0.9 else rj.INDEX)
        eps = float((10.0, 20.0, 50.0, 100.0)[int(rng.integers(4))])
        cons = window_frac(rng, 3600.0 / YEAR, 1.0)
        if rng.random() < 0.3:
            cons.append((2, rj.LTE, int(rng.integers(1, 7))))
        agg = rj.AVG if rng.random() < 0.5 else rj.COUNT
        qs.append(rj.make_query(ps, fn, agg, 1 if agg == rj.AVG else -1, cons, eps=eps, index_res=256))
        nps.append(ps.np)
    assert any(k == 5000 for k in nps) and any(k == 10 for k in nps)
    checked = oracle_checked(table, lib, [("voronoi5000", rj.RASTER, 20.0, rj.AVG, []),
                                          ("voronoi5000", rj.HYBRID, 10.0, rj.COUNT, [(2, rj.LTE, 3)]),
                                          ("tracts2200", rj.INDEX, 10.0, rj.AVG, []),
                                          ("voronoi260", rj.INDEX, 10.0, rj.COUNT, [(2, rj.GTE, 2)]),
                                          ("voronoi40", rj.RASTER, 50.0, rj.COUNT, []),
                                          ("voronoi900", rj.HYBRID, 20.0, rj.AVG, [])])
    for j, c in enumerate(checked):
        qs.insert(40 * j + 7, c[0])
        nps.insert(40 * j + 7, c[1])
    alone = [h.execute_query(table, q, k) for q, k in zip(qs, nps)]
    srv = rj.Server(h, table)
    tickets = []

    def client():
        t_next = time.perf_counter()
        for q, k in zip(qs, nps):
            t_next += rng.exponential(1.0 / 250.0)
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(srv.submit(q, k))

    th = threading.Thread(target=client)
    th.start()
    th.join()
    outs = [srv.poll(tk, wait_ms=120000) for tk in tickets]
    st = srv.stats()
    srv.close()
    for i, (a, b) in enumerate(zip(alone, outs)):
        assert b is not None, i
        assert same(a, b), i
    assert st["completed"] == len(qs) and st["failed"] == 0
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    for j, (q, k, want, agg) in enumerate(checked):
        check_oracle(outs[40 * j + 7], want, agg)
    print(f"served {len(qs)} queries at 250 qps offered: p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms, "
          f"mean batch {st['mean_batch']:.2f}")
