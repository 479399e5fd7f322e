"""Resident collections, time-bucketed tables, fused batches, the autotuner
and the serving loop: every query's results are exactly those it gets alone
(counts and fixed-point sums bit-identical), and the single-query results
are pinned to the oracle here too for a sample of the mix."""
import threading
import time

import numpy as np
import pytest

from oracle import pyoracle as O
import rasterjoin as rj
from common import COLS_TYPES, T0, points, polys

pytestmark = pytest.mark.gpu
N = 2_000_000
YEAR = 31536000


@pytest.fixture(scope="module")
def data():
    pts, t, fare, pas = points(N, 21)
    return pts, (t, fare, pas)


def bucketed(h, data, span=86400):
    pts, cols = data
    return rj.Table(h, pts, cols, COLS_TYPES, time_col=0, time0=T0, bucket_seconds=span)


def window(start_frac, length):
    lo = T0 + int(start_frac * YEAR)
    return [(0, rj.GTE, lo), (0, rj.LTE, lo + length)]


def same(a, b):
    return (np.array_equal(a.counts, b.counts) and np.array_equal(a.hi, b.hi)
            and np.array_equal(a.lo, b.lo))


def polysets(h, names=("voronoi260", "tracts400", "voronoi20", "stars")):
    return [rj.PolySet(h, *polys(n)) for n in names]


def mixed_queries(sets, n, seed, eps_choices=(10.0, 50.0, 100.0), fns=None):
    """A seeded query mix: 60/30/10 raster/hybrid/index, log-uniform time
    windows from 1 hour to 1 year, an optional passengers filter, count or
    avg(fare)."""
    rng = np.random.default_rng(seed)
    fns = fns or [rj.RASTER] * 6 + [rj.HYBRID] * 3 + [rj.INDEX]
    qs, nps, meta = [], [], []
    for _ in range(n):
        k = int(rng.integers(len(sets)))
        fn = int(fns[int(rng.integers(len(fns)))])
        eps = float(eps_choices[int(rng.integers(len(eps_choices)))])
        length = int(np.exp(rng.uniform(np.log(3600), np.log(YEAR))))
        cons = window(rng.uniform(0, 1 - length / YEAR), length)
        if rng.random() < 0.3:
            cons += [(2, rj.LTE, int(rng.integers(1, 7)))]
        agg = rj.AVG if rng.random() < 0.5 else rj.COUNT
        q = rj.make_query(sets[k], fn, agg, 1 if agg == rj.AVG else -1, cons, eps=eps,
                          index_res=256)
        qs.append(q)
        nps.append(sets[k].np)
        meta.append((k, fn, eps, cons, agg))
    return qs, nps, meta


def test_polyset_matches_set_polygons(handler, data):
    h = handler
    pts, cols = data
    t = rj.Table(h, pts, cols, COLS_TYPES)
    xy, off = polys("voronoi260")
    cons = window(0.2, 40 * 86400)
    ref = {}
    h.set_polygons(xy, off)
    h.bind_table(t)
    h.set_query_constraints(cons)
    h.set_aggregation(rj.AVG, 1)
    h.set_accuracy_distance(20.0)
    for fn in (rj.RASTER, rj.INDEX, rj.HYBRID):
        c, _ = h.execute_function(fn)
        ref[fn] = (c, *h.sums_fixed())
    ps = rj.PolySet(h, xy, off)
    h.bind_polyset(ps)
    for fn in (rj.RASTER, rj.INDEX, rj.HYBRID):
        c, _ = h.execute_function(fn)
        hi, lo = h.sums_fixed()
        assert np.array_equal(c, ref[fn][0]) and np.array_equal(hi, ref[fn][1])
        assert np.array_equal(lo, ref[fn][2])
        q = rj.make_query(ps, fn, rj.AVG, 1, cons, eps=20.0, index_res=256)
        r = h.execute_query(t, q, ps.np)
        assert np.array_equal(r.counts, ref[fn][0]) and np.array_equal(r.hi, ref[fn][1])
        assert np.array_equal(r.lo, ref[fn][2])


def test_time_buckets_narrow_rows_and_keep_results(handler, data):
    h = handler
    pts, cols = data
    plain = rj.Table(h, pts, cols, COLS_TYPES)
    tb = bucketed(h, data)
    # bucket rows: every in-window row is inside the narrowed range
    lo, hi = T0 + 100 * 86400 + 5000, T0 + 103 * 86400 + 17
    b, e = tb.time_rows(lo, hi)
    assert 0 <= b < e <= N and e - b < N // 50
    _, tcols = tb.read()
    times = tcols[0]
    inwin = (times >= lo) & (times <= hi)
    assert inwin[:b].sum() == 0 and inwin[e:].sum() == 0
    assert inwin[b:e].sum() == ((cols[0] >= lo) & (cols[0] <= hi)).sum()
    # same results with and without buckets
    sets = polysets(h, ("voronoi260", "stars"))
    qs, nps, _ = mixed_queries(sets, 12, seed=3)
    for q, k in zip(qs, nps):
        a = h.execute_query(plain, q, k)
        c = h.execute_query(tb, q, k)
        assert same(a, c)


def test_batch_equals_single_and_oracle(handler, data):
    h = handler
    tb = bucketed(h, data)
    sets = polysets(h)
    qs, nps, meta = mixed_queries(sets, 40, seed=5)
    single = [h.execute_query(tb, q, k) for q, k in zip(qs, nps)]
    batch = h.execute_batch(tb, qs, nps)
    for i, (a, b) in enumerate(zip(single, batch)):
        assert b.status == 0, i
        assert same(a, b), i
    fused = sum(b.path == rj.PATH_FUSED for b in batch)
    # raster queries over the overlapping star polygons run alone (canvas path)
    alone_expected = sum(1 for m in meta if m[0] == 3 and m[1] == rj.RASTER)
    assert fused == len(qs) - alone_expected
    # oracle: index and hybrid against the CPUJoin restatement, raster against
    # the pixel-center emulation, on a sample of the mix
    pts, cols = data
    checked = 0
    for (k, fn, eps, cons, agg), r in zip(meta, batch):
        if checked >= 6:
            break
        xy, off = polys(("voronoi260", "tracts400", "voronoi20", "stars")[k])
        ps = O.PointSet(pts, cols, COLS_TYPES, cons, agg_type=agg, agg_col=1)
        if fn in (rj.INDEX, rj.HYBRID):
            oc, ohi, olo, _ = O.join_exact(O.Grid(xy, off, 256), ps)
        elif k != 3:
            oc, ohi, olo = O.raster_join(O.plan(O.poly_bbox(xy), eps), xy, off, ps)
        else:
            continue
        assert np.array_equal(r.counts, oc)
        if agg == rj.AVG:
            assert np.array_equal(r.hi, ohi) and np.array_equal(r.lo, olo)
        checked += 1
    assert checked >= 4


def test_batch_graph_replays_match_eager(handler, data):
    h = handler
    tb = bucketed(h, data)
    sets = polysets(h, ("voronoi260", "voronoi20"))
    qa, na, _ = mixed_queries(sets, 10, seed=7)
    qb, nb, _ = mixed_queries(sets, 23, seed=8)
    h.set_option(rj.OPT_GRAPH, 0)
    ea = h.execute_batch(tb, qa, na)
    eb = h.execute_batch(tb, qb, nb)
    h.set_option(rj.OPT_GRAPH, 1)
    for rep in range(3):
        ga = h.execute_batch(tb, qa, na)
        h.execute_query(tb, qb[0], nb[0])  # eager work between replays
        gb = h.execute_batch(tb, qb, nb)
        h.sync()
        for x, y in zip(ea, ga):
            assert same(x, y), rep
        for x, y in zip(eb, gb):
            assert same(x, y), rep


def test_batch_many_passes_and_columns(handler, data):
    """100 queries over 6 columns: a pass holds at most 64 queries, so they run
    as two passes."""
    h = handler
    pts, (t, fare, pas) = data
    cols6 = (t, fare, pas, t.copy(), fare.copy(), pas.copy())
    types6 = COLS_TYPES * 2
    tab = rj.Table(h, pts, cols6, types6, time_col=0, time0=T0, bucket_seconds=7 * 86400)
    ps = rj.PolySet(h, *polys("voronoi260"))
    rng = np.random.default_rng(11)
    qs = []
    for i in range(100):
        length = int(rng.integers(3600, YEAR // 4))
        cons = window(rng.uniform(0, 0.7), length)
        if i % 3 == 0:
            cons += [(5, rj.GTE, 3), (3, rj.GTE, T0)]
            q = rj.make_query(ps, (rj.RASTER, rj.HYBRID)[i % 4 // 2], rj.AVG, 4, cons, eps=50.0,
                              index_res=128)
        else:
            cons += [(2, rj.LTE, 4)]
            q = rj.make_query(ps, (rj.RASTER, rj.INDEX)[i % 4 // 2], rj.AVG, 1, cons, eps=50.0,
                              index_res=128)
        qs.append(q)
    nps = [ps.np] * len(qs)
    batch = h.execute_batch(tab, qs, nps)
    for i, q in enumerate(qs):
        a = h.execute_query(tab, q, ps.np)
        assert batch[i].status == 0 and batch[i].path == rj.PATH_FUSED
        assert same(a, batch[i]), i


def test_batch_errorbounds_and_errors(handler, data):
    h = handler
    tb = bucketed(h, data)
    ps = rj.PolySet(h, *polys("voronoi20"))
    good = rj.make_query(ps, rj.RASTER, rj.COUNT, -1, window(0.1, 86400 * 30), eps=50.0)
    eb = rj.make_query(ps, rj.ERRORBOUNDS, rj.COUNT, -1, window(0.1, 86400 * 30), eps=50.0)
    bad = rj.make_query(ps, rj.RASTER, rj.AVG, 7, (), eps=50.0)  # no column 7
    out = h.execute_batch(tb, [good, eb, bad], [ps.np] * 3)
    assert out[0].status == 0 and out[1].status == 0 and out[2].status == -1  # RJ_ERR_ARG
    single_eb = h.execute_query(tb, eb, ps.np)
    assert np.array_equal(single_eb.bounds, out[1].bounds)
    assert np.array_equal(single_eb.counts, out[1].counts)
    assert out[1].path == rj.PATH_SINGLE
    # the raster count lies in its guaranteed range [count - fp1, count + fn1]
    # around the exact (index) count
    ex = h.execute_query(tb, rj.make_query(ps, rj.INDEX, rj.COUNT, -1, window(0.1, 86400 * 30),
                                           index_res=64), ps.np)
    b = out[1].bounds.reshape(4, -1)
    assert np.array_equal(out[0].counts, out[1].counts)  # label-map path == canvas path
    assert np.all(ex.counts >= out[1].counts - b[3]) and np.all(ex.counts <= out[1].counts + b[2])


def test_autotune_keeps_results(handler, data):
    h = handler
    tb = bucketed(h, data)
    sets = polysets(h, ("voronoi260", "tracts400"))
    qs, nps, _ = mixed_queries(sets, 6, seed=13)
    qs[-1].index_res = 0  # grid resolution left to the engine
    before = h.execute_batch(tb, qs, nps)
    chosen = h.autotune(tb, qs)
    assert chosen["tile_shift"] in (3, 4, 5) and chosen["rows"] in (4, 8)
    assert chosen["workgroups"] in (1024, 2048, 4096) and chosen["chunk_rows"] in (4096, 8192, 16384)
    assert chosen["grid_res"] in (128, 256, 512, 1024)
    after = h.execute_batch(tb, qs, nps)
    for a, b in zip(before[:-1], after[:-1]):
        assert same(a, b)
    # the automatic-resolution query equals a single query at the chosen resolution
    q = qs[-1]
    q.index_res = chosen["grid_res"]
    assert same(h.execute_query(tb, q, nps[-1]), after[-1])


def _serve(h, tb, qs, nps, rate_qps, seed, **cfg):
    srv = rj.Server(h, tb, **cfg)
    rng = np.random.default_rng(seed)
    tickets = [None] * len(qs)

    def client():
        for i, q in enumerate(qs):
            tickets[i] = srv.submit(q, nps[i])
            if rate_qps:
                time.sleep(rng.exponential(1.0 / rate_qps))

    th = threading.Thread(target=client)
    th.start()
    th.join()
    outs = [srv.poll(tk, wait_ms=60000) for tk in tickets]
    st = srv.stats()
    srv.close()
    return outs, st


def test_server_results_alone_equal_in_mix(handler, data):
    h = handler
    tb = bucketed(h, data)
    sets = polysets(h)
    qs, nps, _ = mixed_queries(sets, 60, seed=17)
    alone = [h.execute_query(tb, q, k) for q, k in zip(qs, nps)]
    outs, st = _serve(h, tb, qs, nps, rate_qps=0, seed=1, max_batch=16)  # one burst
    for i, (a, b) in enumerate(zip(alone, outs)):
        assert b is not None and b.status == 0, i
        assert same(a, b), i
    assert st["completed"] == len(qs) and st["failed"] == 0
    assert st["batches"] >= 4 and st["mean_batch"] > 1.0
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    outs2, st2 = _serve(h, tb, qs[:30], nps[:30], rate_qps=400, seed=2)  # Poisson arrivals
    for a, b in zip(alone[:30], outs2):
        assert same(a, b)
    assert st2["completed"] == 30 and st2["qps"] > 0


def test_server_memory_admission_and_eviction(handler, data):
    """A budget that holds a few label maps at a time: queries are admitted as
    memory frees up, cached structures are evicted, results stay exact."""
    h = handler
    tb = bucketed(h, data)
    sets = polysets(h, ("voronoi260", "tracts400", "voronoi20"))
    qs, nps, _ = mixed_queries(sets, 24, seed=19, eps_choices=(10.0, 20.0, 50.0),
                               fns=[rj.RASTER, rj.RASTER, rj.HYBRID])
    alone = [h.execute_query(tb, q, k) for q, k in zip(qs, nps)]
    outs, st = _serve(h, tb, qs, nps, rate_qps=0, seed=3, max_batch=8, mem_mb=700)
    for i, (a, b) in enumerate(zip(alone, outs)):
        assert b is not None and b.status == 0, i
        assert same(a, b), i
    assert st["completed"] == len(qs)


@pytest.mark.parametrize("share", [1, 0])
def test_groups_on_one_label_map_equal_single(handler, data, share):
    """Fused raster queries on the same label map (same collection and
    accuracy) share each chunk's labels; each still applies its own window,
    filter and aggregate. Mixed with hybrid and index queries of the same
    collections, results equal the queries run alone, sharing on and off."""
    h = handler
    tb = bucketed(h, data)
    sets = polysets(h, ("voronoi260", "tracts400"))
    rng = np.random.default_rng(23)
    qs, nps = [], []
    for i in range(30):
        ps = sets[i % 2]
        fn = (rj.RASTER, rj.RASTER, rj.HYBRID, rj.INDEX)[i % 4]
        length = int(np.exp(rng.uniform(np.log(3600), np.log(YEAR))))
        cons = window(rng.uniform(0, 1 - length / YEAR), length)
        if i % 3 == 0:
            cons += [(2, rj.LTE, int(rng.integers(1, 7)))]
        agg = rj.AVG if i % 5 < 2 else rj.COUNT
        qs.append(rj.make_query(ps, fn, agg, 1 if agg == rj.AVG else -1, cons, eps=20.0, index_res=256))
        nps.append(ps.np)
    h.set_option(rj.OPT_BATCH_SHARE, share)
    try:
        batch = h.execute_batch(tb, qs, nps)
    finally:
        h.set_option(rj.OPT_BATCH_SHARE, 1)
    for i, (q, k, b) in enumerate(zip(qs, nps, batch)):
        assert b.status == 0 and b.path == rj.PATH_FUSED, i
        assert same(h.execute_query(tb, q, k), b), i


@pytest.mark.parametrize("share", [1, 0])
def test_plain_members_of_a_group_equal_single(handler, data, share):
    """Members of one label-map group whose constraints a chunk passes in full
    (no constraints, or a window covering the chunk) take one pass of it
    together: counts go to every member, sums only to the members that
    aggregate the same column. Count, avg and sum over two columns, with and
    without windows and a filter: results equal each query run alone."""
    h = handler
    tb = bucketed(h, data)
    ps = polysets(h, ("voronoi260",))[0]
    rng = np.random.default_rng(31)
    aggs = [(rj.COUNT, -1), (rj.AVG, 1), (rj.SUM, 1), (rj.AVG, 2), (rj.SUM, 2), (rj.COUNT, -1)]
    qs = []
    for i in range(24):
        agg, col = aggs[i % len(aggs)]
        if i % 4 == 0:
            cons = []
        elif i % 4 == 1:
            cons = window(0.1, YEAR // 2)
        elif i % 4 == 2:
            length = int(np.exp(rng.uniform(np.log(86400), np.log(YEAR))))
            cons = window(rng.uniform(0, 1 - length / YEAR), length)
        else:
            cons = [(2, rj.LTE, int(rng.integers(1, 7)))]
        qs.append(rj.make_query(ps, rj.RASTER, agg, col, cons, eps=20.0))
    nps = [ps.np] * len(qs)
    h.set_option(rj.OPT_BATCH_SHARE, share)
    try:
        batch = h.execute_batch(tb, qs, nps)
    finally:
        h.set_option(rj.OPT_BATCH_SHARE, 1)
    for i, (q, b) in enumerate(zip(qs, batch)):
        assert b.status == 0 and b.path == rj.PATH_FUSED, i
        assert same(h.execute_query(tb, q, ps.np), b), i
