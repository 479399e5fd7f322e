"""Host-resident (out-of-core) point tables: rows stay in host memory and
stream through the device in chunks per query, as the reference's point
passes do (GLFunction.cpp:329-337, RasterJoin.cpp:214-236). Every join type
gives exactly the results of the same rows in a device-resident table
(counts and fixed-point sums bit-identical, error bounds equal), whatever
the chunk size; the resident results are pinned to the oracle by the other
GPU tests, and a sample is checked against it here too."""
import numpy as np
import pytest

from oracle import pyoracle as O
import rasterjoin as rj
from common import COLS_TYPES, T0, points, polys

pytestmark = pytest.mark.gpu
N = 1_000_003  # not a multiple of any chunk size below
YEAR = 31536000


@pytest.fixture(scope="module")
def data():
    pts, t, fare, pas = points(N, 31)
    return pts, (t, fare, pas)


def same(a, b):
    return (np.array_equal(a.counts, b.counts) and np.array_equal(a.hi, b.hi)
            and np.array_equal(a.lo, b.lo))


def queries(ps, cons):
    out = []
    for fn in (rj.RASTER, rj.HYBRID, rj.INDEX):
        for agg in (rj.COUNT, rj.AVG):
            out.append((fn, rj.make_query(ps, fn, agg, 1 if agg == rj.AVG else -1, cons,
                                          eps=20.0, index_res=256)))
    return out


@pytest.mark.parametrize("chunk", [0, 65536, 300_000])
def test_host_table_equals_resident(handler, data, chunk):
    h = handler
    pts, cols = data
    dev = rj.Table(h, pts, cols, COLS_TYPES)
    host = rj.Table(h, pts, cols, COLS_TYPES, host=True)
    h.set_option(rj.OPT_STREAM_ROWS, chunk)
    try:
        for name in ("voronoi260", "stars"):
            ps = rj.PolySet(h, *polys(name))
            cons = [(0, rj.GTE, T0 + 40 * 86400), (2, rj.LTE, 4)]
            for fn, q in queries(ps, cons):
                a = h.execute_query(dev, q, ps.np)
                b = h.execute_query(host, q, ps.np)
                assert same(a, b), (name, fn)
    finally:
        h.set_option(rj.OPT_STREAM_ROWS, 0)


def test_host_table_errorbounds_and_row_ranges(handler, data):
    h = handler
    pts, cols = data
    xy, off = polys("voronoi20")
    # row ranges select the same rows only when both tables keep input order
    dev = rj.Table(h, pts, cols, COLS_TYPES, organize=False)
    host = rj.Table(h, pts, cols, COLS_TYPES, host=True)
    h.set_option(rj.OPT_STREAM_ROWS, 100_000)
    try:
        h.set_polygons(xy, off)
        h.set_query_constraints([])
        h.set_aggregation(rj.COUNT, -1)
        h.set_accuracy_distance(50.0)
        for rb, re in ((0, N), (12345, 777_777), (500_000, 500_000)):
            out = []
            for t in (dev, host):
                h.bind_table(t, rb, re)
                for fn in (rj.ERRORBOUNDS, rj.RASTER, rj.INDEX):
                    c, _ = h.execute_function(fn)
                    out.append((c, h.get_error_bounds() if fn == rj.ERRORBOUNDS else None))
            for (c1, b1), (c2, b2) in zip(out[:3], out[3:]):
                assert np.array_equal(c1, c2)
                if b1 is not None:
                    assert np.array_equal(b1, b2)
        # the raster pass over [0, N) ran in ceil(N / chunk) passes
        h.bind_table(host)
        h.execute_function(rj.RASTER)
        line = h.print_time_stats(rj.RASTER).split()
        assert int(line[3]) == -(-N // 100_096)
    finally:
        h.set_option(rj.OPT_STREAM_ROWS, 0)


def test_host_table_time_buckets(handler, data):
    h = handler
    pts, cols = data
    dev = rj.Table(h, pts, cols, COLS_TYPES)
    host = rj.Table(h, pts, cols, COLS_TYPES, time_col=0, time0=T0, bucket_seconds=86400, host=True)
    lo, hi = T0 + 100 * 86400 + 5000, T0 + 103 * 86400 + 17
    b, e = host.time_rows(lo, hi)
    _, tcols = host.read()
    inwin = (tcols[0] >= lo) & (tcols[0] <= hi)
    assert inwin[:b].sum() == 0 and inwin[e:].sum() == 0
    assert inwin[b:e].sum() == ((cols[0] >= lo) & (cols[0] <= hi)).sum()
    assert e - b < N // 50
    ps = rj.PolySet(h, *polys("tracts400"))
    h.set_option(rj.OPT_STREAM_ROWS, 4096)
    try:
        for fn, q in queries(ps, [(0, rj.GTE, lo), (0, rj.LTE, hi)]):
            assert same(h.execute_query(dev, q, ps.np), h.execute_query(host, q, ps.np)), fn
    finally:
        h.set_option(rj.OPT_STREAM_ROWS, 0)


def test_host_table_batch_and_oracle(handler, data):
    h = handler
    pts, cols = data
    host = rj.Table(h, pts, cols, COLS_TYPES, host=True)
    xy, off = polys("voronoi260")
    ps = rj.PolySet(h, xy, off)
    cons = [(0, rj.GTE, T0 + 200 * 86400)]
    qs = [q for _, q in queries(ps, cons)]
    h.set_option(rj.OPT_STREAM_ROWS, 262144)
    try:
        single = [h.execute_query(host, q, ps.np) for q in qs]
        batch = h.execute_batch(host, qs, [ps.np] * len(qs))
    finally:
        h.set_option(rj.OPT_STREAM_ROWS, 0)
    for a, b in zip(single, batch):
        assert b.status == 0 and same(a, b)
        assert b.path == rj.PATH_FUSED  # one stream of the table for the whole batch
    # hybrid avg (CPUJoin semantics) against the oracle's exact join
    g = O.Grid(xy, off, 256)
    exp = O.join_exact(g, O.PointSet(pts, cols, COLS_TYPES, cons, agg_type=rj.AVG, agg_col=1))
    r = single[3]
    assert np.array_equal(r.counts, exp[0])
    assert np.array_equal(r.hi, exp[1]) and np.array_equal(r.lo, exp[2])


def test_oversized_points_stream_in_passes():
    """rj_set_points with points beyond the memory budget keeps them in host
    memory and streams them (the reference's passes) instead of failing; an
    index join whose label map does not fit the budget runs on the grid."""
    xy, off = polys("voronoi20")
    pts, t, fare, pas = points(2_000_000, 2)
    cols = (t, fare, pas)
    res = {}
    for mb in (0, 116):  # unlimited, then 16 MB after the 100 MB reserve (26 MB of points)
        h = rj.Handler(gpu_mem_mb=mb)
        h.set_polygons(xy, off)
        h.set_points(pts, cols, COLS_TYPES)
        h.set_aggregation(rj.AVG, 1)
        h.set_accuracy_distance(100.0)
        for fn in (rj.RASTER, rj.INDEX):
            c, _ = h.execute_function(fn)
            res.setdefault(fn, []).append((c, *h.sums_fixed()))
        if mb:
            line = h.print_time_stats(rj.RASTER).split()
            assert int(line[3]) > 1  # passes
        h.close()
    for fn, (a, b) in res.items():
        for x, y in zip(a, b):
            assert np.array_equal(x, y), fn


@pytest.mark.parametrize("chunk", [0, 200_000])
def test_host_table_fused_batch_mix(handler, data, chunk):
    """A mixed batch (raster, hybrid, index; count and avg; time windows and a
    filter; four collections) over a time-bucketed host table streamed in
    chunks: every query equals the same query on the resident table."""
    h = handler
    pts, cols = data
    dev = rj.Table(h, pts, cols, COLS_TYPES, time_col=0, time0=T0, bucket_seconds=7 * 86400)
    host = rj.Table(h, pts, cols, COLS_TYPES, time_col=0, time0=T0, bucket_seconds=7 * 86400, host=True)
    sets = [rj.PolySet(h, *polys(k)) for k in ("voronoi260", "tracts400", "voronoi20", "stars")]
    rng = np.random.default_rng(9)
    qs, nps = [], []
    for i in range(24):
        ps = sets[i % len(sets)]
        length = int(np.exp(rng.uniform(np.log(86400), np.log(YEAR))))
        lo = T0 + int(rng.uniform(0, YEAR - length))
        cons = [(0, rj.GTE, lo), (0, rj.LTE, lo + length)]
        if i % 5 == 0:
            cons.append((2, rj.LTE, 3))
        fn = (rj.RASTER, rj.HYBRID, rj.INDEX)[i % 3]
        agg = rj.AVG if i % 2 else rj.COUNT
        qs.append(rj.make_query(ps, fn, agg, 1 if agg == rj.AVG else -1, cons, eps=(10.0, 50.0)[i % 2],
                                index_res=256))
        nps.append(ps.np)
    h.set_option(rj.OPT_STREAM_ROWS, chunk)
    try:
        batch = h.execute_batch(host, qs, nps)
    finally:
        h.set_option(rj.OPT_STREAM_ROWS, 0)
    fused = 0
    for i, (q, k, b) in enumerate(zip(qs, nps, batch)):
        assert b.status == 0, i
        assert same(h.execute_query(dev, q, k), b), i
        fused += b.path == rj.PATH_FUSED
    # raster queries over the overlapping stars run alone (canvas path)
    assert fused >= 24 - 2
