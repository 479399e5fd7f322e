"""rj_execute_query_noalloc, the entry point that allocates nothing: its
results (counts, exact sums hi/lo, double sums) equal rj_execute_query's
bit for bit for raster, hybrid and both index paths, on two calls in a row
(the second with fewer rows and a smaller polygon collection), including a
hybrid pass whose border queue overflows the scratch's room; and its limits
return status codes."""
import numpy as np
import pytest

import rasterjoin as rj
from common import COLS_TYPES, T0, points, polys

pytestmark = pytest.mark.gpu
YEAR = 31536000


def same(a, b):
    return (a.status == 0 and b.status == 0 and np.array_equal(a.counts, b.counts)
            and np.array_equal(a.hi, b.hi) and np.array_equal(a.lo, b.lo)
            and np.array_equal(a.sums.view(np.uint64), b.sums.view(np.uint64)))


def grid_polygons(nx, ny, box=50000.0):
    """nx * ny axis-aligned squares tiling the box."""
    gx = np.linspace(0.0, box, nx + 1, dtype=np.float32)
    gy = np.linspace(0.0, box, ny + 1, dtype=np.float32)
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    i, j = i.ravel(), j.ravel()
    xy = np.stack([gx[i], gy[j], gx[i + 1], gy[j], gx[i + 1], gy[j + 1], gx[i], gy[j + 1]], 1)
    return xy.astype(np.float32).ravel(), (np.arange(len(i) + 1) * 4).astype(np.int32)


def near_edges(xy, off, n, seed):
    """n points within 7 m of the polygons' edges (border points of a hybrid pass)."""
    rng = np.random.default_rng(seed)
    v = xy.reshape(-1, 2).astype(np.float64)
    prev = np.arange(len(v)) - 1
    prev[off[:-1]] = off[1:] - 1
    e = rng.integers(0, len(v), n)
    a, b = v[prev[e]], v[e]
    mid = a + rng.uniform(0, 1, n)[:, None] * (b - a)
    nrm = np.stack([a[:, 1] - b[:, 1], b[:, 0] - a[:, 0]], 1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True) + 1e-30
    return (mid + nrm * rng.uniform(-7, 7, n)[:, None]).astype(np.float32).ravel()


@pytest.fixture(scope="module")
def data():
    pts, t, fare, pas = points(3_000_000, 41)
    return pts, (t, fare, pas)


@pytest.mark.parametrize("fn,path", [(rj.RASTER, 0), (rj.HYBRID, 0), (rj.INDEX, 0), (rj.INDEX, 1)],
                         ids=["raster", "hybrid", "index-via-labels", "index-grid-pip"])
def test_noalloc_equals_execute_query(handler, data, fn, path):
    h = handler
    h.set_option(rj.OPT_INDEX_PATH, path)
    pts, cols = data
    table = rj.Table(h, pts, cols, COLS_TYPES, time_col=0, time0=T0, bucket_seconds=86400)
    big = rj.PolySet(h, *polys("tracts400"))
    small = rj.PolySet(h, *polys("voronoi20"))
    window = [(0, rj.GTE, T0 + 1_000_000), (0, rj.LTE, T0 + 1_000_000 + YEAR // 2)]
    calls = [
        rj.make_query(big, fn, rj.AVG, 1, window + [(2, rj.LTE, 4)], eps=10.0, index_res=256),
        # the second call: fewer rows (a shorter window and a row range) and fewer polygons
        rj.make_query(small, fn, rj.AVG, 1, [(0, rj.GTE, T0), (0, rj.LTE, T0 + YEAR // 20)], eps=20.0,
                      index_res=128, rows=(1000, 2_000_000)),
        rj.make_query(small, fn, rj.COUNT, -1, [], eps=20.0, index_res=128, rows=(5, 700_000)),
    ]
    nps = [big.np, small.np, small.np]
    want = [h.execute_query(table, q, k) for q, k in zip(calls, nps)]  # prepares the structures
    for q, k, w in zip(calls, nps, want):
        got = h.execute_query_noalloc(table, q, k)
        assert got.path == rj.PATH_SINGLE
        assert same(got, w), (fn, path, k)
    assert want[0].counts.sum() > 10_000
    h.set_option(rj.OPT_INDEX_PATH, 0)


QUEUE_ROOM_400 = (32 * 2**20 - 13056) // 16   # scratch queue entries with 400 polygons


def test_noalloc_border_queue_overflow(handler):
    """About 3M border points: more than the scratch's queue room (about
    2.1M entries with 400 polygons), so the pass reruns with inline exact
    tests; the results still equal rj_execute_query's, then a smaller call
    after it (whose queue fits)."""
    h = handler
    xy, off = polys("tracts400")
    pts = near_edges(xy, off, 3_200_000, 5)
    ids = np.arange(len(pts) // 2, dtype=np.uint32)
    table = rj.Table(h, pts, (ids,), (rj.UINT,))
    ps = rj.PolySet(h, xy, off)
    for rows in ((0, -1), (0, 300_000)):
        q = rj.make_query(ps, rj.HYBRID, rj.AVG, 0, [], eps=10.0, index_res=256, rows=rows)
        h.execute_query(table, q, ps.np)
        want = h.execute_query(table, q, ps.np)   # its queue sized from the first run's border share
        queued = h.exact_points()
        got = h.execute_query_noalloc(table, q, ps.np)
        assert same(got, want), rows
        if rows[1] < 0:
            assert queued > QUEUE_ROOM_400, queued
            assert want.counts.sum() > 1_000_000
        else:
            assert 0 < queued < QUEUE_ROOM_400, queued


def test_noalloc_limits_are_status_codes(handler, data):
    h = handler
    pts, cols = data
    table = rj.Table(h, pts[:200_000], tuple(c[:100_000] for c in cols), COLS_TYPES)
    ps = rj.PolySet(h, *polys("voronoi20"))
    # structures not resident yet
    q = rj.make_query(ps, rj.RASTER, rj.COUNT, -1, [], eps=30.0)
    with pytest.raises(rj.RJError) as e:
        h.execute_query_noalloc(table, q, ps.np)
    assert e.value.code == rj.ERR_STATE
    h.execute_query(table, q, ps.np)
    assert h.execute_query_noalloc(table, q, ps.np).status == 0
    # errorbounds needs a canvas
    qb = rj.make_query(ps, rj.ERRORBOUNDS, rj.COUNT, -1, [], eps=30.0)
    h.execute_query(table, qb, ps.np)
    with pytest.raises(rj.RJError) as e:
        h.execute_query_noalloc(table, qb, ps.np)
    assert e.value.code == rj.ERR_UNSUPPORTED
    # more polygons than the scratch holds results for
    gxy, goff = grid_polygons(300, 220)
    big = rj.PolySet(h, gxy, goff)
    qi = rj.make_query(big, rj.INDEX, rj.COUNT, -1, [], index_res=256)
    h.set_option(rj.OPT_INDEX_PATH, 1)
    h.execute_query(table, qi, big.np)
    with pytest.raises(rj.RJError) as e:
        h.execute_query_noalloc(table, qi, big.np)
    assert e.value.code == rj.ERR_LIMIT
    h.set_option(rj.OPT_INDEX_PATH, 0)
