"""Table organisation (rj_table_create_ex / rj_set_points with organize): the
rows come out ordered by their Morton key (16 bits per axis over the bounds
of the table's finite points), or by (time bucket, Morton key) with a time
column; the order is stable (rows with equal keys keep their input order);
every column moves with its row. The keys are recomputed here with the same float32
steps; a row-number column checks the permutation."""
import numpy as np
import pytest

import rasterjoin as rj
from common import COLS_TYPES, T0, points

pytestmark = pytest.mark.gpu
DAY = 86400


def spread16(v):
    v = v.astype(np.uint64) & 0xFFFF
    v = (v | (v << 8)) & 0x00FF00FF
    v = (v | (v << 4)) & 0x0F0F0F0F
    v = (v | (v << 2)) & 0x33333333
    v = (v | (v << 1)) & 0x55555555
    return v


def morton(xy):
    x, y = xy[:, 0], xy[:, 1]
    ok = np.isfinite(x) & np.isfinite(y)
    if ok.any():
        mn = np.array([x[ok].min(), y[ok].min()], np.float32)
        mx = np.array([x[ok].max(), y[ok].max()], np.float32)
    else:
        mn, mx = np.zeros(2, np.float32), np.ones(2, np.float32)
    one = np.float32(65535.0)
    sx = one / np.maximum(mx[0] - mn[0], np.float32(1e-30))
    sy = one / np.maximum(mx[1] - mn[1], np.float32(1e-30))
    with np.errstate(invalid="ignore", over="ignore"):
        fx = (x - mn[0]) * sx
        fy = (y - mn[1]) * sy
    fx = np.nan_to_num(np.clip(fx, 0, 65535), nan=0.0)
    fy = np.nan_to_num(np.clip(fy, 0, 65535), nan=0.0)
    return spread16(fx.astype(np.uint32)) | (spread16(fy.astype(np.uint32)) << np.uint64(1))


@pytest.mark.parametrize("time_col", [-1, 0])
@pytest.mark.parametrize("n", [2, 1000, 300_001, 3_000_017])
def test_organized_order(handler, time_col, n):
    pts, t, fare, pas = points(n, 61)
    pts = pts.copy()
    t = t.copy()
    rng = np.random.default_rng(n)
    # equal keys: repeated points; times spanning whole days, before the
    # first bucket's start and in the last one
    rep = rng.integers(0, n, n // 5)
    pts.reshape(-1, 2)[rep] = pts.reshape(-1, 2)[rng.integers(0, n, len(rep))]
    t[: n // 10] = (T0 + rng.integers(-2 * DAY, 40 * DAY, n // 10)).astype(np.uint32)
    ids = np.arange(n, dtype=np.uint32)
    cols = (t, fare, pas, ids)
    types = COLS_TYPES + [rj.UINT]
    if time_col >= 0:
        tb = rj.Table(handler, pts, cols, types, time_col=0, time0=T0, bucket_seconds=DAY)
    else:
        tb = rj.Table(handler, pts, cols, types, organize=True)
    xy, got = tb.read()
    tb.close()
    rid = got[3]
    assert np.array_equal(np.sort(rid), ids)
    p2 = pts.reshape(-1, 2)
    assert np.array_equal(xy.reshape(-1, 2), p2[rid])
    for c, ref in zip(got[:3], (t, fare, pas)):
        assert np.array_equal(c, ref[rid])
    key = morton(p2)
    if time_col >= 0:
        tt = t.astype(np.int64)
        b = np.where(tt < T0, 0, (tt - T0) // DAY)
        b = np.minimum(b, b.max())
        key = (b.astype(np.uint64) << np.uint64(32)) | key
    k = key[rid]
    assert np.all(k[1:] >= k[:-1])
    eq = k[1:] == k[:-1]
    assert eq.any() or n < 1000
    assert np.all(rid[1:][eq] > rid[:-1][eq])
