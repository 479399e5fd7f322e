"""The reader of the reference's binary point records (rj_read_points,
Dataset::getNextRecord, db/Dataset.cpp:13-36): taxi (88-byte TaxiRecordData)
and twitter (32-byte Tweet) records, the US bounding-box filter and the
Web-Mercator projection of transformPoint (common/UsefulFuncs.cpp:578-597),
and the attribute columns of getAttributeAsBinary. Host code only: no GPU.

The projection is pinned by the textbook form of Web Mercator's latitude
term, ln(tan(pi/4 + phi/2)), which the reference writes as atanh(sin(phi));
the two agree to double rounding, so the float32 results agree to one ulp."""
import numpy as np
import pytest

import rasterjoin as rj

US = (24.396308, -124.848974, 49.384358, -66.885444)  # min lat, min lon, max lat, max lon


def mercator(lat, lon):
    def world(la, lo):
        x = (lo + 180.0) / 360.0 * 256.0 * 2.0 ** 22
        y = (1.0 - np.log(np.tan(np.pi / 4 + np.radians(la) / 2)) / np.pi) * 128.0 * 2.0 ** 22
        return x, y
    cx, cy = world((US[0] + US[2]) / 2, (US[1] + US[3]) / 2)
    x, y = world(np.float32(lat).astype(np.float64), np.float32(lon).astype(np.float64))
    return (x - cx).astype(np.float32), (y - cy).astype(np.float32)


def close(a, b):
    """Equal to within one float32 ulp."""
    return np.all(np.abs(a.astype(np.float64) - b.astype(np.float64)) <= np.spacing(np.abs(b)).astype(np.float64))


def taxi_records(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, rj.TAXI_RECORD)
    lat = rng.uniform(20.0, 52.0, n)      # some outside the US bbox
    lon = rng.uniform(-130.0, -60.0, n)
    lat[:4] = [US[0], US[2], US[0] - 1e-4, 30.0]   # on the box's edges and just outside
    lon[:4] = [US[1], US[3], -100.0, US[3] + 1e-4]
    r["pick_y"], r["pick_x"] = lat, lon
    r["drop_y"], r["drop_x"] = rng.uniform(25, 49, n), rng.uniform(-120, -70, n)
    r["pickup_time"] = rng.integers(1230768000, 1262304000, n) + (np.uint64(7) << np.uint64(40))
    r["dropoff_time"] = r["pickup_time"] + rng.integers(60, 3600, n).astype(np.uint64)
    r["duration"] = rng.integers(0, 5000, n)
    r["miles"] = rng.uniform(0, 30, n).astype(np.float32)
    r["fare"] = rng.integers(250, 9000, n)
    r["passengers"] = rng.integers(1, 7, n)
    r["vendor"] = b"VTS"
    return r


def test_record_layouts_are_the_references():
    assert rj.TAXI_RECORD.itemsize == 88 and rj.TWEET_RECORD.itemsize == 32
    assert rj.TAXI_RECORD.fields["fare"][1] == 68 and rj.TAXI_RECORD.fields["passengers"][1] == 86


def test_taxi_records(tmp_path):
    r = taxi_records(200_000, 1)
    path = str(tmp_path / "taxi.bin")
    r.tofile(path)
    xy, cols, nrec = rj.read_points(path, rj.DATASET_TAXI, 1, (3, 4, 6, 7, 8, 18))
    assert nrec == len(r)
    lat, lon = r["pick_y"].astype(np.float32), r["pick_x"].astype(np.float32)
    # the float coordinates compare with the double bounds in double (PointF
    # holds floats, transformPoint's bounds are doubles)
    la, lo = lat.astype(np.float64), lon.astype(np.float64)
    keep = (la >= US[0]) & (la <= US[2]) & (lo >= US[1]) & (lo <= US[3])
    assert not keep[2:4].any() and 0 < keep.sum() < len(r)
    k = r[keep]
    assert len(xy) == 2 * keep.sum()
    x, y = mercator(lat[keep], lon[keep])
    assert close(xy[0::2], x) and close(xy[1::2], y)
    t, tdrop, dur, miles, fare, pas = cols
    assert t.dtype == np.uint32 and np.array_equal(t, (k["pickup_time"] & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    assert np.array_equal(tdrop, (k["dropoff_time"] & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    assert np.array_equal(dur, k["duration"]) and miles.dtype == np.float32 and np.array_equal(miles, k["miles"])
    assert np.array_equal(fare, k["fare"].astype(np.uint32)) and np.array_equal(pas, k["passengers"].astype(np.uint32))
    # the dropoff location is passed through unprojected, as (drop_y, drop_x)
    dxy, _, _ = rj.read_points(path, rj.DATASET_TAXI, 2)
    assert np.array_equal(dxy[0::2], k["drop_y"].astype(np.float32))
    assert np.array_equal(dxy[1::2], k["drop_x"].astype(np.float32))
    # at most max_records records
    xy5, _, n5 = rj.read_points(path, rj.DATASET_TAXI, 1, (), max_records=1000)
    assert n5 == 1000 and len(xy5) == 2 * keep[:1000].sum()


def test_twitter_records(tmp_path):
    rng = np.random.default_rng(2)
    n = 50_000
    r = np.zeros(n, rj.TWEET_RECORD)
    r["loc_x"] = rng.uniform(20, 52, n).astype(np.float32)   # lat
    r["loc_y"] = rng.uniform(-130, -60, n).astype(np.float32)  # lon
    r["time"] = rng.integers(1230768000, 1262304000, n)
    r["favCt"] = rng.integers(-5, 1000, n)
    r["rtCt"] = rng.integers(0, 100, n)
    path = str(tmp_path / "tweets.bin")
    r.tofile(path)
    xy, (t, fav, rt), nrec = rj.read_points(path, rj.DATASET_TWITTER, 2, (1, 3, 4))
    la, lo = r["loc_x"].astype(np.float64), r["loc_y"].astype(np.float64)
    keep = (la >= US[0]) & (la <= US[2]) & (lo >= US[1]) & (lo <= US[3])
    x, y = mercator(r["loc_x"][keep], r["loc_y"][keep])
    assert nrec == n and close(xy[0::2], x) and close(xy[1::2], y)
    assert np.array_equal(t, r["time"][keep].astype(np.uint32)) and fav.dtype == np.int32
    assert np.array_equal(fav, r["favCt"][keep]) and np.array_equal(rt, r["rtCt"][keep])


def test_reader_rejects_bad_arguments(tmp_path):
    path = str(tmp_path / "taxi.bin")
    taxi_records(10, 3).tofile(path)
    with pytest.raises(rj.RJError) as e:
        rj.read_points(path, rj.DATASET_TAXI, 1, (9,))   # a 2-byte attribute
    assert e.value.code == rj.ERR_UNSUPPORTED
    with pytest.raises(rj.RJError) as e:
        rj.read_points(path, rj.DATASET_TWITTER, 1)       # twitter's location is attribute 2
    assert e.value.code == rj.ERR_ARG
    with pytest.raises(rj.RJError) as e:
        rj.read_points(str(tmp_path / "missing.bin"), rj.DATASET_TAXI, 1)
    assert e.value.code == rj.ERR_ARG
