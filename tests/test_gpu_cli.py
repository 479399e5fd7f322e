"""The --joinType driver (RasterJoin/main.cpp) end to end: polygons read
from the reference's text format through a polygon-list file (with and
without the Web-Mercator projection of lat/lon vertices), the synthetic
point table, the per-polygon CSVs (main.cpp:178-225) and the timing line
(GLHandler.cpp:214-248) against the same joins through the library API."""
import os
import subprocess

import numpy as np
import pytest

import rasterjoin as rj
from common import T0

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "raster-join_amd", "bin", "raster_join")
N = 300_000

# polygons as the text format holds them: a square given with its closing
# vertex and a repeated vertex, a triangle, and a two-part polygon
POLYS = [
    [[(10000, 10000), (30000, 10000), (30000, 10000), (30000, 30000), (10000, 30000), (10000, 10000)]],
    [[(31000, 5000), (45000, 12000), (33000, 28000)]],
    [[(5000, 32000), (20000, 32000), (20000, 40000)], [(20000, 47000), (5000, 47000)]],
]


def write_polys(d):
    lines = [str(len(POLYS))]
    for p in POLYS:
        lines.append(str(len(p)))
        for part in p:
            lines.append(str(len(part)))
            lines += [f"{x} {y}" for x, y in part]
    with open(os.path.join(d, "polys.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(d, "list.csv"), "w") as f:
        f.write("other,missing.txt\ntest,polys.txt\n")


def api_polys():
    """The same polygons after the reader's clean-up: consecutive duplicates
    and the closing vertex dropped, parts concatenated."""
    xy, off = [], [0]
    for p in POLYS:
        v = []
        for part in p:
            for k, pt in enumerate(part):
                if k > 0 and v and v[-1] == pt:
                    continue
                v.append(pt)
        if v[0] == v[-1]:
            v.pop()
        xy += [c for pt in v for c in pt]
        off.append(len(xy) // 2)
    return np.array(xy, np.float32), np.array(off, np.int32)


def run_cli(tmp, *flags):
    out = os.path.join(tmp, "out")
    os.makedirs(out, exist_ok=True)
    cmd = [CLI, "--polygonList", os.path.join(tmp, "list.csv"), "--polygonDataset", "test",
           "--inputSize", str(N), "--opAggregation", out, "--outputTime", os.path.join(tmp, "time.txt"),
           "--nIter", "2", *flags]
    print("cli:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return out, r.stdout


def read_csv(path):
    rows = [l.split("\t") for l in open(path).read().split("\n") if l]
    return rows


def api_handler(cons=()):
    pts, t, fare, pas = rj.gen_points(N, params=rj.point_params(seed=7))
    h = rj.Handler()
    xy, off = api_polys()
    h.set_polygons(xy, off)
    h.set_points(pts, (t, fare, pas), (rj.UINT, rj.FLOAT, rj.UINT8))
    h.set_query_constraints(cons)
    return h


@pytest.mark.skipif(not os.path.exists(CLI), reason="driver not built")
def test_cli_index_count_and_timing_line(tmp_path):
    d = str(tmp_path)
    write_polys(d)
    out, stdout = run_cli(d, "--joinType", "index", "--indexRes", "128")
    rows = read_csv(os.path.join(out, "raster_0_0.csv"))
    h = api_handler()
    h.set_aggregation(rj.COUNT, -1)
    h.set_poly_index_resolution(128)
    c, _ = h.execute_function(rj.INDEX)
    assert [int(r[0]) for r in rows] == list(range(len(POLYS)))
    assert np.array_equal(np.array([int(r[1]) for r in rows]), c)
    assert c.min() > 0
    line = open(os.path.join(d, "time.txt")).read().strip().split("\n")[-1].split()
    ref = h.print_time_stats(rj.INDEX).split()
    assert len(line) == len(ref) and line[:5] == ref[:5]
    assert line[0] == str(N) and line[1] == str(len(POLYS))
    h.close()


@pytest.mark.skipif(not os.path.exists(CLI), reason="driver not built")
def test_cli_raster_avg_with_time_window(tmp_path):
    d = str(tmp_path)
    write_polys(d)
    t_lo, t_hi = T0 + 1_000_000, T0 + 20_000_000
    out, _ = run_cli(d, "--joinType", "raster", "--accuracy", "20", "--avg", "8",
                     "--startTime", str(t_lo), "--endTime", str(t_hi))
    rows = read_csv(os.path.join(out, f"raster_{t_hi}_20.csv"))
    h = api_handler([(0, rj.GTE, t_lo), (0, rj.LTE, t_hi)])
    h.set_aggregation(rj.AVG, 1)
    h.set_accuracy_distance(20.0)
    c, s = h.execute_function(rj.RASTER)
    avg = np.where(c > 0, s / np.maximum(c, 1), 0.0)
    got = np.array([float(r[1]) for r in rows])
    assert np.allclose(got, avg, rtol=1e-9, atol=0)
    h.close()


@pytest.mark.skipif(not os.path.exists(CLI), reason="driver not built")
def test_cli_errorbounds_csv(tmp_path):
    d = str(tmp_path)
    write_polys(d)
    out, _ = run_cli(d, "--joinType", "errorbounds", "--accuracy", "50")
    rows = read_csv(os.path.join(out, "err_bound_0_50.csv"))
    h = api_handler()
    h.set_aggregation(rj.COUNT, -1)
    h.set_accuracy_distance(50.0)
    c, _ = h.execute_function(rj.ERRORBOUNDS)
    b = h.get_error_bounds()
    np_ = len(POLYS)
    fn_, fp, fn1, fp1 = b[:np_], b[np_:2 * np_], b[2 * np_:3 * np_], b[3 * np_:]
    exp = np.stack([np.arange(np_), c - fp1, c + fn1, c - fp, c + fn_], 1)
    assert np.array_equal(np.array([[int(x) for x in r] for r in rows]), exp)
    # the guaranteed range holds the exact count
    exact, _ = h.execute_function(rj.INDEX)
    assert np.all(exp[:, 1] <= exact) and np.all(exact <= exp[:, 2])
    h.close()


@pytest.mark.skipif(not os.path.exists(CLI), reason="driver not built")
def test_cli_rejects_unknown_join_type(tmp_path):
    r = subprocess.run([CLI, "--joinType", "bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Wrong join type" in r.stderr


def mercator(lat, lon):
    """Web Mercator world coordinates at zoom 22 with the textbook form of
    the latitude term, ln(tan(pi/4 + phi/2)), minus those of the US bounding
    box's center (the reader's transformPoint, UsefulFuncs.cpp:578-612)."""
    def world(la, lo):
        x = (lo + 180.0) / 360.0 * 256.0 * 2.0 ** 22
        y = (1.0 - np.log(np.tan(np.pi / 4 + np.radians(la) / 2)) / np.pi) * 128.0 * 2.0 ** 22
        return x, y
    cx, cy = world((24.396308 + 49.384358) / 2, (-124.848974 + -66.885444) / 2)
    x, y = world(np.float64(np.float32(lat)), np.float64(np.float32(lon)))
    return np.float32(x - cx), np.float32(y - cy)


@pytest.mark.skipif(not os.path.exists(CLI), reason="driver not built")
def test_cli_web_mercator_polygons(tmp_path):
    """--polyTransform reads `lat lon` vertices, projects them, and skips a
    polygon that leaves the US bounding box: lat/lon rectangles become
    axis-aligned rectangles whose counts follow from the points directly."""
    d = str(tmp_path)
    clat, clon = (24.396308 + 49.384358) / 2, (-124.848974 + -66.885444) / 2
    rects = [(0.002, 0.006, 0.002, 0.007), (0.008, 0.0125, 0.004, 0.009), (0.001, 0.004, 0.009, 0.015)]
    lines = [str(len(rects) + 1)]
    boxes = []
    for a, b, c, e in rects:  # lat below the center (y grows southwards), lon east of it
        la0, la1, lo0, lo1 = clat - b, clat - a, clon + c, clon + e
        lines += ["1", "4", f"{la0!r} {lo0!r}", f"{la0!r} {lo1!r}", f"{la1!r} {lo1!r}", f"{la1!r} {lo0!r}"]
        x0, y1 = mercator(la0, lo0)
        x1, y0 = mercator(la1, lo1)
        boxes.append((x0, x1, y0, y1))
        if len(boxes) == 2:  # a polygon with a vertex outside the US bbox: skipped
            lines += ["1", "3", f"{clat!r} {clon!r}", "10.0 -90.0", f"{clat + 0.01!r} {clon!r}"]
    with open(os.path.join(d, "polys.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(d, "list.csv"), "w") as f:
        f.write("test,polys.txt\n")
    out, _ = run_cli(d, "--joinType", "index", "--indexRes", "64", "--polyTransform")
    rows = read_csv(os.path.join(out, "raster_0_0.csv"))
    assert [int(r[0]) for r in rows] == list(range(len(rects)))
    pts = rj.gen_points(N, params=rj.point_params(seed=7))[0]
    x, y = pts[0::2], pts[1::2]
    exp = []
    for x0, x1, y0, y1 in boxes:
        assert 0 < x0 < x1 < 50000 and 0 < y0 < y1 < 50000
        # no point within 0.01 m of an edge: the count does not depend on rounding
        near = (np.minimum(abs(x - x0), abs(x - x1)) < 0.01) | (np.minimum(abs(y - y0), abs(y - y1)) < 0.01)
        assert not near.any()
        exp.append(int(((x > x0) & (x < x1) & (y > y0) & (y < y1)).sum()))
    assert [int(r[1]) for r in rows] == exp and min(exp) > 0


@pytest.mark.skipif(not os.path.exists(CLI), reason="driver not built")
def test_cli_taxi_record_file(tmp_path):
    """--pointFile with taxi records (lat/lon pickups, as the reference's
    datasets hold them): records outside the US bounding box are skipped, the
    pickups projected like the polygons; an index join with a pickup-time
    window, the passengers == 1 filter (--nAttrib 1, main.cpp:85-91) and
    avg(fare) gives the counts and means of the records inside each lat/lon
    rectangle directly."""
    d = str(tmp_path)
    clat, clon = (24.396308 + 49.384358) / 2, (-124.848974 + -66.885444) / 2
    rects = [(0.002, 0.006, 0.002, 0.007), (0.008, 0.0125, 0.004, 0.009), (0.001, 0.004, 0.009, 0.015)]
    lines = [str(len(rects))]
    for a, b, c, e in rects:
        la0, la1, lo0, lo1 = clat - b, clat - a, clon + c, clon + e
        lines += ["1", "4", f"{la0!r} {lo0!r}", f"{la0!r} {lo1!r}", f"{la1!r} {lo1!r}", f"{la1!r} {lo0!r}"]
    with open(os.path.join(d, "polys.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(d, "list.csv"), "w") as f:
        f.write("test,polys.txt\n")
    rng = np.random.default_rng(9)
    n = N  # run_cli reads at most --inputSize N records
    lat = clat - rng.uniform(0.0, 0.014, n)
    lon = clon + rng.uniform(0.0, 0.017, n)
    # no record within 1e-5 degrees of an edge line: rounding cannot move one across
    m = 1e-5
    la, lo = lat.astype(np.float32).astype(np.float64), lon.astype(np.float32).astype(np.float64)
    near = np.zeros(n, bool)
    for a, b, c, e in rects:
        for v in (clat - b, clat - a):
            near |= abs(la - v) < m
        for v in (clon + c, clon + e):
            near |= abs(lo - v) < m
    lat, lon, la, lo = lat[~near], lon[~near], la[~near], lo[~near]
    n = len(lat)
    r = np.zeros(n, rj.TAXI_RECORD)
    lat[:1000], lon[:1000] = 10.0, -100.0   # outside the US bbox: skipped
    la[:1000], lo[:1000] = 10.0, -100.0
    r["pick_y"], r["pick_x"] = lat, lon
    t0 = 1230768000
    r["pickup_time"] = rng.integers(t0, t0 + 31536000, n)
    r["fare"] = rng.integers(250, 9000, n)
    r["passengers"] = rng.integers(1, 7, n)
    path = os.path.join(d, "taxi.bin")
    r.tofile(path)
    lo_t, hi_t = t0 + 1_000_000, t0 + 20_000_000
    t = r["pickup_time"]
    sel = (t >= lo_t) & (t <= hi_t) & (r["passengers"] == 1)
    exp_c, exp_avg = [], []
    for a, b, c, e in rects:
        la0, la1, lo0, lo1 = clat - b, clat - a, clon + c, clon + e
        k = (la > la0) & (la < la1) & (lo > lo0) & (lo < lo1) & sel
        exp_c.append(int(k.sum()))
        exp_avg.append(r["fare"][k].astype(np.float64).sum() / max(1, k.sum()))
    assert min(exp_c) > 100
    flags = ["--joinType", "index", "--indexRes", "128", "--polyTransform", "--pointFile", path,
             "--pointDataset", "taxi", "--locAttrib", "1", "--startTime", str(lo_t), "--endTime", str(hi_t),
             "--nAttrib", "1"]
    out, _ = run_cli(d, *flags)
    rows = read_csv(os.path.join(out, f"raster_{hi_t}_0.csv"))
    assert [int(x[1]) for x in rows] == exp_c
    out, _ = run_cli(d, *flags, "--avg", "8")
    rows = read_csv(os.path.join(out, f"raster_{hi_t}_0.csv"))
    assert np.allclose([float(x[1]) for x in rows], exp_avg, rtol=1e-9, atol=0)
    xy, _, nrec = rj.read_points(path, rj.DATASET_TAXI, 1)
    assert nrec == n and len(xy) == 2 * (n - 1000)
