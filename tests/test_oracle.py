"""Pins the C oracle (oracle/rj_oracle.c) before it is trusted as the checker:
against the reference's own compiled GridIndexF (golden fixtures and, when the
reference tree is present, live), against brute force, closed forms and the
invariants the raster rules must satisfy. CPU only."""
import hashlib
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import pyoracle as O
import rasterjoin as rj

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


# ---------------------------------------------------------------------------
# CPUJoin restatement vs the reference's GridIndexF
# ---------------------------------------------------------------------------

def test_toy_closed_form():
    """baseline.hpp:139-217: square [5.1,5.2]^2 holds the 2 points at
    (5.15,5.15), square [4.9,5.1]^2 the 3 points at (5,5)."""
    d = np.load(os.path.join(GOLD, "toy_squares.npz"))
    g = O.Grid(d["xy"], d["off"], 4)
    counts, _, _, _ = O.join_exact(g, O.PointSet(d["pts"]))
    assert counts.tolist() == [2, 3]


@pytest.mark.parametrize("name", ["toy_squares", "squares_edges", "voronoi20"])
def test_golden_fixture(name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    g = O.Grid(d["xy"], d["off"], int(d["res"]))
    cell_off, cand = g.export()
    assert np.array_equal(cell_off, d["cell_off"])
    assert np.array_equal(cand, d["cand"])
    counts, _, _, _ = O.join_exact(g, O.PointSet(d["pts"]))
    assert np.array_equal(counts, d["counts"])


def test_index_1M_voronoi260_golden():
    """1M points, 260 Voronoi polygons, grid 256: the reference's counts."""
    d = np.load(os.path.join(GOLD, "index_1M_voronoi260.npz"))
    cp = rj.POLY_SETS["voronoi260"]
    xy, off = rj.gen_voronoi(cp["nsites"], seed=cp["seed"], spacing=cp["spacing"], jitter=cp["jitter"])
    pts, _, _, _ = rj.gen_points(int(d["npts"]), params=rj.point_params(seed=int(d["point_seed"])),
                                 columns=False)
    assert sha(xy, off) == str(d["poly_sha"]), "polygon generator changed"
    assert sha(pts) == str(d["pts_sha"]), "point generator changed"
    g = O.Grid(xy, off, int(d["res"]))
    cell_off, cand = g.export()
    assert len(cand) == int(d["ncand"])
    assert sha(cell_off, cand) == str(d["cand_sha"])
    counts, _, _, _ = O.join_exact(g, O.PointSet(pts))
    assert np.array_equal(counts, d["counts"])


def star_polygons(n, seed, box=1000.0, rmin=20.0, rmax=200.0, vmin=3, vmax=40):
    """Random star-shaped polygons (they overlap each other)."""
    rng = np.random.default_rng(seed)
    xy, off = [], [0]
    for _ in range(n):
        k = int(rng.integers(vmin, vmax + 1))
        cx, cy = rng.uniform(rmax, box - rmax, 2)
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        r = rng.uniform(rmin, rmax, k)
        pts = np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], 1).astype(np.float32)
        keep = [0] + [i for i in range(1, k) if not np.array_equal(pts[i], pts[i - 1])]
        pts = pts[keep]
        xy.append(pts.ravel())
        off.append(off[-1] + len(pts))
    return np.concatenate(xy), np.array(off, np.int32)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("case", ["voronoi", "stars", "tracts"])
def test_oracle_matches_reference_live(case, tmp_path):
    if case == "voronoi":
        xy, off = rj.gen_voronoi(60, seed=5, spacing=(120.0, 120.0))
        res, npts = 128, 200_000
        pts, _, _, _ = rj.gen_points(npts, params=rj.point_params(seed=9), columns=False)
    elif case == "tracts":
        xy, off = rj.gen_voronoi(150, seed=8, spacing=(3.0, 140.0))
        res, npts = 300, 100_000
        pts, _, _, _ = rj.gen_points(npts, params=rj.point_params(seed=10), columns=False)
    else:
        xy, off = star_polygons(40, seed=3)
        res, npts = 50, 100_000
        rng = np.random.default_rng(4)
        pts = rng.uniform(-10, 1010, 2 * npts).astype(np.float32)
    ref = O.RefGrid(xy, off, res, tmp_path)
    ours = O.Grid(xy, off, res)
    rc, rcand = ref.export()
    oc, ocand = ours.export()
    assert np.array_equal(rc, oc) and np.array_equal(rcand, ocand)
    counts, _, _, _ = O.join_exact(ours, O.PointSet(pts))
    assert np.array_equal(counts, ref.join_count(pts))


def test_grid_join_equals_brute_force():
    xy, off = rj.gen_voronoi(30, seed=2, spacing=(200.0, 200.0))
    pts, _, _, _ = rj.gen_points(50_000, params=rj.point_params(seed=3), columns=False)
    g = O.Grid(xy, off, 64)
    counts, _, _, _ = O.join_exact(g, O.PointSet(pts))
    assert np.array_equal(counts, O.join_brute(xy, off, O.PointSet(pts)))


# ---------------------------------------------------------------------------
# filters and exact sums
# ---------------------------------------------------------------------------

def filter_mask(cols, types, cons):
    m = np.ones(len(cols[0]), bool)
    for col, op, val in cons:
        a = cols[col]
        if types[col] == rj.FLOAT:
            v = np.float32(val)
        elif types[col] == rj.INT:
            v = np.int32(val)
        else:
            v = np.uint32(val)
            a = a.astype(np.uint32)
        m &= [a == v, a < v, a <= v, a > v, a >= v][op]
    return m


def test_filters_match_numpy_predicate():
    xy, off = rj.gen_voronoi(12, seed=4, spacing=(800.0, 800.0))
    pts, t, fare, pas = rj.gen_points(40_000, params=rj.point_params(seed=5))
    rng = np.random.default_rng(0)
    ival = rng.integers(-50, 50, len(t)).astype(np.int32)
    cols = [t, fare, pas, ival]
    types = [rj.UINT, rj.FLOAT, rj.UINT8, rj.INT]
    g = O.Grid(xy, off, 32)
    cases = [
        [(0, rj.GTE, 1240000000), (0, rj.LTE, 1250000000), (2, rj.LTE, 2)],
        [(1, rj.GT, 12.5)], [(1, rj.LTE, float(fare[7]))], [(2, rj.EQ, 1)], [(2, rj.LT, 1)],
        [(3, rj.LT, 0)], [(3, rj.GTE, -3), (3, rj.LTE, 3)], [(0, rj.GT, int(t.max()))],
    ]
    for cons in cases:
        m = filter_mask(cols, types, cons)
        got, _, _, _ = O.join_exact(g, O.PointSet(pts, cols, types, cons))
        want, _, _, _ = O.join_exact(g, O.PointSet(pts.reshape(-1, 2)[m].ravel()))
        assert np.array_equal(got, want), cons


def test_fixed_point_sums_exact_and_close_to_f64():
    xy, off = rj.gen_voronoi(8, seed=6, spacing=(1000.0, 1000.0))
    pts, t, fare, pas = rj.gen_points(20_000, params=rj.point_params(seed=8))
    g = O.Grid(xy, off, 16)
    ps = O.PointSet(pts, [t, fare, pas], [rj.UINT, rj.FLOAT, rj.UINT8], (), agg_type=rj.AVG, agg_col=1)
    counts, hi, lo, f64 = O.join_exact(g, ps)
    # exact: hi + lo/2^32 equals the rational sum of the float values
    pxy = pts.reshape(-1, 2)
    for p in range(len(counts)):
        v = pxy.shape[0]
        sel = [i for i in range(v) if O.join_brute(xy[2 * off[p]:2 * off[p + 1]],
                                                   np.array([0, off[p + 1] - off[p]], np.int32),
                                                   O.PointSet(pxy[i])).item()] if p == 0 else None
        if sel is not None:
            exact = sum(Fraction(float(fare[i])) for i in sel)
            assert Fraction(int(hi[p])) + Fraction(int(lo[p]), 2 ** 32) == exact
        s = float(hi[p]) + float(lo[p]) / 2 ** 32
        assert abs(s - f64[p]) <= 1e-9 * max(1.0, abs(f64[p]))
    # integer columns: sums are exact integers
    ps2 = O.PointSet(pts, [t, fare, pas], [rj.UINT, rj.FLOAT, rj.UINT8], (), agg_type=rj.AVG, agg_col=2)
    c2, hi2, lo2, f2 = O.join_exact(g, ps2)
    assert np.all(lo2 == 0) and np.array_equal(hi2.astype(np.float64), f2)


# ---------------------------------------------------------------------------
# raster rules: coverage, pixel formula, boundary, bounds
# ---------------------------------------------------------------------------

def small_plan(W):
    return O.plan(np.array([0.0, 0.0, 1000.0, 1000.0]), 1000.0 * np.sqrt(2) / W)


@pytest.mark.parametrize("seed", range(6))
def test_edge_flag_coverage_equals_per_pixel_crossing(seed):
    p = small_plan(64)
    xy, off = star_polygons(5, seed=seed, box=1000.0, rmin=30, rmax=400, vmin=3, vmax=60)
    rng = np.random.default_rng(seed)
    # also a polygon with vertices on pixel centers and horizontal edges
    c = (np.array([3, 3, 30, 3, 30, 20, 10, 20, 10, 40, 3, 40], np.float64) + 0.5) * 1000.0 / 64
    xy = np.concatenate([xy, c.astype(np.float32)])
    off = np.append(off, off[-1] + 6).astype(np.int32)
    for k in range(len(off) - 1):
        v = xy[2 * off[k]:2 * off[k + 1]]
        assert np.array_equal(O.coverage(p, v, 0), O.coverage(p, v, 1))
        assert np.array_equal(O.coverage(p, v, 2), O.coverage(p, v, 3))
    del rng


def test_tessellation_covers_each_pixel_center_once():
    xy, off = rj.gen_voronoi(25, seed=12, spacing=(300.0, 300.0))
    p = O.plan(O.poly_bbox(xy).astype(np.float64), 400.0)
    tot = np.zeros((p.H, p.W), np.int32)
    for k in range(len(off) - 1):
        tot += O.coverage(p, xy[2 * off[k]:2 * off[k + 1]], 1)
    # shared chains snap identically: no pixel center in two cells; the only
    # uncovered centers are on the bbox max edge side
    assert tot.max() == 1
    assert (tot == 0).sum() <= p.W + p.H


def test_pixel_formula_and_plan():
    bbox = np.array([0.0, 0.0, 50000.0, 50000.0])
    p = O.plan(bbox, 10.0)
    assert (p.W, p.H) == (7072, 7072)          # ceil(50000 / (10 / sqrt(2))) pixels a side
    p5 = O.plan(bbox, 5.0)
    assert (p5.W, p5.splitx) == (14144, 2)     # two 7072 tiles per axis
    assert O.pixel(p, 0.0, 0.0) == 0
    assert O.pixel(p, np.float32(49999.996), 1.0) == p.W - 1
    assert O.pixel(p, 50000.0, 1.0) == -1      # max edge is outside
    assert O.pixel(p, -0.001, 1.0) == -1


def test_pixel_aligned_squares_raster_is_exact():
    """Squares on pixel boundaries: pixel-center coverage = exact membership."""
    p = small_plan(100)  # 10-unit pixels
    squares = []
    for (i0, j0, i1, j1) in [(0, 0, 10, 10), (10, 0, 35, 10), (5, 20, 45, 70), (50, 50, 100, 100)]:
        squares.append(np.array([i0, j0, i1, j0, i1, j1, i0, j1], np.float32) * 10)
    xy = np.concatenate(squares)
    off = np.arange(0, 4 * len(squares) + 1, 4, dtype=np.int32)
    rng = np.random.default_rng(1)
    pts = rng.uniform(0, 1000, (200_000, 2)).astype(np.float32)
    on_grid = np.any(np.abs(pts / 10 - np.round(pts / 10)) < 1e-3, axis=1)
    pts = pts[~on_grid].ravel()
    counts, _, _ = O.raster_join(p, xy, off, O.PointSet(pts))
    exact, _, _, _ = O.join_exact(O.Grid(xy, off, 16, bbox=[0, 0, 1000, 1000]), O.PointSet(pts))
    assert np.array_equal(counts, exact)


def np_pip(poly, x, y):
    c = np.zeros(len(x), bool)
    n = len(poly)
    for i in range(n):
        xi, yi = poly[i]
        xj, yj = poly[i - 1]
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = (xj - xi) * (y - yi) / (yj - yi) + xi
        c ^= ((yi > y) != (yj > y)) & (x < xint)
    return c


def seg_dist(poly, x, y):
    d = np.full(len(x), np.inf)
    for i in range(len(poly)):
        a, b = poly[i - 1], poly[i]
        ab = b - a
        t = np.clip(((x - a[0]) * ab[0] + (y - a[1]) * ab[1]) / max(ab @ ab, 1e-30), 0, 1)
        d = np.minimum(d, np.hypot(x - (a[0] + t * ab[0]), y - (a[1] + t * ab[1])))
    return d


def test_raster_misassignments_lie_within_eps_of_a_boundary():
    eps = 150.0
    xy, off = rj.gen_voronoi(15, seed=31, spacing=(400.0, 400.0))
    p = O.plan(O.poly_bbox(xy).astype(np.float64), eps)
    pts, _, _, _ = rj.gen_points(30_000, params=rj.point_params(seed=32), columns=False)
    P = pts.reshape(-1, 2).astype(np.float64)
    pix = np.array([O.pixel(p, float(a), float(b)) for a, b in pts.reshape(-1, 2)])
    for k in range(len(off) - 1):
        v = xy[2 * off[k]:2 * off[k + 1]]
        cov = O.coverage(p, v, 1).ravel()
        raster_in = (pix >= 0) & (cov[np.maximum(pix, 0)] == 1)
        poly = v.reshape(-1, 2).astype(np.float64)
        exact_in = np_pip(poly, P[:, 0], P[:, 1])
        bad = raster_in != exact_in
        if bad.any():
            assert seg_dist(poly, P[bad, 0], P[bad, 1]).max() <= eps * (1 + 1e-6)


def touched_exact(X, Y, W, H):
    """Pixels whose closed square meets the snapped segment, by exact
    rational geometry (separating axes)."""
    out = set()
    for (ax, ay), (bx, by) in zip(zip(X, Y), zip(np.roll(X, 1), np.roll(Y, 1))):
        for j in range(max(0, min(ay, by) // 256 - 1), min(H, max(ay, by) // 256 + 2)):
            for i in range(max(0, min(ax, bx) // 256 - 1), min(W, max(ax, bx) // 256 + 2)):
                x0, x1, y0, y1 = 256 * i, 256 * (i + 1), 256 * j, 256 * (j + 1)
                if max(ax, bx) < x0 or min(ax, bx) > x1 or max(ay, by) < y0 or min(ay, by) > y1:
                    continue
                s = [(bx - ax) * (cy - ay) - (by - ay) * (cx - ax)
                     for cx, cy in ((x0, y0), (x1, y0), (x0, y1), (x1, y1))]
                if min(s) <= 0 <= max(s):
                    out.add((j, i))
    return out


@pytest.mark.parametrize("seed", range(3))
def test_boundary_rule_is_a_tight_superset_of_touched_pixels(seed):
    p = small_plan(48)
    xy, off = star_polygons(3, seed=seed + 10, box=1000.0, rmin=50, rmax=400)
    for k in range(len(off) - 1):
        v = xy[2 * off[k]:2 * off[k + 1]]
        X, Y = O.snap(p, v)
        exact = touched_exact(X.tolist(), Y.tolist(), p.W, p.H)
        got0 = O.boundary(p, v, 0)
        got = {tuple(a) for a in np.argwhere(got0)}
        assert exact <= got
        ex = np.zeros((p.H, p.W), np.uint8)
        for (j, i) in exact:
            ex[max(0, j - 1):j + 2, max(0, i - 1):i + 2] = 1
        assert np.all(ex[got0 == 1] == 1)        # within one pixel of exact
        # dilation radius 1 of the rule
        got1 = O.boundary(p, v, 1)
        dil = np.zeros_like(got0)
        for (j, i) in np.argwhere(got0):
            dil[max(0, j - 1):j + 2, max(0, i - 1):i + 2] = 1
        assert np.array_equal(dil, got1)


@pytest.mark.parametrize("eps", [20.0, 100.0])
def test_errorbounds_contain_exact_counts(eps):
    cp = rj.POLY_SETS["voronoi260"]
    xy, off = rj.gen_voronoi(cp["nsites"], seed=cp["seed"], spacing=cp["spacing"], jitter=cp["jitter"])
    pts, _, _, _ = rj.gen_points(300_000, params=rj.point_params(seed=13), columns=False)
    p = O.plan(O.poly_bbox(xy).astype(np.float64), eps)
    counts, _, _, canvas = O.raster_join(p, xy, off, O.PointSet(pts), want_canvas=True)
    b = O.errorbounds(p, xy, off, canvas)
    np_ = len(off) - 1
    fn, fp, fn1, fp1 = b[:np_], b[np_:2 * np_], b[2 * np_:3 * np_], b[3 * np_:]
    exact, _, _, _ = O.join_exact(O.Grid(xy, off, 256), O.PointSet(pts))
    assert np.all(counts - fp1 <= exact) and np.all(exact <= counts + fn1)
    assert np.all(fn <= fn1) and np.all(fp <= fp1)
    assert (counts != exact).sum() > 0     # the raster is approximate at this eps


def test_errorbounds_estimate_from_area_shares():
    """fn/fp: ceil(min(a, 1 - a) * count) over each polygon's boundary pixels
    outside/inside it, a = the polygon's exact area share of the pixel
    (O.area, pinned by the three tests below); fn1/fp1 the whole counts."""
    p = small_plan(40)
    xy, off = star_polygons(2, seed=41, box=1000.0, rmin=100, rmax=400)
    rng = np.random.default_rng(2)
    pts = rng.uniform(0, 1000, 2 * 50_000).astype(np.float32)
    counts, _, _, canvas = O.raster_join(p, xy, off, O.PointSet(pts), want_canvas=True)
    b = O.errorbounds(p, xy, off, canvas)
    np_ = len(off) - 1
    can = canvas.reshape(p.H, p.W).astype(np.int64)
    one = O.AREA_ONE
    for k in range(np_):
        v = xy[2 * off[k]:2 * off[k + 1]]
        cov = O.coverage(p, v, 0)
        a = np.minimum(np.abs(O.area(p, v).astype(np.int64)), one)
        m = np.minimum(a, one - a)
        add = (m * can + one - 1) >> 24
        bnd = O.boundary(p, v, 1)
        sel_in = (bnd == 1) & (cov == 1)
        sel_out = (bnd == 1) & (cov == 0)
        assert (add[sel_out] > 0).any() and (add[sel_in] > 0).any()
        assert b[k] == add[sel_out].sum()
        assert b[np_ + k] == add[sel_in].sum()
        assert b[2 * np_ + k] == can[sel_out].sum()
        assert b[3 * np_ + k] == can[sel_in].sum()
        assert counts[k] == can[cov == 1].sum()


def shoelace2(X, Y):
    """Twice the signed area of an integer ring (exact)."""
    return sum(int(X[i - 1]) * int(Y[i]) - int(X[i]) * int(Y[i - 1]) for i in range(len(X)))


def row_pieces(X, Y, j):
    """Edges whose open y-range meets row j's [256j, 256j + 256]."""
    n = 0
    for i in range(len(X)):
        y0, y1 = sorted((int(Y[i - 1]), int(Y[i])))
        if y0 != y1 and y0 < 256 * j + 256 and y1 > 256 * j:
            n += 1
    return n


@pytest.mark.parametrize("seed", range(4))
def test_area_shares_sum_to_the_ring_area(seed):
    """Summed over the canvas, the per-pixel areas give the snapped ring's
    shoelace area (each (edge, row) piece floors at most one unit per
    column it spans, none right of it); covered pixels far from the boundary
    hold exactly one pixel, outside ones exactly zero."""
    p = small_plan(96)
    xy, off = star_polygons(3, seed=50 + seed, box=1000.0, rmin=40, rmax=420, vmin=3, vmax=60)
    for k in range(len(off) - 1):
        v = xy[2 * off[k]:2 * off[k + 1]]
        X, Y = O.snap(p, v)
        a = O.area(p, v).astype(np.int64)
        want = shoelace2(X, Y) * 128          # (2A / 2) sub^2 x 256 units
        spans = sum(abs(int(X[i]) - int(X[i - 1])) // 256 + 2 for i in range(len(X)))
        assert abs(int(a.sum()) - want) <= spans
        assert want > 0 and a.min() > -spans  # counter-clockwise ring: positive shares
        cov = O.coverage(p, v, 1)
        far = O.boundary(p, v, 1) == 0
        assert np.all(a[far & (cov == 1)] == O.AREA_ONE)
        assert np.all(a[far & (cov == 0)] == 0)


def clip_area(poly, x0, y0, x1, y1):
    """Exact area (Fraction) of a simple ring inside the box, by
    Sutherland-Hodgman clipping against the box's four sides and the
    shoelace formula."""
    def clip(pts, inside, cut):
        out = []
        for i in range(len(pts)):
            cur, prv = pts[i], pts[i - 1]
            if inside(cur):
                if not inside(prv):
                    out.append(cut(prv, cur))
                out.append(cur)
            elif inside(prv):
                out.append(cut(prv, cur))
        return out

    def at_x(a, b, x):
        return (x, a[1] + (b[1] - a[1]) * (x - a[0]) / (b[0] - a[0]))

    def at_y(a, b, y):
        return (a[0] + (b[0] - a[0]) * (y - a[1]) / (b[1] - a[1]), y)

    pts = poly
    for inside, cut in ((lambda q: q[0] >= x0, lambda a, b: at_x(a, b, x0)),
                        (lambda q: q[0] <= x1, lambda a, b: at_x(a, b, x1)),
                        (lambda q: q[1] >= y0, lambda a, b: at_y(a, b, y0)),
                        (lambda q: q[1] <= y1, lambda a, b: at_y(a, b, y1))):
        if not pts:
            return Fraction(0)
        pts = clip(pts, inside, cut)
    return sum(pts[i - 1][0] * pts[i][1] - pts[i][0] * pts[i - 1][1] for i in range(len(pts))) / 2


@pytest.mark.parametrize("seed", range(3))
def test_area_shares_equal_exact_clipping(seed):
    """Every pixel's area against the exact rational area of the snapped ring
    clipped to the pixel, to within one unit (2^-24 of a pixel) per edge
    crossing the pixel's row."""
    p = small_plan(14)
    xy, off = star_polygons(2, seed=70 + seed, box=1000.0, rmin=60, rmax=420, vmin=3, vmax=14)
    for k in range(len(off) - 1):
        v = xy[2 * off[k]:2 * off[k + 1]]
        X, Y = O.snap(p, v)
        ring = [(Fraction(int(x)), Fraction(int(y))) for x, y in zip(X, Y)]
        a = O.area(p, v)
        partial = 0
        for j in range(p.H):
            tol = row_pieces(X, Y, j) + 1
            for i in range(p.W):
                exact = clip_area(ring, 256 * i, 256 * j, 256 * i + 256, 256 * j + 256)
                assert abs(int(a[j, i]) - exact * 256) <= tol, (k, j, i)
                partial += 0 < exact < 65536
        assert partial >= 10


def ref_compute_fraction(p1, p2, lb, rt):
    """line-bound.frag:27-187 (ComputeOutCode, computeArea, computeFraction)
    in float32: Cohen-Sutherland clip of the segment to the pixel, then the
    area on one side of the clipped chord as a share of the pixel. Returns
    (accept, computed, frac)."""
    f = np.float32
    x0, y0, x1, y1 = f(p1[0]), f(p1[1]), f(p2[0]), f(p2[1])
    xmin, ymin, xmax, ymax = f(lb[0]), f(lb[1]), f(rt[0]), f(rt[1])

    def code(x, y):
        c = 0
        if x < xmin:
            c |= 1
        elif x > xmax:
            c |= 2
        if y < ymin:
            c |= 4
        elif y > ymax:
            c |= 8
        return c

    o0, o1 = code(x0, y0), code(x1, y1)
    while True:
        if (o0 | o1) == 0:
            break
        if o0 & o1:
            return False, False, 0.0
        oo = o0 if o0 else o1
        if oo & 8:
            x, y = x0 + (x1 - x0) * (ymax - y0) / (y1 - y0), ymax
        elif oo & 4:
            x, y = x0 + (x1 - x0) * (ymin - y0) / (y1 - y0), ymin
        elif oo & 2:
            x, y = xmax, y0 + (y1 - y0) * (xmax - x0) / (x1 - x0)
        else:
            x, y = xmin, y0 + (y1 - y0) * (xmin - x0) / (x1 - x0)
        if oo == o0:
            x0, y0, o0 = x, y, code(x, y)
        else:
            x1, y1, o1 = x, y, code(x, y)
    if x0 > x1:
        x0, x1, y0, y1 = x1, x0, y1, y0

    def tri(a, b, c):
        return abs(f(0.5) * abs((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])))

    quad = None
    if x0 == xmin:
        if x1 == xmax:
            quad = ((xmin, ymin), (xmax, ymin), (x1, y1), (x0, y0))
        elif y1 == ymax:
            t = ((xmin, ymax), (x0, y0), (x1, y1))
        elif y1 == ymin:
            t = ((xmin, ymin), (x0, y0), (x1, y1))
        else:
            return True, False, 0.5
    elif x1 == xmax:
        if y0 == ymax:
            t = ((xmax, ymax), (x0, y0), (x1, y1))
        elif y0 == ymin:
            t = ((xmax, ymin), (x1, y1), (x0, y0))
        else:
            return True, False, 0.5
    elif y0 == ymin and y1 == ymax:
        quad = ((xmin, ymin), (xmin, ymax), (x1, y1), (x0, y0))
    elif y1 == ymin and y0 == ymax:
        quad = ((xmin, ymin), (xmin, ymax), (x0, y0), (x1, y1))
    else:
        return True, False, 0.5
    if quad is not None:
        area = tri(quad[0], quad[1], quad[2]) + tri(quad[0], quad[2], quad[3])
    else:
        area = tri(*t)
    return True, True, float(area / ((ymax - ymin) * (xmax - xmin)))


@pytest.mark.parametrize("seed", range(5))
def test_area_share_of_a_one_edge_pixel_equals_the_reference_fraction(seed):
    """Where one edge crosses a pixel from side to side, min(a, 1 - a) of the
    polygon's area share equals line-bound.frag's folded computeFraction of
    that edge (its float32 arithmetic, to 5e-5)."""
    W = 64
    p = small_plan(W)
    rng = np.random.default_rng(90 + seed)
    u = 1000.0 / W
    ya, yb = rng.uniform(3 * u, (W - 3) * u, 2)
    xa, xb = 3.4 * u, (W - 3.4) * u
    if seed == 4:
        xa, xb, ya, yb = ya, yb, 3.4 * u, (W - 3.4) * u   # a steep edge
    # a simple ring whose other edges stay within 4 pixels of the canvas sides
    quad = np.array([xa, ya, xb, yb, 999.0, 999.0, 1.0 if seed != 4 else 999.0,
                     999.0 if seed != 4 else 1.0], np.float32)
    a = np.abs(O.area(p, quad).astype(np.int64))
    X, Y = O.snap(p, quad)
    e0, e1 = (X[0] / 256.0, Y[0] / 256.0), (X[1] / 256.0, Y[1] / 256.0)
    one = O.AREA_ONE
    checked = 0
    for j in range(5, W - 5):
        for i in range(5, W - 5):
            acc, comp, fr = ref_compute_fraction(e0, e1, (i, j), (i + 1, j + 1))
            if not (acc and comp):
                continue
            share = min(a[j, i], one - a[j, i]) / one
            assert abs(share - min(fr, 1.0 - fr)) <= 5e-5, (i, j, share, fr)
            checked += 1
    assert checked >= 40
