"""Measurement aids. Kernel timing (RJ_OPT_TIMING): 1 records every kernel
of the last execute; 2 records only the kernels that stream the table, summed
over every execute since the option was set, and leaves the results
unchanged. rj_exact_points: the size of the last execute's border queue."""
import numpy as np
import pytest

import rasterjoin as rj
from common import COLS_TYPES, points, polys

pytestmark = pytest.mark.gpu


def test_timing_modes(handler):
    h = handler
    xy, off = polys("voronoi260")
    pts, t, fare, pas = points(500_000, 3)
    h.set_polygons(xy, off)
    h.set_points(pts, (t, fare, pas), COLS_TYPES)
    h.set_query_constraints(())
    h.set_aggregation(rj.AVG, 1)
    h.set_accuracy_distance(20.0)
    h.set_option(rj.OPT_CACHE, 0)
    try:
        h.set_option(rj.OPT_TIMING, 1)
        for _ in range(2):
            c1, s1 = h.execute_function(rj.RASTER)
        k1 = h.kernel_times()
        assert k1["points_label"][1] == 1 and "band_label" in k1
        assert all(ms >= 0 for ms, _ in k1.values())
        h.set_option(rj.OPT_TIMING, 2)
        assert h.kernel_times() == {}
        for _ in range(3):
            c2, s2 = h.execute_function(rj.RASTER)
        k2 = h.kernel_times()
        assert set(k2) == {"points_label"} and k2["points_label"][1] == 3
        assert k2["points_label"][0] > 0
        assert np.array_equal(c1, c2) and np.array_equal(s1, s2)
        with pytest.raises(rj.RJError):
            h.set_option(rj.OPT_TIMING, 3)
    finally:
        h.set_option(rj.OPT_TIMING, 0)
        h.set_option(rj.OPT_CACHE, 1)


def test_exact_points_counts_the_border_queue(handler):
    """rj_exact_points: the points a hybrid join queued for the exact pass
    (some, never all, of the selected rows); -1 after a raster join."""
    h = handler
    xy, off = polys("voronoi260")
    pts = points(400_000, 4)[0]
    h.set_polygons(xy, off)
    h.set_points(pts, (), ())
    h.set_query_constraints(())
    h.set_aggregation(rj.COUNT, -1)
    h.set_accuracy_distance(20.0)
    h.set_poly_index_resolution(256)
    c, _ = h.execute_function(rj.HYBRID)
    k = h.exact_points()
    assert 0 < k < c.sum()
    h.execute_function(rj.RASTER)
    assert h.exact_points() == -1
