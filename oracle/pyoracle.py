"""ctypes wrapper of the C oracle (oracle/rj_oracle.c) and of the reference's
own GridIndexF build (oracle/_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgridindexf_ref.so")

_lib = None
_ref = None


class Constraint(C.Structure):
    _fields_ = [("attrib_id", C.c_int32), ("type", C.c_int32), ("val", C.c_uint32),
                ("dummy", C.c_int32)]


class Points(C.Structure):
    _fields_ = [("n", C.c_int64), ("xy", C.c_void_p), ("ncols", C.c_int32),
                ("cols", C.c_void_p), ("types", C.c_void_p), ("ncons", C.c_int32),
                ("cons", C.c_void_p), ("agg_type", C.c_int32), ("agg_col", C.c_int32)]


class Plan(C.Structure):
    _fields_ = [("W", C.c_int32), ("H", C.c_int32), ("splitx", C.c_int32), ("splity", C.c_int32),
                ("minx", C.c_double), ("miny", C.c_double), ("maxx", C.c_double),
                ("maxy", C.c_double), ("minx_f", C.c_float), ("miny_f", C.c_float),
                ("maxx_f", C.c_float), ("maxy_f", C.c_float), ("sx_f", C.c_float),
                ("sy_f", C.c_float), ("fxs", C.c_double), ("fys", C.c_double),
                ("cell", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_grid_build.restype = C.c_void_p
        _lib.orc_grid_build.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                        C.c_int32, C.c_int32, C.c_int32]
        _lib.orc_grid_ncand.restype = C.c_int64
        _lib.orc_grid_ncand.argtypes = [C.c_void_p]
        _lib.orc_grid_export.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_grid_free.argtypes = [C.c_void_p]
        _lib.orc_join_exact.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_join_brute.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                        C.c_void_p]
        _lib.orc_plan_make.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p]
        _lib.orc_pixel.restype = C.c_int64
        _lib.orc_pixel.argtypes = [C.c_void_p, C.c_float, C.c_float]
        _lib.orc_snap.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        _lib.orc_coverage.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        _lib.orc_raster_join.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p]
        _lib.orc_boundary.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        _lib.orc_area.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
        _lib.orc_errorbounds.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                         C.c_void_p, C.c_void_p]
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
        _ref.ref_grid_build.restype = C.c_void_p
        _ref.ref_grid_build.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                        C.c_int32, C.c_char_p]
        _ref.ref_grid_ncand.restype = C.c_int64
        _ref.ref_grid_ncand.argtypes = [C.c_void_p]
        _ref.ref_grid_export.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _ref.ref_join_count.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        _ref.ref_grid_free.argtypes = [C.c_void_p]
    return _ref


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def poly_bbox(xy):
    """Collection bbox as PolyHandler::readPolygons computes it (float min/max)."""
    xs, ys = xy[0::2], xy[1::2]
    return np.array([xs.min(), ys.min(), xs.max(), ys.max()], dtype=np.float32)


class Grid:
    """Candidate lists of GridIndexF::buildGrid (cell index = y*xs + x)."""

    def __init__(self, xy, off, res, bbox=None, nthreads=0):
        self.xy = np.ascontiguousarray(xy, np.float32)
        self.off = np.ascontiguousarray(off, np.int32)
        self.np = len(off) - 1
        self.res = res
        self.bbox = poly_bbox(self.xy) if bbox is None else np.asarray(bbox, np.float32)
        self.h = lib().orc_grid_build(_p(self.xy), _p(self.off), self.np, _p(self.bbox), res, res,
                                      nthreads)

    def export(self):
        n = lib().orc_grid_ncand(self.h)
        cell_off = np.zeros(self.res * self.res + 1, np.int32)
        cand = np.zeros(max(n, 1), np.int32)
        lib().orc_grid_export(self.h, _p(cell_off), _p(cand))
        return cell_off, cand[:n]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_grid_free(self.h)
            self.h = None


class PointSet:
    """Keeps numpy buffers alive for an orc_points struct."""

    def __init__(self, xy, cols=(), types=(), cons=(), agg_type=0, agg_col=0):
        self.xy = np.ascontiguousarray(xy, np.float32)
        self.cols = [np.ascontiguousarray(c) for c in cols]
        self.types = np.array(list(types) or [0], np.int32)
        self.colptr = (C.c_void_p * max(1, len(self.cols)))(*[_p(c) for c in self.cols])
        arr = (Constraint * max(1, len(cons)))()
        for i, (aid, typ, val) in enumerate(cons):
            arr[i].attrib_id = aid
            arr[i].type = typ
            if isinstance(val, float):
                arr[i].val = int(np.array([val], np.float32).view(np.uint32)[0])
            else:
                arr[i].val = int(np.array([val], np.int64).astype(np.uint32)[0]) if val < 0 \
                    else int(val) & 0xFFFFFFFF
        self.cons = arr
        self.s = Points(len(self.xy) // 2, _p(self.xy), len(self.cols),
                        C.cast(self.colptr, C.c_void_p), _p(self.types), len(cons),
                        C.cast(self.cons, C.c_void_p), agg_type, agg_col)


def join_exact(grid, ps, nthreads=0):
    np_ = grid.np
    counts = np.zeros(np_, np.int64)
    hi = np.zeros(np_, np.int64)
    lo = np.zeros(np_, np.uint64)
    f64 = np.zeros(np_, np.float64)
    lib().orc_join_exact(C.c_void_p(grid.h), C.byref(ps.s), nthreads, _p(counts), _p(hi), _p(lo),
                         _p(f64))
    return counts, hi, lo, f64


def join_brute(xy, off, ps):
    off = np.ascontiguousarray(off, np.int32)
    xy = np.ascontiguousarray(xy, np.float32)
    counts = np.zeros(len(off) - 1, np.int64)
    lib().orc_join_brute(_p(xy), _p(off), len(off) - 1, C.byref(ps.s), _p(counts))
    return counts


def plan(bbox, eps, ground_res=1.0):
    p = Plan()
    b = np.asarray(bbox, np.float64)
    rc = lib().orc_plan_make(_p(b), eps, ground_res, C.byref(p))
    if rc != 0:
        raise ValueError("bad plan")
    return p


def pixel(p, x, y):
    return lib().orc_pixel(C.byref(p), x, y)


def snap(p, xy):
    xy = np.ascontiguousarray(xy, np.float32)
    n = len(xy) // 2
    X = np.zeros(n, np.int64)
    Y = np.zeros(n, np.int64)
    lib().orc_snap(C.byref(p), _p(xy), n, _p(X), _p(Y))
    return X, Y


def coverage(p, xy, method):
    xy = np.ascontiguousarray(xy, np.float32)
    sub = 4 if method >= 2 else 1
    out = np.zeros((p.H * sub, p.W * sub), np.uint8)
    lib().orc_coverage(C.byref(p), _p(xy), len(xy) // 2, method, _p(out))
    return out


AREA_ONE = 1 << 24


def area(p, xy):
    """Signed area of the snapped ring inside each pixel, H x W int32 in
    units of 2^-24 of a pixel (orc_area)."""
    xy = np.ascontiguousarray(xy, np.float32)
    out = np.zeros((p.H, p.W), np.int32)
    lib().orc_area(C.byref(p), _p(xy), len(xy) // 2, _p(out))
    return out


def boundary(p, xy, r=1):
    xy = np.ascontiguousarray(xy, np.float32)
    out = np.zeros((p.H, p.W), np.uint8)
    lib().orc_boundary(C.byref(p), _p(xy), len(xy) // 2, r, _p(out))
    return out


def raster_join(p, pxy, off, ps, want_canvas=False):
    pxy = np.ascontiguousarray(pxy, np.float32)
    off = np.ascontiguousarray(off, np.int32)
    np_ = len(off) - 1
    counts = np.zeros(np_, np.int64)
    hi = np.zeros(np_, np.int64)
    lo = np.zeros(np_, np.uint64)
    canvas = np.zeros(p.W * p.H, np.uint32)
    lib().orc_raster_join(C.byref(p), _p(pxy), _p(off), np_, C.byref(ps.s), _p(counts), _p(hi),
                          _p(lo), _p(canvas))
    if want_canvas:
        return counts, hi, lo, canvas
    return counts, hi, lo


def errorbounds(p, pxy, off, canvas):
    pxy = np.ascontiguousarray(pxy, np.float32)
    off = np.ascontiguousarray(off, np.int32)
    np_ = len(off) - 1
    out = np.zeros(4 * np_, np.int64)
    lib().orc_errorbounds(C.byref(p), _p(pxy), _p(off), np_, _p(canvas), _p(out))
    return out


class RefGrid:
    """The reference's GridIndexF (compiled from the reference tree)."""

    def __init__(self, xy, off, res, tmpdir, bbox=None):
        self.xy = np.ascontiguousarray(xy, np.float32)
        self.off = np.ascontiguousarray(off, np.int32)
        self.np = len(off) - 1
        self.res = res
        self.bbox = poly_bbox(self.xy) if bbox is None else np.asarray(bbox, np.float32)
        prefix = os.path.join(str(tmpdir), "refgrid").encode()
        self.h = ref().ref_grid_build(_p(self.xy), _p(self.off), self.np, _p(self.bbox), res,
                                      prefix)

    def export(self):
        n = ref().ref_grid_ncand(self.h)
        cell_off = np.zeros(self.res * self.res + 1, np.int32)
        cand = np.zeros(max(n, 1), np.int32)
        ref().ref_grid_export(self.h, _p(cell_off), _p(cand))
        return cell_off, cand[:n]

    def join_count(self, pts_xy):
        pts_xy = np.ascontiguousarray(pts_xy, np.float32)
        counts = np.zeros(self.np, np.int64)
        ref().ref_join_count(self.h, len(pts_xy) // 2, _p(pts_xy), _p(counts))
        return counts

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_grid_free(self.h)
            self.h = None
