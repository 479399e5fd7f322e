"""Python binding of librasterjoin.so (the C ABI in include/rasterjoin.h).

`Handler` mirrors the reference's join API, the GLHandler singleton
(RasterJoin/join/GLHandler.hpp:20-78): set the data, the accuracy and the
polygon index resolution, then `execute_function(fn)` for one of the four
FunctionType values, `get_error_bounds(fn)` and `print_time_stats(fn)`.
All joins run on the GPU through the library; there is no CPU fallback, and
loading fails loudly if the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "librasterjoin.so")

# FunctionType (GLHandler.hpp:23-29)
RASTER, INDEX, HYBRID, ERRORBOUNDS = 0, 1, 2, 3
FUNCTIONS = {"raster": RASTER, "index": INDEX, "hybrid": HYBRID, "errorbounds": ERRORBOUNDS}
# AttributeType (Common.h:20) + u8 storage
UINT, INT, FLOAT, UINT8 = 1, 2, 3, 4
# ConstraintType (Common.h:22)
EQ, LT, LTE, GT, GTE = 0, 1, 2, 3, 4
# Aggregation (Common.h:26)
COUNT, AVG, SUM = 0, 1, 2
# options
OPT_RASTER_PATH, OPT_TIMING, OPT_HYBRID_DILATION, OPT_CACHE, OPT_ORGANIZE = 1, 2, 3, 4, 5
OPT_TILE_SHIFT, OPT_PASS_WORKGROUPS, OPT_BATCH_ROWS, OPT_BATCH_WORKGROUPS, OPT_BATCH_CHUNK_ROWS, OPT_GRAPH = \
    6, 7, 8, 9, 10, 11
OPT_INDEX_PATH, OPT_EXACT_QUEUE, OPT_STREAM_ROWS, OPT_BATCH_HYBRID_ROWS, OPT_BATCH_SHARE = 12, 13, 14, 15, 16
OPT_GRID_DIRECT = 17
# how a query ran
PATH_SINGLE, PATH_FUSED = 0, 1
ERR_ARG, ERR_STATE, ERR_NOMEM, ERR_DEVICE, ERR_LIMIT, ERR_UNSUPPORTED = -1, -2, -3, -4, -5, -6

ERRORS = {0: "ok", -1: "bad argument", -2: "bad state", -3: "out of memory",
          -4: "device error", -5: "limit exceeded", -6: "unsupported"}

_lib = None


class RJError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)} ({code}): {msg}")
        self.code = code


class Constraint(C.Structure):
    """QueryConstraint (Common.h:34-39): 16 bytes."""
    _fields_ = [("attrib_id", C.c_int32), ("type", C.c_int32), ("val", C.c_uint32),
                ("dummy", C.c_int32)]


class PointParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("box", C.c_double), ("ncenters", C.c_int32),
                ("frac_cluster", C.c_double), ("sigma_lo", C.c_double), ("sigma_hi", C.c_double),
                ("t0", C.c_uint32), ("tspan", C.c_uint32), ("fare_mu", C.c_double),
                ("fare_sigma", C.c_double)]


class VoronoiParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("nsites", C.c_int32), ("box", C.c_double),
                ("spacing_lo", C.c_double), ("spacing_hi", C.c_double), ("jitter", C.c_double)]


class TableOptions(C.Structure):
    _fields_ = [("organize", C.c_int32), ("time_col", C.c_int32), ("time0", C.c_uint32),
                ("bucket_seconds", C.c_uint32), ("host", C.c_int32)]


class Query(C.Structure):
    """rj_query: one self-contained query."""
    _fields_ = [("polys", C.c_void_p), ("fn", C.c_int32), ("agg_type", C.c_int32),
                ("agg_col", C.c_int32), ("ncons", C.c_int32), ("cons", Constraint * 10),
                ("eps_m", C.c_double), ("ground_res", C.c_double), ("index_res", C.c_int32),
                ("reserved", C.c_int32), ("row_begin", C.c_int64), ("row_end", C.c_int64)]


class Result(C.Structure):
    _fields_ = [("counts", C.c_void_p), ("sums", C.c_void_p), ("sum_hi", C.c_void_p),
                ("sum_lo", C.c_void_p), ("bounds", C.c_void_p), ("status", C.c_int32),
                ("path", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch", C.c_int32), ("max_pending", C.c_int32), ("mem_mb", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("batches", C.c_int64), ("mean_batch", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("max_ms", C.c_double), ("qps", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C raster-join_amd`")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        sig = {
            "rj_create": [i64, i32, vp], "rj_destroy": [vp], "rj_set_polygons": [vp, vp, vp, i32, vp],
            "rj_set_points": [vp, i64, vp, i32, vp, vp], "rj_set_constraints": [vp, vp, i32],
            "rj_set_aggregation": [vp, i32, i32], "rj_set_accuracy": [vp, f64, f64],
            "rj_set_index_res": [vp, i32, i32], "rj_execute": [vp, i32, vp, vp, vp],
            "rj_get_sums_fixed": [vp, vp, vp], "rj_error_bounds": [vp, vp],
            "rj_time_stats": [vp, i32, vp, C.c_size_t], "rj_canvas_info": [vp, vp, vp],
            "rj_set_option": [vp, i32, i64], "rj_kernel_times": [vp, i32, vp, C.c_size_t, vp, vp],
            "rj_sync": [vp], "rj_table_create": [vp, i64, vp, i32, vp, vp, i32, vp],
            "rj_table_rows": [vp, vp], "rj_table_read": [vp, i64, i64, vp, vp],
            "rj_table_destroy": [vp], "rj_bind_table": [vp, vp, i64, i64],
            "rj_gen_points": [vp, i64, i64, vp, vp, vp, vp], "rj_gen_voronoi": [vp, vp],
            "rj_polys_size": [vp, vp, vp], "rj_polys_copy": [vp, vp, vp],
            "rj_table_create_ex": [vp, i64, vp, i32, vp, vp, vp, vp],
            "rj_table_time_rows": [vp, C.c_uint32, C.c_uint32, vp, vp],
            "rj_polyset_create": [vp, vp, vp, i32, vp, vp], "rj_polyset_count": [vp, vp],
            "rj_polyset_bytes": [vp, vp], "rj_polyset_destroy": [vp], "rj_bind_polyset": [vp, vp],
            "rj_execute_query": [vp, vp, vp, vp], "rj_execute_batch": [vp, vp, i32, vp, vp],
            "rj_execute_query_noalloc": [vp, vp, vp, vp],
            "rj_polyset_grid": [vp, vp, i32, vp, vp, vp, vp, vp],
            "rj_read_points": [C.c_char_p, i32, i32, i32, vp, i64, vp, vp, vp, vp, vp],
            "rj_autotune": [vp, vp, i32, vp, vp],
            "rj_server_create": [vp, vp, vp, vp], "rj_server_submit": [vp, vp, vp, vp],
            "rj_server_poll": [vp, C.c_uint64, i32], "rj_server_stats_get": [vp, vp],
            "rj_server_stats_reset": [vp], "rj_server_destroy": [vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.rj_last_error.argtypes = [vp]
        L.rj_last_error.restype = C.c_char_p
        L.rj_polys_free.argtypes = [vp]
        L.rj_polys_free.restype = None
        L.rj_exact_points.argtypes = [vp]
        L.rj_exact_points.restype = C.c_int64
        _lib = L
    return _lib


def _p(a):
    """Address of a numpy array's data for a void* argument (an int: cheaper
    than a ctypes pointer object on every call)."""
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------------------
# synthetic inputs (rasterjoin_datagen.h)
# ---------------------------------------------------------------------------

def point_params(seed=7, box=50000.0, ncenters=40, frac_cluster=0.7, sigma=(200.0, 2000.0),
                 t0=1230768000, tspan=31536000, fare=(2.5, 0.6)):
    return PointParams(seed, box, ncenters, frac_cluster, sigma[0], sigma[1], t0, tspan, fare[0],
                       fare[1])


def gen_points(n, first=0, params=None, columns=True):
    """Seeded point table rows [first, first+n): xy (2n f32), time u32, fare f32,
    passengers u8."""
    params = params or point_params()
    xy = np.empty(2 * n, np.float32)
    if columns:
        t = np.empty(n, np.uint32)
        f = np.empty(n, np.float32)
        ps = np.empty(n, np.uint8)
    else:
        t = f = ps = None
    rc = lib().rj_gen_points(C.byref(params), first, n, _p(xy), _p(t), _p(f), _p(ps))
    if rc:
        raise RJError(rc, "rj_gen_points")
    return xy, t, f, ps


def gen_voronoi(nsites, seed=1, spacing=(60.0, 60.0), jitter=30.0, box=50000.0):
    """Voronoi polygon set with shared jittered edges: (xy f32 CSR, offsets i32)."""
    p = VoronoiParams(seed, nsites, box, spacing[0], spacing[1], jitter)
    h = C.c_void_p()
    rc = lib().rj_gen_voronoi(C.byref(p), C.byref(h))
    if rc:
        raise RJError(rc, "rj_gen_voronoi")
    np_, nv = C.c_int32(), C.c_int64()
    lib().rj_polys_size(h, C.byref(np_), C.byref(nv))
    xy = np.empty(2 * nv.value, np.float32)
    off = np.empty(np_.value + 1, np.int32)
    lib().rj_polys_copy(h, _p(xy), _p(off))
    lib().rj_polys_free(h)
    return xy, off


# The benchmark configurations (point counts, polygon sets, accuracy).
POLY_SETS = {
    "voronoi260": dict(nsites=260, seed=1, spacing=(60.0, 60.0), jitter=30.0),
    "tracts2200": dict(nsites=2200, seed=3, spacing=(2.2, 140.0), jitter=30.0),
    # the query mixes' collections of 10 to 5000 polygons
    "voronoi10": dict(nsites=10, seed=101, spacing=(400.0, 400.0), jitter=30.0),
    "voronoi40": dict(nsites=40, seed=102, spacing=(150.0, 150.0), jitter=30.0),
    "voronoi900": dict(nsites=900, seed=104, spacing=(40.0, 40.0), jitter=30.0),
    "voronoi5000": dict(nsites=5000, seed=106, spacing=(25.0, 25.0), jitter=20.0),
}


def poly_set(name):
    """The (xy, off) of a named collection of POLY_SETS."""
    cp = POLY_SETS[name]
    return gen_voronoi(cp["nsites"], seed=cp["seed"], spacing=cp["spacing"], jitter=cp["jitter"])


TAXI_RECORD = np.dtype([("db_idx", "<u8"), ("pick_x", "<f8"), ("pick_y", "<f8"), ("drop_x", "<f8"),
                        ("drop_y", "<f8"), ("pickup_time", "<u8"), ("dropoff_time", "<u8"), ("vendor", "S4"),
                        ("duration", "<u4"), ("miles", "<f4"), ("fare", "<u2"), ("surcharge", "<u2"),
                        ("mta_tax", "<u2"), ("tip", "<u2"), ("toll", "<u2"), ("total", "<u2"),
                        ("medallion_id", "<u2"), ("license_id", "<u2"), ("store_and_forward", "u1"),
                        ("payment_type", "u1"), ("passengers", "u1"), ("rate_code", "u1")])
TWEET_RECORD = np.dtype([("user_id", "<i8"), ("time", "<i8"), ("loc_x", "<f4"), ("loc_y", "<f4"),
                         ("favCt", "<i4"), ("rtCt", "<i4")])
DATASET_TAXI, DATASET_TWITTER = 0, 1


def read_points(path, dataset, loc_attrib, attribs=(), max_records=0):
    """rj_read_points: the kept rows' locations (2 floats per row) and the
    attribute columns (uint32/int32/float32 arrays) of a file of the
    reference's point records; returns (xy, cols, records read)."""
    na = len(attribs)
    attribs = np.ascontiguousarray(list(attribs) or [0], np.int32)
    nrec, kept = C.c_int64(), C.c_int64()
    p = os.fsencode(path)
    rc = lib().rj_read_points(p, dataset, loc_attrib, na, _p(attribs), max_records, C.byref(nrec), None, None,
                              None, None)
    if rc:
        raise RJError(rc, "rj_read_points")
    xy = np.zeros(2 * max(1, nrec.value), np.float32)
    raw = [np.zeros(max(1, nrec.value), np.uint32) for _ in range(na)]
    colptr = (C.c_void_p * max(1, na))(*[_p(c) for c in raw])
    types = np.zeros(max(1, na), np.int32)
    rc = lib().rj_read_points(p, dataset, loc_attrib, na, _p(attribs), max_records, C.byref(nrec), C.byref(kept),
                              _p(xy), C.cast(colptr, C.c_void_p), _p(types))
    if rc:
        raise RJError(rc, "rj_read_points")
    k = kept.value
    view = {UINT: np.uint32, INT: np.int32, FLOAT: np.float32}
    cols = [c[:k].view(view[int(t)]) for c, t in zip(raw, types[:na])]
    return xy[:2 * k], cols, nrec.value


def make_constraints(cons):
    """cons: iterable of (column, ConstraintType, value); float values are
    stored as float bits, ints as int/uint."""
    cons = list(cons)
    arr = (Constraint * max(1, len(cons)))()
    for i, (col, typ, val) in enumerate(cons):
        arr[i].attrib_id = col
        arr[i].type = typ
        if isinstance(val, (float, np.floating)):
            arr[i].val = int(np.array([val], np.float32).view(np.uint32)[0])
        else:
            arr[i].val = int(val) & 0xFFFFFFFF
    return arr, len(cons)


# ---------------------------------------------------------------------------
# join API
# ---------------------------------------------------------------------------

class Table:
    """A point table (rj_table): resident in device memory, or with
    host=True kept in host memory and streamed through the device in
    passes per query."""

    def __init__(self, handler, xy, cols=(), types=(), organize=True, time_col=-1, time0=0,
                 bucket_seconds=0, host=False):
        """time_col >= 0 (a UINT column) orders the rows by (time bucket,
        Morton key) with buckets of bucket_seconds starting at time0."""
        self.h = C.c_void_p()
        xy = np.ascontiguousarray(xy, np.float32)
        cols = [np.ascontiguousarray(c) for c in cols]
        colptr = (C.c_void_p * max(1, len(cols)))(*[_p(c) for c in cols])
        typ = np.array(list(types) or [0], np.int32)
        if time_col >= 0 or host:
            opt = TableOptions(int(organize), time_col, time0, bucket_seconds, int(host))
            handler._check(lib().rj_table_create_ex(handler.ctx, len(xy) // 2, _p(xy), len(cols),
                                                    C.cast(colptr, C.c_void_p), _p(typ),
                                                    C.byref(opt), C.byref(self.h)))
        else:
            handler._check(lib().rj_table_create(handler.ctx, len(xy) // 2, _p(xy), len(cols),
                                                 C.cast(colptr, C.c_void_p), _p(typ), int(organize),
                                                 C.byref(self.h)))
        self.n = len(xy) // 2
        self.types = list(types)
        self.host = bool(host)

    def time_rows(self, t_lo, t_hi):
        b, e = C.c_int64(), C.c_int64()
        rc = lib().rj_table_time_rows(self.h, t_lo, t_hi, C.byref(b), C.byref(e))
        if rc:
            raise RJError(rc, "rj_table_time_rows")
        return b.value, e.value

    def read(self, row0=0, n=None):
        n = self.n - row0 if n is None else n
        xy = np.empty(2 * n, np.float32)
        cols = []
        for t in self.types:
            cols.append(np.empty(n, np.uint8 if t == UINT8 else
                                 (np.float32 if t == FLOAT else
                                  (np.int32 if t == INT else np.uint32))))
        colptr = (C.c_void_p * max(1, len(cols)))(*[_p(c) for c in cols])
        rc = lib().rj_table_read(self.h, row0, n, _p(xy), C.cast(colptr, C.c_void_p))
        if rc:
            raise RJError(rc, "rj_table_read")
        return xy, cols

    def close(self):
        if self.h:
            lib().rj_table_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class PolySet:
    """A resident polygon collection (rj_polyset) of one Handler."""

    def __init__(self, handler, xy, off, bbox=None):
        self.h = C.c_void_p()
        xy = np.ascontiguousarray(xy, np.float32)
        off = np.ascontiguousarray(off, np.int32)
        b = None if bbox is None else np.ascontiguousarray(bbox, np.float64)
        handler._check(lib().rj_polyset_create(handler.ctx, _p(xy), _p(off), len(off) - 1, _p(b),
                                               C.byref(self.h)))
        self.np = len(off) - 1
        self.handler = handler

    def grid(self, res):
        """The device-built CPUJoin grid at resolution res: (cell_off,
        cand, deficient cells)."""
        nc, nd = C.c_int64(), C.c_int64()
        h = self.handler
        h._check(lib().rj_polyset_grid(h.ctx, self.h, res, None, None, C.byref(nc), None, C.byref(nd)))
        cell_off = np.zeros(res * res + 1, np.int32)
        cand = np.zeros(max(1, nc.value), np.int32)
        defi = np.zeros(max(1, nd.value), np.int32)
        h._check(lib().rj_polyset_grid(h.ctx, self.h, res, _p(cell_off), _p(cand), C.byref(nc), _p(defi),
                                       C.byref(nd)))
        return cell_off, cand[:nc.value], defi[:nd.value]

    def device_bytes(self):
        b = C.c_int64()
        lib().rj_polyset_bytes(self.h, C.byref(b))
        return b.value

    def close(self):
        # a collection outlives its context on the C side, so it can always be destroyed
        if self.h:
            lib().rj_polyset_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()


def make_query(polys, fn, agg=COUNT, agg_col=-1, cons=(), eps=10.0, ground_res=0.0, index_res=0,
               rows=(0, -1)):
    """An rj_query record for a PolySet."""
    q = Query()
    q.polys = polys.h
    q.fn = fn
    q.agg_type = agg
    q.agg_col = agg_col
    arr, n = make_constraints(cons)
    q.ncons = n
    for i in range(n):
        q.cons[i] = arr[i]
    q.eps_m = eps
    q.ground_res = ground_res
    q.index_res = index_res
    q.row_begin, q.row_end = rows
    return q


class QueryResult:
    """Output buffers of one query (counts, sums, exact sums, bounds)."""

    def __init__(self, np_):
        self.counts = np.zeros(np_, np.int64)
        self.sums = np.zeros(np_, np.float64)
        self.hi = np.zeros(np_, np.int64)
        self.lo = np.zeros(np_, np.uint64)
        self.bounds = np.zeros(4 * np_, np.int64)
        self.rec = Result(_p(self.counts), _p(self.sums), _p(self.hi), _p(self.lo), _p(self.bounds),
                          0, 0)

    @property
    def status(self):
        return self.rec.status

    @property
    def path(self):
        return self.rec.path


class Server:
    """Serving loop (rj_server) over one resident table."""

    def __init__(self, handler, table, max_batch=0, max_pending=0, mem_mb=0):
        self.h = C.c_void_p()
        cfg = ServerConfig(max_batch, max_pending, mem_mb)
        handler._check(lib().rj_server_create(handler.ctx, table.h, C.byref(cfg), C.byref(self.h)))
        self._inflight = {}

    def submit(self, q, np_):
        res = QueryResult(np_)
        t = C.c_uint64()
        rc = lib().rj_server_submit(self.h, C.byref(q), C.byref(res.rec), C.byref(t))
        if rc:
            raise RJError(rc, "rj_server_submit")
        self._inflight[t.value] = res
        return t.value

    def poll(self, ticket, wait_ms=0):
        """The QueryResult once done, else None."""
        rc = lib().rj_server_poll(self.h, ticket, int(wait_ms))
        if rc == 0:
            return None
        res = self._inflight.pop(ticket)
        if rc < 0:
            raise RJError(rc, f"query {ticket} failed")
        return res

    def stats(self):
        st = ServerStats()
        lib().rj_server_stats_get(self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        lib().rj_server_stats_reset(self.h)

    def close(self):
        if self.h:
            lib().rj_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Handler:
    """GLHandler-shaped front end of one rj_ctx."""

    def __init__(self, gpu_mem_mb=0, inmem=True):
        self.ctx = C.c_void_p()
        rc = lib().rj_create(int(gpu_mem_mb), int(bool(inmem)), C.byref(self.ctx))
        if rc:
            raise RJError(rc, "rj_create (is a GPU visible?)")
        self.np = 0
        self._keep = []

    def _check(self, rc):
        if rc:
            raise RJError(rc, lib().rj_last_error(self.ctx).decode())
        return rc

    def close(self):
        if self.ctx:
            lib().rj_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    # --- setters (DataHandler / PolyHandler / GLHandler) ---
    def set_polygons(self, xy, off, bbox=None):
        xy = np.ascontiguousarray(xy, np.float32)
        off = np.ascontiguousarray(off, np.int32)
        b = None if bbox is None else np.ascontiguousarray(bbox, np.float64)
        self._check(lib().rj_set_polygons(self.ctx, _p(xy), _p(off), len(off) - 1, _p(b)))
        self.np = len(off) - 1

    def set_points(self, xy, cols=(), types=()):
        xy = np.ascontiguousarray(xy, np.float32)
        cols = [np.ascontiguousarray(c) for c in cols]
        colptr = (C.c_void_p * max(1, len(cols)))(*[_p(c) for c in cols])
        typ = np.array(list(types) or [0], np.int32)
        self._check(lib().rj_set_points(self.ctx, len(xy) // 2, _p(xy), len(cols),
                                        C.cast(colptr, C.c_void_p), _p(typ)))

    def bind_polyset(self, ps):
        self._check(lib().rj_bind_polyset(self.ctx, ps.h))
        self.np = ps.np

    def bind_table(self, table, row_begin=0, row_end=None):
        row_end = table.n if row_end is None else row_end
        self._check(lib().rj_bind_table(self.ctx, table.h, row_begin, row_end))
        self._keep.append(table)

    def set_query_constraints(self, cons):
        arr, n = make_constraints(cons)
        self._check(lib().rj_set_constraints(self.ctx, C.cast(arr, C.c_void_p), n))

    def set_aggregation(self, agg=COUNT, col=-1):
        self._check(lib().rj_set_aggregation(self.ctx, agg, col))

    def set_accuracy_distance(self, meters, ground_res=1.0):
        self._check(lib().rj_set_accuracy(self.ctx, float(meters), float(ground_res)))

    def set_poly_index_resolution(self, rx, ry=None):
        self._check(lib().rj_set_index_res(self.ctx, rx, rx if ry is None else ry))

    def set_option(self, key, value):
        self._check(lib().rj_set_option(self.ctx, key, int(value)))

    # --- execution ---
    def execute_function(self, fn, legacy=False):
        """Returns (counts int64[np], sums float64[np]) or with legacy=True the
        reference's int32[3*np] result vector as well."""
        counts = np.zeros(self.np, np.int64)
        sums = np.zeros(self.np, np.float64)
        leg = np.zeros(3 * self.np, np.int32) if legacy else None
        self._check(lib().rj_execute(self.ctx, fn, _p(counts), _p(sums), _p(leg)))
        return (counts, sums, leg) if legacy else (counts, sums)

    def sums_fixed(self):
        hi = np.zeros(self.np, np.int64)
        lo = np.zeros(self.np, np.uint64)
        self._check(lib().rj_get_sums_fixed(self.ctx, _p(hi), _p(lo)))
        return hi, lo

    def get_error_bounds(self, fn=ERRORBOUNDS):
        if fn != ERRORBOUNDS:
            return np.zeros(0, np.int64)
        out = np.zeros(4 * self.np, np.int64)
        self._check(lib().rj_error_bounds(self.ctx, _p(out)))
        return out

    def print_time_stats(self, fn):
        buf = C.create_string_buffer(1024)
        self._check(lib().rj_time_stats(self.ctx, fn, buf, 1024))
        return buf.value.decode()

    def canvas_size(self):
        W, H = C.c_int32(), C.c_int32()
        self._check(lib().rj_canvas_info(self.ctx, C.byref(W), C.byref(H)))
        return W.value, H.value

    def kernel_times(self):
        out = {}
        name = C.create_string_buffer(64)
        ms = C.c_double()
        n = C.c_int32()
        k = lib().rj_kernel_times(self.ctx, -1, None, 0, None, None)
        for i in range(max(0, k)):
            lib().rj_kernel_times(self.ctx, i, name, 64, C.byref(ms), C.byref(n))
            out[name.value.decode()] = (ms.value, n.value)
        return out

    def exact_points(self):
        """Points the last execute queued for the separate exact pass (border
        points of hybrid and label-accelerated index joins), or -1."""
        return int(lib().rj_exact_points(self.ctx))

    def sync(self):
        self._check(lib().rj_sync(self.ctx))

    # --- self-contained queries ---
    def execute_query(self, table, q, np_):
        res = QueryResult(np_)
        rc = lib().rj_execute_query(self.ctx, table.h, C.byref(q), C.byref(res.rec))
        self._check(rc)
        return res

    def execute_query_noalloc(self, table, q, np_):
        """rj_execute_query_noalloc: the query on resident structures, with its
        per-call buffers in the library's static scratch array."""
        res = QueryResult(np_)
        rc = lib().rj_execute_query_noalloc(self.ctx, table.h, C.byref(q), C.byref(res.rec))
        self._check(rc)
        return res

    def execute_batch(self, table, queries, nps):
        """queries: list of Query; nps: polygon count of each. Returns a list of
        QueryResult (check .status)."""
        n = len(queries)
        qa = (Query * max(1, n))(*queries)
        outs = [QueryResult(k) for k in nps]
        ra = (Result * max(1, n))(*[o.rec for o in outs])
        self._check(lib().rj_execute_batch(self.ctx, table.h, n, C.cast(qa, C.c_void_p),
                                           C.cast(ra, C.c_void_p)))
        for i, o in enumerate(outs):
            o.rec = ra[i]
        return outs

    def autotune(self, table, queries):
        n = len(queries)
        qa = (Query * max(1, n))(*queries)
        chosen = np.zeros(5, np.int32)
        self._check(lib().rj_autotune(self.ctx, table.h, n, C.cast(qa, C.c_void_p), _p(chosen)))
        return dict(zip(("tile_shift", "rows", "workgroups", "chunk_rows", "grid_res"), chosen.tolist()))
