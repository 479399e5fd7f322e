/* rasterjoin.h -- C ABI of the MI355X raster-join library (librasterjoin.so).
 *
 * The entry points replace the reference's join API, the GLHandler singleton
 * and its GLFunction plugins (RasterJoin/join/GLHandler.hpp:20-78), with plain
 * pointers and sizes. Reference paths are relative to the reference root.
 *
 * Conventions
 *  - Every function returns an int status: RJ_OK (0) or a negative RJ_ERR_*.
 *    Nothing calls exit(); rj_last_error() explains the last failure.
 *  - Output buffers are owned by the caller.
 *  - One context per host thread; several contexts may share a device. Work is
 *    ordered on the context's own HIP stream; entry points return after the
 *    results they report are on the host.
 */
#ifndef RASTERJOIN_H
#define RASTERJOIN_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RJ_OK 0
#define RJ_ERR_ARG (-1)         /* bad argument */
#define RJ_ERR_STATE (-2)       /* missing polygons / points, wrong call order */
#define RJ_ERR_NOMEM (-3)       /* device allocation failed or memory budget exhausted */
#define RJ_ERR_DEVICE (-4)      /* HIP runtime error */
#define RJ_ERR_LIMIT (-5)       /* too many constraints or columns (GLFunction.cpp:250-253, :308) */
#define RJ_ERR_UNSUPPORTED (-6) /* function not supported for this input */

/* FunctionType, GLHandler.hpp:23-29 */
enum { RJ_RASTER = 0, RJ_INDEX = 1, RJ_HYBRID = 2, RJ_ERRORBOUNDS = 3 };
/* AttributeType, common/Common.h:20, plus a 1-byte unsigned storage type */
enum { RJ_ATTR_LOCATION = 0, RJ_ATTR_UINT = 1, RJ_ATTR_INT = 2, RJ_ATTR_FLOAT = 3, RJ_ATTR_UINT8 = 4 };
/* ConstraintType, common/Common.h:22 */
enum { RJ_EQ = 0, RJ_LT = 1, RJ_LTE = 2, RJ_GT = 3, RJ_GTE = 4 };
/* Aggregation, common/Common.h:26 */
enum { RJ_AGG_COUNT = 0, RJ_AGG_AVG = 1, RJ_AGG_SUM = 2 };

#define RJ_MAX_CONSTRAINTS 10 /* GLFunction.cpp:308 */
#define RJ_MAX_COLUMNS 8      /* columns of one point table (a query may use at most 5) */

/* 16-byte filter record, the layout of QueryConstraint (common/Common.h:34-39).
 * attrib_id is the index of the column in the table's column list. Uint and
 * Int columns compare as integers, Float columns as floats. */
typedef struct rj_constraint {
    int32_t attrib_id;
    int32_t type;
    union { int32_t ival; float fval; uint32_t uval; } val;
    int32_t dummy;
} rj_constraint;

typedef struct rj_ctx rj_ctx;
typedef struct rj_table rj_table;
typedef struct rj_polyset rj_polyset;
typedef struct rj_server rj_server;

/* GLHandler::getInstance(gpuMemMB, inMemory), GLHandler.cpp:64-69, and the
 * constructor it calls (GLHandler.cpp:43-54; the budget keeps its 100 MB
 * reserve, :51). */
int rj_create(int64_t gpu_mem_mb, int32_t inmem, rj_ctx** out);
/* GLHandler::~GLHandler, GLHandler.cpp:56-62 */
int rj_destroy(rj_ctx* ctx);
const char* rj_last_error(const rj_ctx* ctx);

/* Polygon collection: PolyHandler::addPolygonCollection, db/PolyHandler.cpp:49-75
 * (CSR vertices `polys` + `pindexes`). xy: 2*off[np] floats; polygons are
 * open rings (no repeated closing vertex) of >= 3 vertices. bbox (minx, miny,
 * maxx, maxy) may be NULL: it is then the vertex min/max, as readPolygons
 * computes it (PolyHandler.cpp:171-182). */
int rj_set_polygons(rj_ctx* ctx, const float* xy, const int32_t* off, int32_t np, const double* bbox);

/* Point columns: DataHandler / QueryResult::getAttribute (db/DataHandler.hpp:14-41,
 * db/QueryResult.hpp:56-79). xy is the location column (2n floats); cols[k]
 * has n entries of types[k] (RJ_ATTR_*). The data is copied into a context-
 * owned device table; host buffers may be freed afterwards. */
int rj_set_points(rj_ctx* ctx, int64_t n, const float* xy, int32_t ncols, const void* const* cols,
                  const int32_t* types);

/* Point records of the reference's binary datasets, read as
 * Dataset::getNextRecord reads them (db/Dataset.cpp:13-36): 88-byte
 * TaxiRecordData records (db/TaxiRecord.hpp:15-46) or 32-byte Tweet records
 * (db/TwitterRecord.hpp:19-28), native byte order. A record is kept when its
 * index location (taxi: pick_y, pick_x; twitter: loc_x, loc_y; as lat, lon
 * floats) passes transformPoint's US bounding-box filter
 * (common/UsefulFuncs.cpp:578-597); the location is then the projected one
 * (Web Mercator at zoom 22, minus the projection of the bbox center).
 * loc_attrib is the location attribute (taxi: 1 pickup, 2 dropoff, which the
 * reference does not project; twitter: 2); attribs[k] are 4-byte attributes,
 * written to cols[k] as getAttributeAsBinary gives them, types[k] receiving
 * their RJ_ATTR_* type: taxi 3 pickup time, 4 dropoff time, 6 duration, 8 fare
 * (cents), 18 passengers (Uint), 7 miles (Float); twitter 1 time (Uint),
 * 3 favorites, 4 retweets (Int). Other attributes: RJ_ERR_UNSUPPORTED.
 * At most max_records records are read (0: all). With xy NULL only
 * *n_records (the records that would be read) is set; otherwise xy (2 floats
 * per row) and cols[k] need room for *n_records rows and *n_kept receives the
 * rows kept. Host code only (no device work). */
enum { RJ_DATASET_TAXI = 0, RJ_DATASET_TWITTER = 1 };
int rj_read_points(const char* path, int32_t dataset, int32_t loc_attrib, int32_t nattr, const int32_t* attribs,
                   int64_t max_records, int64_t* n_records, int64_t* n_kept, float* xy, void* const* cols,
                   int32_t* types);

/* Resident point tables (replacing the disk backend, db/HashGridIndex.cpp:78-137).
 * organize = 1 orders the rows by a spatial (Morton) key so that splats and
 * lookups stay local; join results do not depend on row order. */
int rj_table_create(rj_ctx* ctx, int64_t n, const float* xy, int32_t ncols, const void* const* cols,
                    const int32_t* types, int32_t organize, rj_table** out);
int rj_table_rows(const rj_table* t, int64_t* n);
/* Copy rows [row0, row0+n) back to the host (any pointer may be NULL). */
int rj_table_read(const rj_table* t, int64_t row0, int64_t n, float* xy, void* const* cols);
int rj_table_destroy(rj_table* t);
/* Use rows [row_begin, row_end) of a table as the context's points. */
int rj_bind_table(rj_ctx* ctx, rj_table* t, int64_t row_begin, int64_t row_end);

/* Time-bucketed tables (replacing the backend's time partitions,
 * DataHandler::executeQuery, DataHandler.cpp:86-96, and the 10-day x 1 km
 * partitions of BackendIndex/main.cpp:23-71). With time_col >= 0 (a Uint
 * column) rows are ordered by (bucket, Morton key), bucket = (t - time0) /
 * bucket_seconds clamped to the table's buckets, so a query whose constraints
 * bound that column reads only the rows of its buckets; its constraints are
 * still applied exactly.
 * host = 1 keeps the rows in host memory (pinned when the driver grants it):
 * each query streams its rows through the device in chunks of
 * RJ_OPT_STREAM_ROWS rows, as the reference's point passes do
 * (GLFunction.cpp:329-337, RasterJoin.cpp:214-236), with the copy of one
 * chunk overlapped with the pass over the previous one. Host tables are
 * ordered by time bucket only (no Morton order); their queries in a batch
 * run one at a time. rj_set_points makes such a table when the points do
 * not fit the memory budget. */
typedef struct rj_table_options {
    int32_t organize;        /* 1: Morton order (within each bucket) */
    int32_t time_col;        /* -1: no buckets */
    uint32_t time0;          /* start of bucket 0 */
    uint32_t bucket_seconds; /* bucket width, > 0 when time_col >= 0 */
    int32_t host;            /* 1: rows stay in host memory and stream through the device */
} rj_table_options;
int rj_table_create_ex(rj_ctx* ctx, int64_t n, const float* xy, int32_t ncols, const void* const* cols,
                       const int32_t* types, const rj_table_options* opt, rj_table** out);
/* Rows of a bucketed table that may hold times in [t_lo, t_hi] (all rows
 * without buckets). */
int rj_table_time_rows(const rj_table* t, uint32_t t_lo, uint32_t t_hi, int64_t* row_begin, int64_t* row_end);

/* Resident polygon collections. A collection keeps its snapped polygons,
 * label maps (per accuracy) and grid indexes (per resolution) in device
 * memory across queries; cached structures of collections not in use are
 * evicted when the memory budget runs short. Several collections per
 * context; a collection is used only with the context that created it.
 * Input as rj_set_polygons. */
int rj_polyset_create(rj_ctx* ctx, const float* xy, const int32_t* off, int32_t np, const double* bbox,
                      rj_polyset** out);
int rj_polyset_count(const rj_polyset* ps, int32_t* np);
/* Device bytes the collection holds now (polygons + cached structures). */
int rj_polyset_bytes(const rj_polyset* ps, int64_t* bytes);
int rj_polyset_destroy(rj_polyset* ps);
/* Use a collection as the context's polygons (instead of rj_set_polygons). */
int rj_bind_polyset(rj_ctx* ctx, rj_polyset* ps);

/* DataHandler::setQueryConstraints (DataHandler.cpp:64-66); at most 10. */
int rj_set_constraints(rj_ctx* ctx, const rj_constraint* c, int32_t n);
/* DataHandler::setAggregation (DataHandler.cpp:68-71): type RJ_AGG_*, col = column index. */
int rj_set_aggregation(rj_ctx* ctx, int32_t type, int32_t col);
/* GLHandler::setAccuracyDistance (GLHandler.cpp:250-257) -> GLFunction::setMaxCellSize
 * (GLFunction.cpp:26-30): pixel side = (eps/sqrt 2)/ground_res. ground_res is
 * meters per coordinate unit (1 for projected meters). */
int rj_set_accuracy(rj_ctx* ctx, double eps_m, double ground_res);
/* GLHandler::setPolyIndexResolution (GLHandler.cpp:259-262). The grid is
 * square; ry is accepted and ignored as in the reference (GLFunction.cpp:76). */
int rj_set_index_res(rj_ctx* ctx, int32_t rx, int32_t ry);

/* GLHandler::executeFunction (GLHandler.cpp:188-204). counts: np entries.
 * sums (may be NULL): np doubles, the aggregate sum per polygon (avg = sum /
 * count). legacy3np (may be NULL): the reference's int32 result layout
 * [count | sum/100 | decimal*10] (polygon.frag:19-24, main.cpp:206-209). */
int rj_execute(rj_ctx* ctx, int32_t fn, int64_t* counts, double* sums, int32_t* legacy3np);
/* Exact sums of the last execute: sum = hi + lo * 2^-32 per polygon. */
int rj_get_sums_fixed(rj_ctx* ctx, int64_t* hi, uint64_t* lo);
/* GLHandler::getErrorBounds (GLHandler.cpp:206-212): 4*np values
 * [fn | fp | fn1 | fp1] of the last RJ_ERRORBOUNDS execute, summed over each
 * polygon's boundary pixels (pixels an edge touches, dilated by one) whose
 * center is outside (fn) or inside (fp): fn1/fp1 add the pixel's count,
 * fn/fp add ceil(min(a, 1 - a) * count) with a the polygon's exact area share
 * of the pixel (line-bound.frag:178-183, 247-257). */
int rj_error_bounds(rj_ctx* ctx, int64_t* out4np);
/* GLHandler::printTimeStats (GLHandler.cpp:214-248): the tab-separated timing
 * line (minimum over executions, ms). */
int rj_time_stats(rj_ctx* ctx, int32_t fn, char* buf, size_t len);

/* One self-contained query (no reference counterpart: the reference runs one
 * query at a time through GLHandler's state). row_end < 0 reads to the end of
 * the table; index_res 0 leaves the grid resolution to the engine (256 unless
 * rj_autotune chose another with the same results); ground_res 0 means 1. */
typedef struct rj_query {
    rj_polyset* polys;
    int32_t fn;        /* RJ_RASTER, RJ_INDEX, RJ_HYBRID or RJ_ERRORBOUNDS */
    int32_t agg_type;  /* RJ_AGG_* */
    int32_t agg_col;
    int32_t ncons;
    rj_constraint cons[RJ_MAX_CONSTRAINTS];
    double eps_m;
    double ground_res;
    int32_t index_res;
    int32_t reserved;
    int64_t row_begin, row_end;
} rj_query;

/* Where a query's results go; every pointer may be NULL. */
typedef struct rj_result {
    int64_t* counts;   /* np */
    double* sums;      /* np: aggregate sums */
    int64_t* sum_hi;   /* np: exact sums = hi + lo * 2^-32 */
    uint64_t* sum_lo;  /* np */
    int64_t* bounds;   /* 4*np, RJ_ERRORBOUNDS only */
    int32_t status;    /* RJ_OK or the query's error code */
    int32_t path;      /* RJ_PATH_*: how it ran */
} rj_result;
enum { RJ_PATH_SINGLE = 0, RJ_PATH_FUSED = 1 };

/* One query with the single-query kernels. */
int rj_execute_query(rj_ctx* ctx, rj_table* t, const rj_query* q, rj_result* out);
/* rj_execute_query without allocating: no device or pinned host memory is
 * allocated or freed (no hipMalloc, hipHostMalloc, hipFree or allocating
 * library routine). No counterpart in the reference; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The per-call buffers (per-polygon results,
 * the hybrid border queue and its counter) live in a static per-device scratch
 * array of 32 MiB that the library holds until the process ends; every call on
 * a device shares it, so calls that overlap on one device race (one context
 * serializes its own calls). Limits: a device-resident table, RJ_RASTER
 * (label-map path), RJ_HYBRID or RJ_INDEX, at most 65536 polygons
 * (RJ_ERR_LIMIT beyond); RJ_ERRORBOUNDS and the canvas path return
 * RJ_ERR_UNSUPPORTED. The query's polygon structures must be resident (run it
 * once with rj_execute_query, caching on), else RJ_ERR_STATE. A border queue
 * larger than the scratch's room reruns the pass with inline exact tests.
 * Results equal rj_execute_query's exactly. */
int rj_execute_query_noalloc(rj_ctx* ctx, rj_table* t, const rj_query* q, rj_result* out);
/* A batch of queries over one table: queries that can share a scan run in
 * fused passes (up to 64 queries per pass, one read of the union of their
 * rows); errorbounds, the canvas path and queries over
 * overlapping polygons run alone. Results equal rj_execute_query's exactly.
 * Returns RJ_OK when the batch ran; out[i].status holds each query's code. */
int rj_execute_batch(rj_ctx* ctx, rj_table* t, int32_t nq, const rj_query* qs, rj_result* out);
/* Autotuner of the fused pass: times label tile sizes (8/16/32 pixels), rows
 * per thread, workgroup counts, chunk sizes and, for queries with index_res 0,
 * grid resolutions on this batch, checks that every candidate returns the
 * same results, and keeps the fastest in the context. chosen (may be NULL):
 * 5 values, tile shift, rows per thread, workgroups, rows per chunk,
 * automatic grid resolution. */
int rj_autotune(rj_ctx* ctx, rj_table* t, int32_t nq, const rj_query* qs, int32_t* chosen);

/* Serving loop (no reference counterpart; the nearest is the queue left as a
 * TODO at GLHandler.cpp:195). A thread admits submitted queries while device
 * memory and pass slots allow (memory model: the label maps, grids and results
 * a query would add), runs them in fused passes that overlap with the
 * preparation of the next, and fills each query's rj_result. Buffers named by
 * the rj_result must stay valid until rj_server_poll reports the query done.
 * While a server runs, its context must not be used by other threads except
 * through rj_server_* and rj_polyset_create. */
typedef struct rj_server_config {
    int32_t max_batch;    /* queries per fused pass, 1..64 (0: 64) */
    int32_t max_pending;  /* queued queries before rj_server_submit refuses (0: 65536) */
    int64_t mem_mb;       /* device memory for admitted queries (0: the context budget) */
} rj_server_config;
typedef struct rj_server_stats {
    int64_t submitted, completed, failed, batches;
    double mean_batch;       /* queries per fused pass */
    double p50_ms, p99_ms, max_ms;  /* submit -> result latency of completed queries */
    double qps;              /* completed / (last completion - first submit) */
} rj_server_stats;
int rj_server_create(rj_ctx* ctx, rj_table* t, const rj_server_config* cfg, rj_server** out);
int rj_server_submit(rj_server* s, const rj_query* q, rj_result* out, uint64_t* ticket);
/* 1: done (rj_result filled), 0: still pending after wait_ms, < 0: error. */
int rj_server_poll(rj_server* s, uint64_t ticket, int32_t wait_ms);
int rj_server_stats_get(rj_server* s, rj_server_stats* st);
/* Resets the latency record (counters stay). */
int rj_server_stats_reset(rj_server* s);
/* Finishes every admitted query, then stops the thread. */
int rj_server_destroy(rj_server* s);

/* The CPUJoin grid index of a collection at resolution res as the engine
 * builds it on the device (GridIndexF::buildGrid, CPUJoin/GridIndexF.cpp:60-155;
 * the reference's own GPU index is GLFunction::createPolyIndex,
 * GLFunction.cpp:65-156): cell_off (res*res + 1 entries, cell = y*res + x) and
 * the candidate lists cand (*ncand entries, polygons ascending within a cell),
 * and the *ndef cells whose list misses a polygon that meets the cell
 * (deficient, ascending). Builds the grid when it is not resident. Arrays may
 * be NULL (then only the sizes are returned). A test and measurement aid. */
int rj_polyset_grid(rj_ctx* ctx, rj_polyset* ps, int32_t res, int32_t* cell_off, int32_t* cand, int64_t* ncand,
                    int32_t* deficient, int64_t* ndef);

/* Canvas geometry of the current polygons + accuracy (RasterJoin.cpp:51-78). */
int rj_canvas_info(rj_ctx* ctx, int32_t* W, int32_t* H);

/* Tuning and path selection (no reference counterpart).
 *   RJ_OPT_RASTER_PATH:     0 auto, 1 canvas splat + polygon scan, 2 polygon label map
 *                           (raster and errorbounds joins; auto takes the label map
 *                           unless the polygons overlap on the canvas)
 *   RJ_OPT_TIMING:          1 records device time per kernel of each execute with HIP events;
 *                           2 only the kernels that stream the table (point passes, fused
 *                           passes), totals kept over executes until the option is set again
 *   RJ_OPT_HYBRID_DILATION: border dilation radius in pixels, 0 = derived from the
 *                           float error bound of the exact test (at least 1)
 *   RJ_OPT_CACHE:           1 (default) keeps polygon-side structures (snapped
 *                           polygons, label maps, grid index) across executes;
 *                           0 rebuilds them on every execute as the reference does
 *   RJ_OPT_ORGANIZE:        1 (default) rj_set_points orders rows spatially
 *   RJ_OPT_TILE_SHIFT:      label tile side 2^v pixels, v in 3..5; 0 (default): 8 px for
 *                           raster label maps, 16 px for hybrid maps
 *   RJ_OPT_PASS_WORKGROUPS: workgroups of a single-query point pass; 0 (default): one per
 *                           12288 rows of a spatially ordered table (4096..49152), else 4096
 *   RJ_OPT_BATCH_ROWS:      rows per thread of the fused label-map and index kernels, 4 (default) or 8
 *   RJ_OPT_BATCH_HYBRID_ROWS: rows per thread of the fused hybrid kernel, 4 (default) or 8
 *   RJ_OPT_BATCH_SHARE:     1 (default) fused raster queries on the same label map (same
 *                           collection and accuracy) compute each row's label once per chunk;
 *                           0 every query computes its own
 *   RJ_OPT_BATCH_WORKGROUPS: persistent workgroups of fused passes (64..65536)
 *   RJ_OPT_BATCH_CHUNK_ROWS: rows per chunk of fused passes (1024 .. 2^22)
 *   RJ_OPT_GRAPH:           1 (default) fused passes replay a captured HIP graph
 *   RJ_OPT_INDEX_PATH:      0 (default) index joins run the hybrid pass (label map +
 *                           exact test of the points near an edge: the same CPUJoin
 *                           result); 1 the exact test of every point over its grid
 *                           cell's candidates, as IndexJoin (index.glsl:186-249)
 *   RJ_OPT_EXACT_QUEUE:     1 (default) single-query hybrid passes queue their border
 *                           points for a separate exact pass; 0 tests them inline
 *   RJ_OPT_STREAM_ROWS:     rows per streamed chunk of host tables (0, the default:
 *                           32M rows or what two chunks of the budget left hold)
 *   RJ_OPT_GRID_DIRECT:     1: grid builds test every cell of a polygon's box against
 *                           all its edges; 0 (default): per-vertex and per-edge marks
 *                           (the same lists; grids already built are kept) */
enum { RJ_OPT_RASTER_PATH = 1, RJ_OPT_TIMING = 2, RJ_OPT_HYBRID_DILATION = 3, RJ_OPT_CACHE = 4,
       RJ_OPT_ORGANIZE = 5, RJ_OPT_TILE_SHIFT = 6, RJ_OPT_PASS_WORKGROUPS = 7, RJ_OPT_BATCH_ROWS = 8,
       RJ_OPT_BATCH_WORKGROUPS = 9, RJ_OPT_BATCH_CHUNK_ROWS = 10, RJ_OPT_GRAPH = 11,
       RJ_OPT_INDEX_PATH = 12, RJ_OPT_EXACT_QUEUE = 13, RJ_OPT_STREAM_ROWS = 14, RJ_OPT_BATCH_HYBRID_ROWS = 15,
       RJ_OPT_BATCH_SHARE = 16, RJ_OPT_GRID_DIRECT = 17 };
int rj_set_option(rj_ctx* ctx, int32_t key, int64_t value);

/* Per-kernel device time of the last execute (timing 1), or of every execute
 * since the option was set (timing 2): name i, total ms and launches. Returns
 * the number of kernels recorded, or < 0. */
int rj_kernel_times(rj_ctx* ctx, int32_t i, char* name, size_t name_len, double* ms, int32_t* launches);

/* Points the last rj_execute queued for the separate exact pass (the border
 * points of a hybrid or label-accelerated index join that the exact kernel
 * tests against the polygons' edges), or -1 when it queued none (other join
 * types, the exact test inline). A measurement aid with no reference
 * counterpart: the work of the border pass. */
int64_t rj_exact_points(rj_ctx* ctx);

/* Synchronize the context's stream. */
int rj_sync(rj_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
