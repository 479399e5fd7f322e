// engine.hpp -- the join engine behind the C ABI: device buffers under a
// memory budget, resident point tables, resident polygon collections with
// their per-canvas and per-grid structures, single-query execution and fused
// batches.
//
// One rj_ctx plays the role of the reference's GLHandler + GLFunction objects
// (RasterJoin/join/GLHandler.cpp, GLFunction.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "rasterjoin.h"
#include "rj_impl.hpp"

namespace rj {

// Device memory of a context: the budget (the reference's gpuMemMB,
// GLFunction.cpp:177-192) and a pool of freed buffers by size class, which
// alloc reuses before calling hipMalloc. A first query on a new collection
// allocates a few dozen buffers, and hipMalloc / hipFree take 40-130 us each.
// Pooled blocks count against the budget and are freed when a buffer would
// not fit otherwise. Every buffer of a context is used on the context's
// stream (or behind a synchronisation of its copy stream), so a block
// released while queued work still reads it is only written again by work
// queued after that work.
struct Budget {
    int64_t limit = 0;   // bytes; <= 0: unlimited
    int64_t used = 0;    // bytes of live buffers
    int64_t cached = 0;  // bytes of pooled blocks
    std::multimap<size_t, void*> pool;
    Budget() = default;
    Budget(const Budget&) = delete;
    Budget& operator=(const Budget&) = delete;
    ~Budget() { trim(0); }
    void* take(size_t cls);          // a pooled block of size class cls, or null
    bool put(void* p, size_t cls);   // false: the block is not pooled (the caller frees it)
    void trim(int64_t keep);         // free pooled blocks until at most `keep` bytes stay
};

// Allocation size classes: powers of two up to 1 MiB, then eight classes per
// power of two (at most 12.5 % above the request).
size_t size_class(size_t bytes);

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    std::shared_ptr<Budget> budget;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Polygons of one collection on one canvas (one accuracy).
struct Canvas {
    CanvasPlan plan;
    PolyRaster pr;
    DevBuf d_pc, d_XY;  // regions + bands; snapped vertices X then Y
    int32_t *X = nullptr, *Y = nullptr, *region = nullptr;
    Band *bands_sum = nullptr, *bands_bounds = nullptr;
    // raster label map. lab_state: 0 not checked, 1 polygons disjoint on the
    // canvas, 2 some pixel in two polygons (the label map cannot be used)
    bool lab_built = false;
    int lab_state = 0;
    int lab16 = 0;
    int lab_tshift = kTileShiftDefault;
    DevBuf d_lab, d_tiles;  // tiles buffer ends with 2 overlap counters
    // hybrid label map (border pixels marked); depends on the grid resolution
    // through the cells whose candidate lists are incomplete
    int hyb_res = -1;
    int hyb16 = 0;
    int hyb_tshift = kTileShiftDefault;
    bool hyb_halo = false;  // border = touched pixels + the halo test (else dilated)
    double hyb_border = 0.125;  // share of a pass's rows last queued for the exact test
    float hyb_delta = 0;    // halo width in pixels
    DevBuf d_hlab, d_htiles, d_mask, d_rects;
    // error-bounds label map: the raster label map with every boundary pixel
    // (touched by an edge, dilated by one) flagged, for errorbounds without a
    // canvas splat; blab_state as lab_state
    bool blab_built = false;
    int blab_state = 0;
    int blab16 = 0;
    int blab_tshift = kTileShiftDefault;
    DevBuf d_blab, d_btiles, d_bmask;
    bool bb_built = false;  // error-bounds bands planned and uploaded (bands_bounds)
    DevBuf d_bb;
    uint64_t last_use = 0;
    size_t device_bytes() const;
};

struct Grid {
    GridIndex g;
    int np = 0;         // polygons of the collection
    int64_t ncand = 0;  // entries of the candidate lists
    DevBuf d_cell_off, d_cand, d_bands, d_edges;  // d_bands: band_lo | band_hi | pband | band_off
    uint64_t last_use = 0;
    size_t device_bytes() const { return d_cell_off.bytes + d_cand.bytes + d_bands.bytes + d_edges.bytes; }
};

// One query as the engine runs it.
struct QuerySpec {
    rj_polyset* P = nullptr;
    int fn = RJ_RASTER;
    int exec = RJ_RASTER;  // the pass that runs it: fn, or RJ_HYBRID for an index join
                           // answered by the label-accelerated exact classification
    int agg_type = RJ_AGG_COUNT, agg_col = -1;
    std::vector<rj_constraint> cons;
    double eps = 100.0, ground_res = 1.0;
    int index_res = 256;
    int64_t rb = 0, re = 0;  // rows of the table (before time narrowing)
};

// Largest collection whose per-polygon sums a point pass keeps in LDS indexed
// by polygon; larger ones use hashed slots.
int lds_direct_cap(bool avg);
// Accuracy of the canvas a query runs on (hybrid passes: 8192 pixels on the
// long side of the collection's bbox).
double canvas_eps(const QuerySpec& s);
// Looks up, once per device (current), the device address of the no-alloc
// entry point's scratch array (noalloc.hip).
hipError_t noalloc_scratch_setup(int device);

// Results of one query: counts, exact sums hi + lo * 2^-32, error bounds.
struct QueryOut {
    std::vector<int64_t> counts, hi, bounds;
    std::vector<uint64_t> lo;
    int path = 0;  // RJ_PATH_*
};

}  // namespace rj

struct rj_ctx;

// A resident polygon collection (rj_polyset in the C ABI).
struct rj_polyset {
    rj_ctx* ctx = nullptr;
    rj::PolySet ps;
    rj::DevBuf d_xy, d_off, d_prev;
    std::map<double, std::unique_ptr<rj::Canvas>> canvases;  // by pixel side
    std::map<int, std::unique_ptr<rj::Grid>> grids;          // by resolution
    int pins = 0;  // in use by a batch in flight: not evicted
    size_t device_bytes() const;
};

// A resident point table (rj_table in the C ABI).
struct rj_table {
    int device = 0;
    int64_t n = 0;
    int32_t ncols = 0;
    int32_t types[RJ_MAX_COLUMNS] = {0};
    void* xy = nullptr;
    void* cols[RJ_MAX_COLUMNS] = {nullptr};
    float bbox[4] = {0, 0, 0, 0};  // bounds of the finite locations
    bool nonfinite = false;         // some location is not finite
    uint32_t float_in_range = 0;    // bit c: float column c is finite and in [-2^31, 2^31)
    bool organized = false;
    // time buckets: rows are ordered by (bucket, Morton key); bucket b holds
    // rows [bucket_off[b], bucket_off[b+1]). bucket(t) = (t - t0) / span,
    // clamped to [0, nbuckets-1] (so it is monotone in t).
    int32_t time_col = -1;
    uint32_t t0 = 0, span = 0;
    std::vector<int64_t> bucket_off;
    // host-resident (out-of-core) table: the rows stay in host memory
    // (xy, then each column, at hoff[]) and stream through the device in
    // chunks per pass; hdata is pinned unless hipHostMalloc refused it
    bool host = false;
    char* hdata = nullptr;
    bool hpinned = false;
    size_t hoff[RJ_MAX_COLUMNS + 1] = {0};
    size_t row_bytes() const;
    ~rj_table();
    size_t bytes() const;
    // Rows that may hold times in [lo, hi] (all rows without buckets).
    void time_rows(uint32_t lo, uint32_t hi, int64_t* b, int64_t* e) const;
};

struct rj_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::shared_ptr<rj::Budget> budget = std::make_shared<rj::Budget>();
    int inmem = 0;
    std::string err;
    std::recursive_mutex mu;  // entry points and the serving thread

    // polygons: rj_set_polygons (owned) or a bound collection
    std::unique_ptr<rj_polyset> own_polys;
    rj_polyset* polys = nullptr;
    std::set<rj_polyset*> polysets;  // every collection of this context

    // points
    rj_table* table = nullptr;
    bool own_table = false;
    int64_t rb = 0, re = 0;

    // query state of the GLHandler-style interface
    std::vector<rj_constraint> cons;
    int agg_type = RJ_AGG_COUNT, agg_col = -1;
    double eps = 100.0, ground_res = 1.0;
    int index_res = 256;

    // options
    int raster_path = 0;
    bool timing = false;
    bool timing_pass_only = false;  // timing 2: point passes only, totals kept across executes
    int64_t last_exact = -1;        // points the last single query queued for the exact pass
    int hybrid_dilation = 0;
    bool cache = true;
    bool organize = true;
    int opt_cap = 0;                       // workgroups of a single-query pass (0: by rows, kPassRowsPerWorkgroup)
    int opt_tshift = 0;  // label tile side 2^tshift (0: 8 px for raster maps, 16 px for hybrid maps)
    int raster_tshift() const { return opt_tshift ? opt_tshift : 3; }
    int hybrid_tshift() const { return opt_tshift ? opt_tshift : rj::kTileShiftDefault; }
    int opt_batch_rows = 4;                // rows per thread of the fused label-map and index kernels
    int opt_batch_hybrid_rows = 4;         // rows per thread of the fused hybrid kernel
    bool opt_batch_share = true;           // fused members on one label map share its lookups
    bool opt_grid_direct = false;          // grid build: every cell against every edge
    int opt_batch_wg = 2048;               // persistent workgroups of a fused pass
    int opt_batch_chunk_rows = 8192;       // rows per chunk of a fused pass
    bool opt_graph = true;                 // fused passes replay a captured HIP graph
    int opt_auto_res = 256;                // grid resolution of queries that leave it open
    int opt_index_path = 0;                // 0: index joins by the hybrid pass, 1: per-point grid PIP
    void set_exec(rj::QuerySpec& s) const { s.exec = (s.fn == RJ_INDEX && opt_index_path == 0) ? RJ_HYBRID : s.fn; }

    // canvas (canvas path, errorbounds)
    rj::DevBuf d_cnt, d_chi, d_clo;
    // single-query results
    rj::DevBuf d_res;
    rj::DevBuf d_queue;  // hybrid border queue (float4 entries, then its counter)
    bool opt_queue = true;
    // out-of-core streaming: two pinned staging buffers, two device chunks
    int64_t opt_stream_rows = 0;  // rows per chunk (0: from the budget)
    void* h_chunk[2] = {nullptr, nullptr};
    size_t h_chunk_bytes[2] = {0, 0};
    rj::DevBuf d_chunk[2];
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_h2d[2] = {nullptr, nullptr}, ev_pass[2] = {nullptr, nullptr};
    int stream_passes = 0;  // chunks of the last execute
    unsigned long long* h_res = nullptr;  // pinned
    size_t h_res_bytes = 0;
    void* h_stage = nullptr;  // pinned staging of polygon uploads
    size_t h_stage_bytes = 0;
    hipEvent_t stage_ev = nullptr;
    bool stage_busy = false;
    rj::QueryOut last;
    int last_fn = -1;

    // fused passes: two slots (one being filled while the other runs)
    struct PassSlot {
        rj::DevBuf d_block, d_res;
        rj::BatchBlock* h_block = nullptr;  // pinned
        unsigned long long* h_res = nullptr;  // pinned
        size_t h_res_cap = 0;
        hipGraphExec_t graph = nullptr;
        const void* graph_key[2] = {nullptr, nullptr};
        int graph_rows = 0, graph_wg = 0;
        int64_t graph_nrows = 0;
        size_t graph_lds = 0;
        unsigned graph_classes = 0;
        int graph_hrows = 0;
        hipEvent_t done = nullptr;
        hipEvent_t blk = nullptr;  // host-table passes: the last block upload
        std::vector<int> members;  // query indices of the pass
        std::vector<int64_t> res_off;
        std::vector<uint32_t> qcap;  // border-queue capacity per member (0: inline test)
        std::vector<int64_t> qrows;  // rows each member reads
        std::vector<rj::Canvas*> qcanvas;
        rj::DevBuf d_queue;
        int64_t res_words = 0;
        bool busy = false;
    } slots[2];

    // timing
    struct Launch {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Launch> launches;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    struct KStat {
        std::string name;
        double ms;
        int launches;
    };
    std::vector<KStat> kstats;
    struct Times {
        std::vector<double> execute, ptMem, ptRender, polyMem, polyRender, setup, tri, polyIndex, backend;
        int64_t pts = 0;
        int np = 0;
        int passes = 1;
        int ncons = 0;
    } times[4];
    double cur_ptmem = 0, cur_polymem = 0, cur_setup = 0, cur_polyindex = 0;
    uint64_t clock = 0;  // LRU stamps of cached structures

    ~rj_ctx();

    int fail(int code, const std::string& m);
    int hip_fail(hipError_t e, const char* what);
    int alloc(rj::DevBuf& b, size_t bytes, const char* what);
    int upload(rj::DevBuf& b, const void* src, size_t bytes, const char* what);
    int pinned(void** p, size_t* cur, size_t bytes, const char* what);
    hipEvent_t event();
    // the kernels that stream the table (point passes, splats, fused passes)
    static bool pass_kernel(const std::string& name)
    {
        return name.rfind("points", 0) == 0 || name.rfind("splat", 0) == 0 || name.rfind("batch", 0) == 0;
    }
    template <class F>
    int launch(const char* name, F f)
    {
        hipEvent_t a = nullptr, b = nullptr;
        const bool timed = timing && (!timing_pass_only || pass_kernel(name));
        if (timed) {
            a = event();
            b = event();
            (void)hipEventRecord(a, stream);
        }
        f();
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, name);
        if (timed) {
            (void)hipEventRecord(b, stream);
            launches.push_back({name, a, b});
        }
        return RJ_OK;
    }
    void collect_times(double* pt_ms, double* poly_ms);

    // polygon structures (cached per collection)
    int make_polyset(const float* xy, const int32_t* off, int32_t np, const double* bbox,
                     std::unique_ptr<rj_polyset>* out);
    int canvas_for(rj_polyset* P, double eps_m, double gres, rj::Canvas** out);
    int grid_for(rj_polyset* P, int res, rj::Grid** out);
    // builds grid G of P at resolution res on the device (gridbuild.hip)
    int grid_device(rj_polyset* P, int res, rj::Grid* G);
    int raster_labels(rj_polyset* P, rj::Canvas* C, bool* fresh);
    int bounds_labels(rj_polyset* P, rj::Canvas* C);
    int bounds_for(rj_polyset* P, rj::Canvas* C);
    int hybrid_labels(rj_polyset* P, rj::Canvas* C, rj::Grid* G);
    int build_labels(rj_polyset* P, rj::Canvas* C, void* lab, int label16, rj::DevBuf& tiles, int tshift,
                     bool count);
    int hybrid_radius(const rj::CanvasPlan& plan) const;
    static uint32_t queue_cap(const rj::Canvas* C, int64_t rows);
    double hybrid_delta(const rj::CanvasPlan& plan) const;
    void hybrid_args(const rj::Canvas* C, rj::PointArgs* pa) const;
    int make_room(int64_t bytes);
    void forget_polyset(rj_polyset* P);

    // queries
    int build_query(const rj_table* t, const rj::QuerySpec& s, rj::DevQuery* q);
    void narrow_rows(const rj_table* t, const rj::QuerySpec& s, int64_t* b, int64_t* e) const;
    void interior_rows(const rj_table* t, const rj::QuerySpec& s, int64_t* b, int64_t* e) const;
    rj::DevCols dev_cols(const rj_table* t) const;
    rj::DevPolys dev_polys(const rj_polyset* P, const rj::Canvas* C) const;
    rj::DevGrid dev_grid(const rj::Grid* G) const;
    int ensure_canvas_buffers(const rj::CanvasPlan& plan, bool sums);
    // One point pass over rows [rb, re) of t (streamed in chunks for a host table).
    int point_pass(const char* name, int mode, rj::PointArgs pa, rj_table* t, int64_t rb, int64_t re);
    // streaming of host tables through the two device chunk buffers
    struct StreamLayout {
        int64_t chunk = 0;  // rows per chunk
        int nparts = 0;
        int part_of[RJ_MAX_COLUMNS + 1];  // table part (0: x,y; c + 1: column c) of each chunk part
        size_t off[RJ_MAX_COLUMNS + 1], width[RJ_MAX_COLUMNS + 1];
        size_t total = 0;
    };
    int stream_setup(rj_table* t, const bool* need, int64_t rows, StreamLayout* L);
    int stream_copy(rj_table* t, const StreamLayout& L, int k, int64_t r, int64_t n);
    void stream_args(const rj_table* t, const StreamLayout& L, int k, int64_t r0, rj::PointArgs* a) const;
    int64_t estimate_bytes(const rj::QuerySpec& s) const;

    // the single-query path (plain kernels); out receives the results
    int run_single(rj_table* t, const rj::QuerySpec& s, rj::QueryOut* out);
    // GLHandler::executeFunction on the context's state
    int execute(int fn, int64_t* out_counts, double* out_sums, int32_t* legacy);

    // fused batches
    // builds the polygon structures of a batch and pins its collections (the
    // caller unpins them once the batch is done)
    int batch_prepare(rj_table* t, const rj::QuerySpec* qs, int nq, std::vector<int>* fused,
                      std::vector<int>* alone, std::vector<int>* status);
    int batch_launch(int slot, rj_table* t, const rj::QuerySpec* qs, const std::vector<int>& members);
    static constexpr int64_t kQueueEntriesMax = 1LL << 28;  // border-queue entries of one fused pass
    void group_passes(rj_table* t, const rj::QuerySpec* qs, const std::vector<int>& fused, int max_batch,
                      std::vector<std::vector<int>>* passes);
    void batch_segments(rj::BatchBlock& B, const std::vector<int64_t>& qb, const std::vector<int64_t>& qe, int64_t lo,
                        int64_t hi) const;
    int batch_stream(PassSlot& S, rj_table* t, const std::vector<int64_t>& qb, const std::vector<int64_t>& qe,
                     uint32_t* counter, size_t lds, unsigned classes);
    int batch_finish(int slot, rj_table* t, const rj::QuerySpec* qs, rj::QueryOut* outs);
    // every query of the batch: fused passes, then the queries that run alone;
    // status[i] receives each query's own code
    int execute_batch(rj_table* t, const rj::QuerySpec* qs, int nq, rj::QueryOut* outs, int* status);
    // autoq: queries whose grid resolution is left to the tuner
    int autotune(rj_table* t, const rj::QuerySpec* qs, int nq, const std::vector<int>& autoq, int32_t* chosen);
};

// QuerySpec from the C ABI query record (spec_check reports the error in
// `why` and leaves the context untouched).
int spec_check(const rj_ctx* ctx, const rj_table* t, const rj_query* q, rj::QuerySpec* s, std::string* why);
int spec_from_query(rj_ctx* ctx, const rj_table* t, const rj_query* q, rj::QuerySpec* s);
// Copy a query's results to the caller's rj_result buffers.
void rj_fill_result(const rj::QueryOut& o, int np, int fn, int agg, rj_result* r);
