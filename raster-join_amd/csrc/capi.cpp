// capi.cpp -- the C ABI of include/rasterjoin.h over the engine (engine.cpp):
// argument checks, point tables, polygon collections, the GLHandler-style
// state setters, single queries, batches and the autotuner. The serving loop
// is in server.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace rj {
size_t organize_scratch_bytes(int64_t n);
int float_columns_in_range(hipStream_t s, int64_t n, void* const* cols, const int32_t* types, int ncols,
                           uint32_t* flags, uint32_t* mask);
int organize_rows(hipStream_t s, int64_t n, float* xy, void** cols, const int32_t* types, int ncols,
                  const float bbox[4], const uint32_t* time, uint32_t t0, uint32_t span, uint32_t nb,
                  int64_t* off_dev, void* scratch);
}  // namespace rj

using namespace rj;

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

size_t type_bytes(int t) { return t == RJ_ATTR_UINT8 ? 1 : 4; }

constexpr uint32_t kMaxBuckets = 1u << 20;

void table_bbox(rj_table* t, int64_t n, const float* xy);

int table_fill(rj_ctx* ctx, rj_table* t, int64_t n, const float* xy, int32_t ncols, const void* const* cols,
               const int32_t* types)
{
    t->n = n;
    t->ncols = ncols;
    hipError_t e = hipMalloc(&t->xy, std::max<size_t>(16, (size_t)n * 8));
    if (e != hipSuccess) return ctx->hip_fail(e, "point table");
    if (n) e = hipMemcpyAsync(t->xy, xy, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream);
    for (int c = 0; c < ncols && e == hipSuccess; c++) {
        t->types[c] = types[c];
        size_t b = (size_t)n * type_bytes(types[c]);
        e = hipMalloc(&t->cols[c], std::max<size_t>(16, b));
        if (e == hipSuccess && n) e = hipMemcpyAsync(t->cols[c], cols[c], b, hipMemcpyHostToDevice, ctx->stream);
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "point table");
    table_bbox(t, n, xy);
    bool any_float = false;
    for (int c = 0; c < ncols; c++) any_float = any_float || types[c] == RJ_ATTR_FLOAT;
    if (any_float) {
        uint32_t* flags = nullptr;
        e = hipMalloc((void**)&flags, (size_t)ncols * 4);
        if (e != hipSuccess) return ctx->hip_fail(e, "column ranges");
        const int r = float_columns_in_range(ctx->stream, n, t->cols, t->types, ncols, flags, &t->float_in_range);
        (void)hipFree(flags);
        if (r) return ctx->hip_fail((hipError_t)r, "column ranges");
    }
    return RJ_OK;
}

// Bounds of the finite points (0..1 when there are none) and whether some
// point is not finite; large tables on 8 host threads.
void table_bbox(rj_table* t, int64_t n, const float* xy)
{
    struct Part {
        float mn[2], mx[2];
        bool any = false, bad = false;
    };
    const int nth = n >= (1 << 22) ? 8 : 1;
    std::vector<Part> part(nth);
    auto run = [&](int k) {
        Part& P = part[k];
        for (int64_t i = n * k / nth; i < n * (k + 1) / nth; i++) {
            const float x = xy[2 * i], y = xy[2 * i + 1];
            if (!std::isfinite(x) || !std::isfinite(y)) {
                P.bad = true;
                continue;
            }
            if (!P.any) {
                P.mn[0] = P.mx[0] = x; P.mn[1] = P.mx[1] = y;
                P.any = true;
            }
            P.mn[0] = std::min(P.mn[0], x); P.mx[0] = std::max(P.mx[0], x);
            P.mn[1] = std::min(P.mn[1], y); P.mx[1] = std::max(P.mx[1], y);
        }
    };
    if (nth == 1) {
        run(0);
    } else {
        std::vector<std::thread> th;
        for (int k = 0; k < nth; k++) th.emplace_back(run, k);
        for (auto& x : th) x.join();
    }
    float mn[2] = {0, 0}, mx[2] = {1, 1};
    bool first = true, bad = false;
    for (const Part& P : part) {
        bad = bad || P.bad;
        if (!P.any) continue;
        if (first) {
            mn[0] = P.mn[0]; mn[1] = P.mn[1]; mx[0] = P.mx[0]; mx[1] = P.mx[1];
            first = false;
        }
        mn[0] = std::min(mn[0], P.mn[0]); mx[0] = std::max(mx[0], P.mx[0]);
        mn[1] = std::min(mn[1], P.mn[1]); mx[1] = std::max(mx[1], P.mx[1]);
    }
    t->bbox[0] = mn[0]; t->bbox[1] = mn[1]; t->bbox[2] = mx[0]; t->bbox[3] = mx[1];
    t->nonfinite = bad;
}

// A host-resident table: the rows are copied into one host allocation
// (pinned when the driver grants it), xy then each column, and stream
// through the device per query. With a time column the rows are put in
// bucket order (a stable counting sort) so that a time-constrained query
// streams only its buckets' rows.
int host_fill(rj_ctx* ctx, rj_table* t, int64_t n, const float* xy, int32_t ncols, const void* const* cols,
              const int32_t* types)
{
    t->n = n;
    t->ncols = ncols;
    t->host = true;
    const int nparts = ncols + 1;
    size_t w[RJ_MAX_COLUMNS + 1];
    size_t total = 0;
    for (int part = 0; part < nparts; part++) {
        if (part) t->types[part - 1] = types[part - 1];
        w[part] = part == 0 ? 8 : type_bytes(types[part - 1]);
        t->hoff[part] = total;
        total += ((size_t)n * w[part] + 255) / 256 * 256;
    }
    total = std::max<size_t>(total, 256);
    void* p = nullptr;
    if (hipHostMalloc(&p, total, hipHostMallocDefault) == hipSuccess) {
        t->hpinned = true;
    } else {
        (void)hipGetLastError();
        p = std::malloc(total);
        if (!p) return ctx->fail(RJ_ERR_NOMEM, "host point table (" + std::to_string(total) + " bytes)");
    }
    t->hdata = static_cast<char*>(p);
    auto src = [&](int part) { return part == 0 ? static_cast<const char*>((const void*)xy) : static_cast<const char*>(cols[part - 1]); };
    std::vector<int64_t> perm;  // bucket order of the rows
    if (t->time_col >= 0 && n > 0) {
        const uint32_t* tm = static_cast<const uint32_t*>(cols[t->time_col]);
        uint32_t mx = 0;
        for (int64_t i = 0; i < n; i++) mx = std::max(mx, tm[i]);
        const uint64_t last = mx < t->t0 ? 0 : (uint64_t)(mx - t->t0) / t->span;
        if (last + 1 > kMaxBuckets) return ctx->fail(RJ_ERR_LIMIT, "more than 2^20 time buckets");
        const size_t nb = (size_t)last + 1;
        auto bucket = [&](uint32_t v) -> size_t { return v < t->t0 ? 0 : (size_t)((v - t->t0) / t->span); };
        t->bucket_off.assign(nb + 1, 0);
        for (int64_t i = 0; i < n; i++) t->bucket_off[bucket(tm[i]) + 1]++;
        for (size_t b = 0; b < nb; b++) t->bucket_off[b + 1] += t->bucket_off[b];
        std::vector<int64_t> next(t->bucket_off.begin(), t->bucket_off.end() - 1);
        perm.resize((size_t)n);
        for (int64_t i = 0; i < n; i++) perm[(size_t)next[bucket(tm[i])]++] = i;
    }
    // one host thread per part
    std::vector<std::thread> th;
    for (int part = 0; part < nparts; part++)
        th.emplace_back([&, part] {
            char* dst = t->hdata + t->hoff[part];
            const char* sp = src(part);
            if (perm.empty()) {
                if (n) std::memcpy(dst, sp, (size_t)n * w[part]);
                return;
            }
            for (int64_t i = 0; i < n; i++) std::memcpy(dst + (size_t)i * w[part], sp + (size_t)perm[(size_t)i] * w[part], w[part]);
        });
    for (auto& x : th) x.join();
    return RJ_OK;
}

// Spatial order (and time buckets when the table has a time column).
int table_organize(rj_ctx* ctx, rj_table* t, const uint32_t* host_time)
{
    // the radix sort counts items in an int: larger tables keep their order
    if (t->n < 2 || t->n >= ((int64_t)1 << 31)) {
        if (t->time_col >= 0)
            return ctx->fail(RJ_ERR_LIMIT, "time buckets need 2 .. 2^31-1 rows");
        return RJ_OK;
    }
    uint32_t nb = 0;
    if (t->time_col >= 0) {
        uint32_t mx = 0;
        for (int64_t i = 0; i < t->n; i++) mx = std::max(mx, host_time[i]);
        const uint64_t last = mx < t->t0 ? 0 : (uint64_t)(mx - t->t0) / t->span;
        if (last + 1 > kMaxBuckets) return ctx->fail(RJ_ERR_LIMIT, "more than 2^20 time buckets");
        nb = (uint32_t)last + 1;
    }
    void* scratch = nullptr;
    int64_t* off_dev = nullptr;
    hipError_t e = hipMalloc(&scratch, organize_scratch_bytes(t->n));
    if (e != hipSuccess) return ctx->hip_fail(e, "organize scratch");
    if (nb) {
        e = hipMalloc((void**)&off_dev, ((size_t)nb + 1) * 8);
        if (e != hipSuccess) {
            (void)hipFree(scratch);
            return ctx->hip_fail(e, "bucket offsets");
        }
    }
    const uint32_t* dtime = nb ? static_cast<const uint32_t*>(t->cols[t->time_col]) : nullptr;
    int r = organize_rows(ctx->stream, t->n, (float*)t->xy, t->cols, t->types, t->ncols, t->bbox, dtime, t->t0,
                          t->span, nb, off_dev, scratch);
    if (r == 0 && nb) {
        t->bucket_off.assign((size_t)nb + 1, 0);
        r = (int)hipMemcpyAsync(t->bucket_off.data(), off_dev, ((size_t)nb + 1) * 8, hipMemcpyDeviceToHost,
                                ctx->stream);
    }
    e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(scratch);
    if (off_dev) (void)hipFree(off_dev);
    if (r != 0) return ctx->hip_fail((hipError_t)r, "organize");
    if (e != hipSuccess) return ctx->hip_fail(e, "organize");
    t->organized = true;
    return RJ_OK;
}

int check_columns(rj_ctx* ctx, int64_t n, const float* xy, int32_t ncols, const void* const* cols,
                  const int32_t* types)
{
    if (n < 0 || (n > 0 && !xy)) return ctx->fail(RJ_ERR_ARG, "bad point array");
    if (ncols < 0 || ncols > RJ_MAX_COLUMNS) return ctx->fail(RJ_ERR_LIMIT, "too many columns");
    if (n >= ((int64_t)1 << 32)) return ctx->fail(RJ_ERR_LIMIT, "at most 2^32-1 points per table");
    for (int c = 0; c < ncols; c++) {
        if (!types || types[c] < RJ_ATTR_UINT || types[c] > RJ_ATTR_UINT8)
            return ctx->fail(RJ_ERR_ARG, "bad column type");
        if (n > 0 && (!cols || !cols[c])) return ctx->fail(RJ_ERR_ARG, "missing column");
    }
    return RJ_OK;
}

int make_table(rj_ctx* ctx, int64_t n, const float* xy, int32_t ncols, const void* const* cols, const int32_t* types,
               const rj_table_options* opt, rj_table** out)
{
    int rc = check_columns(ctx, n, xy, ncols, cols, types);
    if (rc) return rc;
    if (opt->time_col >= 0) {
        if (opt->time_col >= ncols || types[opt->time_col] != RJ_ATTR_UINT)
            return ctx->fail(RJ_ERR_ARG, "the time column must be a Uint column");
        if (opt->bucket_seconds == 0) return ctx->fail(RJ_ERR_ARG, "bucket width must be > 0");
    }
    rj_table* t = new rj_table();
    (void)hipGetDevice(&t->device);
    t->time_col = opt->time_col;
    t->t0 = opt->time0;
    t->span = opt->bucket_seconds;
    if (opt->host) {
        rc = host_fill(ctx, t, n, xy, ncols, cols, types);
        if (!rc) table_bbox(t, n, xy);
        if (rc) delete t;
        else *out = t;
        return rc;
    }
    rc = table_fill(ctx, t, n, xy, ncols, cols, types);
    if (!rc && (opt->organize || opt->time_col >= 0))
        rc = table_organize(ctx, t, opt->time_col >= 0 ? static_cast<const uint32_t*>(cols[opt->time_col]) : nullptr);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = ctx->hip_fail(e, "table upload");
    }
    if (rc) {
        delete t;
        return rc;
    }
    *out = t;
    return RJ_OK;
}

void fill_result(const QueryOut& o, int np, int fn, int agg, rj_result* r)
{
    const bool avg = agg != RJ_AGG_COUNT && fn != RJ_ERRORBOUNDS;
    for (int p = 0; p < np; p++) {
        if (r->counts) r->counts[p] = o.counts[p];
        if (r->sums) r->sums[p] = avg ? fixed_to_double(o.hi[p], o.lo[p]) : 0.0;
        if (r->sum_hi) r->sum_hi[p] = avg ? o.hi[p] : 0;
        if (r->sum_lo) r->sum_lo[p] = avg ? o.lo[p] : 0;
    }
    if (r->bounds && fn == RJ_ERRORBOUNDS && o.bounds.size() == 4 * (size_t)np)
        std::copy(o.bounds.begin(), o.bounds.end(), r->bounds);
    r->path = o.path;
}

}  // namespace

void rj_fill_result(const rj::QueryOut& o, int np, int fn, int agg, rj_result* r) { fill_result(o, np, fn, agg, r); }

#define RJ_GUARD(ctx)                                   \
    if (!(ctx)) return RJ_ERR_ARG;                      \
    std::lock_guard<std::recursive_mutex> lk_((ctx)->mu); \
    (void)hipSetDevice((ctx)->device)

extern "C" {

int rj_create(int64_t gpu_mem_mb, int32_t inmem, rj_ctx** out)
{
    if (!out) return RJ_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return RJ_ERR_DEVICE;
    }
    rj_ctx* c = new rj_ctx();
    (void)hipGetDevice(&c->device);
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return RJ_ERR_DEVICE;
    }
    // GLHandler keeps 100 MB of the budget for other data (GLHandler.cpp:51).
    c->budget->limit = gpu_mem_mb > 0 ? (gpu_mem_mb - 100) * 1024LL * 1024LL : 0;
    if (gpu_mem_mb > 0 && c->budget->limit <= 0) c->budget->limit = 1;
    c->inmem = inmem;
    // the no-alloc entry point's per-device scratch address (its lookup
    // error, if any, is reported by rj_execute_query_noalloc)
    (void)rj::noalloc_scratch_setup(c->device);
    *out = c;
    return RJ_OK;
}

int rj_destroy(rj_ctx* ctx)
{
    if (!ctx) return RJ_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return RJ_OK;
}

const char* rj_last_error(const rj_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rj_set_polygons(rj_ctx* ctx, const float* xy, const int32_t* off, int32_t np, const double* bbox)
{
    RJ_GUARD(ctx);
    double t0 = now_ms();
    std::unique_ptr<rj_polyset> P;
    int rc = ctx->make_polyset(xy, off, np, bbox, &P);
    if (rc) return rc;
    if (ctx->own_polys) ctx->forget_polyset(ctx->own_polys.get());
    ctx->own_polys = std::move(P);
    ctx->polys = ctx->own_polys.get();
    ctx->cur_polymem = now_ms() - t0;
    return RJ_OK;
}

int rj_polyset_create(rj_ctx* ctx, const float* xy, const int32_t* off, int32_t np, const double* bbox,
                      rj_polyset** out)
{
    RJ_GUARD(ctx);
    if (!out) return RJ_ERR_ARG;
    std::unique_ptr<rj_polyset> P;
    int rc = ctx->make_polyset(xy, off, np, bbox, &P);
    if (rc) return rc;
    *out = P.release();
    return RJ_OK;
}

int rj_polyset_count(const rj_polyset* ps, int32_t* np)
{
    if (!ps || !np) return RJ_ERR_ARG;
    *np = ps->ps.np;
    return RJ_OK;
}

int rj_polyset_bytes(const rj_polyset* ps, int64_t* bytes)
{
    if (!ps || !bytes) return RJ_ERR_ARG;
    *bytes = (int64_t)ps->device_bytes();
    return RJ_OK;
}

int rj_polyset_destroy(rj_polyset* ps)
{
    if (!ps) return RJ_ERR_ARG;
    rj_ctx* ctx = ps->ctx;
    if (ctx) {
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        if (ps->pins > 0) return ctx->fail(RJ_ERR_STATE, "collection in use by a query in flight");
        if (ctx->own_polys.get() == ps) return ctx->fail(RJ_ERR_ARG, "the context's own polygons");
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->forget_polyset(ps);
    }
    delete ps;
    return RJ_OK;
}

int rj_bind_polyset(rj_ctx* ctx, rj_polyset* ps)
{
    RJ_GUARD(ctx);
    if (!ps || ps->ctx != ctx) return ctx->fail(RJ_ERR_ARG, "collection of another context");
    ctx->polys = ps;
    return RJ_OK;
}

int rj_set_points(rj_ctx* ctx, int64_t n, const float* xy, int32_t ncols, const void* const* cols, const int32_t* types)
{
    RJ_GUARD(ctx);
    int rc = check_columns(ctx, n, xy, ncols, cols, types);
    if (rc) return rc;
    size_t bytes = (size_t)n * 8;
    for (int c = 0; c < ncols; c++) bytes += (size_t)n * type_bytes(types[c]);
    double t0 = now_ms();
    if (ctx->own_table) delete ctx->table;
    ctx->table = nullptr;
    ctx->own_table = false;
    // points that do not fit the memory budget stay in host memory and
    // stream through the device in passes (GLFunction.cpp:329-337)
    const bool host = ctx->budget->limit > 0 && ctx->budget->used + (int64_t)bytes > ctx->budget->limit;
    rj_table_options opt{ctx->organize ? 1 : 0, -1, 0, 0, host ? 1 : 0};
    rj_table* t = nullptr;
    if ((rc = make_table(ctx, n, xy, ncols, cols, types, &opt, &t))) return rc;
    ctx->table = t;
    ctx->own_table = true;
    ctx->rb = 0;
    ctx->re = n;
    ctx->cur_ptmem = now_ms() - t0;
    return RJ_OK;
}

int rj_table_create(rj_ctx* ctx, int64_t n, const float* xy, int32_t ncols, const void* const* cols,
                    const int32_t* types, int32_t organize, rj_table** out)
{
    RJ_GUARD(ctx);
    if (!out) return RJ_ERR_ARG;
    rj_table_options opt{organize, -1, 0, 0, 0};
    return make_table(ctx, n, xy, ncols, cols, types, &opt, out);
}

int rj_table_create_ex(rj_ctx* ctx, int64_t n, const float* xy, int32_t ncols, const void* const* cols,
                       const int32_t* types, const rj_table_options* opt, rj_table** out)
{
    RJ_GUARD(ctx);
    if (!out || !opt) return RJ_ERR_ARG;
    return make_table(ctx, n, xy, ncols, cols, types, opt, out);
}

int rj_table_time_rows(const rj_table* t, uint32_t t_lo, uint32_t t_hi, int64_t* row_begin, int64_t* row_end)
{
    if (!t || !row_begin || !row_end) return RJ_ERR_ARG;
    t->time_rows(t_lo, t_hi, row_begin, row_end);
    return RJ_OK;
}

int rj_table_rows(const rj_table* t, int64_t* n)
{
    if (!t || !n) return RJ_ERR_ARG;
    *n = t->n;
    return RJ_OK;
}

int rj_table_read(const rj_table* t, int64_t row0, int64_t n, float* xy, void* const* cols)
{
    if (!t || row0 < 0 || n < 0 || row0 + n > t->n) return RJ_ERR_ARG;
    if (t->host) {
        if (xy && n) std::memcpy(xy, t->hdata + t->hoff[0] + row0 * 8, (size_t)n * 8);
        for (int c = 0; c < t->ncols && cols; c++) {
            if (!cols[c] || !n) continue;
            const size_t w = type_bytes(t->types[c]);
            std::memcpy(cols[c], t->hdata + t->hoff[c + 1] + row0 * w, (size_t)n * w);
        }
        return RJ_OK;
    }
    (void)hipSetDevice(t->device);
    hipError_t e = hipSuccess;
    if (xy && n) e = hipMemcpy(xy, (const char*)t->xy + row0 * 8, (size_t)n * 8, hipMemcpyDeviceToHost);
    for (int c = 0; c < t->ncols && cols && e == hipSuccess; c++) {
        if (!cols[c] || !n) continue;
        size_t w = type_bytes(t->types[c]);
        e = hipMemcpy(cols[c], (const char*)t->cols[c] + row0 * w, (size_t)n * w, hipMemcpyDeviceToHost);
    }
    return e == hipSuccess ? RJ_OK : RJ_ERR_DEVICE;
}

int rj_table_destroy(rj_table* t)
{
    if (!t) return RJ_ERR_ARG;
    (void)hipSetDevice(t->device);
    (void)hipDeviceSynchronize();
    delete t;
    return RJ_OK;
}

int rj_bind_table(rj_ctx* ctx, rj_table* t, int64_t row_begin, int64_t row_end)
{
    RJ_GUARD(ctx);
    if (!t || row_begin < 0 || row_end < row_begin || row_end > t->n) return ctx->fail(RJ_ERR_ARG, "bad row range");
    if (ctx->own_table) delete ctx->table;
    ctx->table = t;
    ctx->own_table = false;
    ctx->rb = row_begin;
    ctx->re = row_end;
    return RJ_OK;
}

int rj_set_constraints(rj_ctx* ctx, const rj_constraint* c, int32_t n)
{
    RJ_GUARD(ctx);
    if (n < 0 || (n > 0 && !c)) return ctx->fail(RJ_ERR_ARG, "bad constraints");
    if (n > RJ_MAX_CONSTRAINTS) return ctx->fail(RJ_ERR_LIMIT, "at most 10 constraints");
    ctx->cons.assign(c, c + n);
    return RJ_OK;
}

int rj_set_aggregation(rj_ctx* ctx, int32_t type, int32_t col)
{
    RJ_GUARD(ctx);
    if (type < RJ_AGG_COUNT || type > RJ_AGG_SUM) return ctx->fail(RJ_ERR_ARG, "bad aggregation");
    if (type != RJ_AGG_COUNT && col < 0) return ctx->fail(RJ_ERR_ARG, "aggregation needs a column");
    ctx->agg_type = type;
    ctx->agg_col = col;
    return RJ_OK;
}

int rj_set_accuracy(rj_ctx* ctx, double eps_m, double ground_res)
{
    RJ_GUARD(ctx);
    if (!(eps_m > 0) || !(ground_res > 0)) return ctx->fail(RJ_ERR_ARG, "accuracy and ground resolution must be > 0");
    ctx->eps = eps_m;
    ctx->ground_res = ground_res;
    return RJ_OK;
}

int rj_set_index_res(rj_ctx* ctx, int32_t rx, int32_t ry)
{
    RJ_GUARD(ctx);
    (void)ry;
    if (rx <= 0 || rx > 16384) return ctx->fail(RJ_ERR_ARG, "index resolution must be in 1..16384");
    ctx->index_res = rx;
    return RJ_OK;
}

int rj_set_option(rj_ctx* ctx, int32_t key, int64_t value)
{
    RJ_GUARD(ctx);
    switch (key) {
    case RJ_OPT_RASTER_PATH:
        if (value < 0 || value > 2) return ctx->fail(RJ_ERR_ARG, "raster path is 0, 1 or 2");
        ctx->raster_path = (int)value;
        return RJ_OK;
    case RJ_OPT_TIMING:
        if (value < 0 || value > 2) return ctx->fail(RJ_ERR_ARG, "timing 0, 1 or 2");
        ctx->timing = value != 0;
        ctx->timing_pass_only = value == 2;
        ctx->kstats.clear();
        return RJ_OK;
    case RJ_OPT_HYBRID_DILATION:
        if (value < 0 || value > 31) return ctx->fail(RJ_ERR_ARG, "dilation radius is 0 (auto) .. 31");
        if (ctx->hybrid_dilation != (int)value)
            for (rj_polyset* P : ctx->polysets)
                for (auto& c : P->canvases) c.second->hyb_res = -1;
        ctx->hybrid_dilation = (int)value;
        return RJ_OK;
    case RJ_OPT_CACHE: ctx->cache = value != 0; return RJ_OK;
    case RJ_OPT_ORGANIZE: ctx->organize = value != 0; return RJ_OK;
    case RJ_OPT_TILE_SHIFT:
        if (value != 0 && (value < 3 || value > 5)) return ctx->fail(RJ_ERR_ARG, "tile shift is 0 (auto), 3, 4 or 5");
        ctx->opt_tshift = (int)value;
        return RJ_OK;
    case RJ_OPT_PASS_WORKGROUPS:
        if (value < 0 || value > 65536) return ctx->fail(RJ_ERR_ARG, "workgroups 0 (default) .. 65536");
        ctx->opt_cap = (int)value;
        return RJ_OK;
    case RJ_OPT_BATCH_ROWS:
        if (value != 4 && value != 8) return ctx->fail(RJ_ERR_ARG, "rows per thread is 4 or 8");
        ctx->opt_batch_rows = (int)value;
        return RJ_OK;
    case RJ_OPT_BATCH_SHARE: ctx->opt_batch_share = value != 0; return RJ_OK;
    case RJ_OPT_BATCH_HYBRID_ROWS:
        if (value != 4 && value != 8) return ctx->fail(RJ_ERR_ARG, "rows per thread is 4 or 8");
        ctx->opt_batch_hybrid_rows = (int)value;
        return RJ_OK;
    case RJ_OPT_BATCH_WORKGROUPS:
        if (value < 64 || value > 65536) return ctx->fail(RJ_ERR_ARG, "workgroups 64 .. 65536");
        ctx->opt_batch_wg = (int)value;
        return RJ_OK;
    case RJ_OPT_BATCH_CHUNK_ROWS:
        if (value < 1024 || value > (1 << 22)) return ctx->fail(RJ_ERR_ARG, "chunk rows 1024 .. 2^22");
        ctx->opt_batch_chunk_rows = (int)value;
        return RJ_OK;
    case RJ_OPT_GRAPH: ctx->opt_graph = value != 0; return RJ_OK;
    case RJ_OPT_GRID_DIRECT: ctx->opt_grid_direct = value != 0; return RJ_OK;
    case RJ_OPT_EXACT_QUEUE: ctx->opt_queue = value != 0; return RJ_OK;
    case RJ_OPT_STREAM_ROWS:
        if (value < 0 || value > ((int64_t)1 << 32)) return ctx->fail(RJ_ERR_ARG, "stream rows out of range");
        ctx->opt_stream_rows = value;
        return RJ_OK;
    case RJ_OPT_INDEX_PATH:
        if (value < 0 || value > 1) return ctx->fail(RJ_ERR_ARG, "index path is 0 or 1");
        ctx->opt_index_path = (int)value;
        return RJ_OK;
    default: return ctx->fail(RJ_ERR_ARG, "unknown option");
    }
}

int rj_execute(rj_ctx* ctx, int32_t fn, int64_t* counts, double* sums, int32_t* legacy3np)
{
    RJ_GUARD(ctx);
    return ctx->execute(fn, counts, sums, legacy3np);
}

int rj_get_sums_fixed(rj_ctx* ctx, int64_t* hi, uint64_t* lo)
{
    RJ_GUARD(ctx);
    if (ctx->last_fn < 0) return ctx->fail(RJ_ERR_STATE, "no execute yet");
    if (hi) std::copy(ctx->last.hi.begin(), ctx->last.hi.end(), hi);
    if (lo) std::copy(ctx->last.lo.begin(), ctx->last.lo.end(), lo);
    return RJ_OK;
}

int rj_error_bounds(rj_ctx* ctx, int64_t* out4np)
{
    RJ_GUARD(ctx);
    if (ctx->last_fn != RJ_ERRORBOUNDS) return ctx->fail(RJ_ERR_STATE, "last execute was not RJ_ERRORBOUNDS");
    if (out4np) std::copy(ctx->last.bounds.begin(), ctx->last.bounds.end(), out4np);
    return RJ_OK;
}

int rj_execute_query(rj_ctx* ctx, rj_table* t, const rj_query* q, rj_result* out)
{
    RJ_GUARD(ctx);
    if (!t || !q || !out) return ctx->fail(RJ_ERR_ARG, "null table, query or result");
    QuerySpec s;
    int rc = spec_from_query(ctx, t, q, &s);
    if (rc) {
        out->status = rc;
        return rc;
    }
    ctx->launches.clear();
    ctx->pool_used = 0;
    QueryOut o;
    rc = ctx->run_single(t, s, &o);
    out->status = rc;
    if (rc) return rc;
    if (ctx->timing) {
        double a, b;
        ctx->collect_times(&a, &b);
    }
    fill_result(o, s.P->ps.np, s.fn, s.agg_type, out);
    return RJ_OK;
}

int rj_execute_batch(rj_ctx* ctx, rj_table* t, int32_t nq, const rj_query* qs, rj_result* out)
{
    RJ_GUARD(ctx);
    if (!t || nq < 0 || (nq > 0 && (!qs || !out))) return ctx->fail(RJ_ERR_ARG, "bad batch");
    std::vector<QuerySpec> specs(nq);
    std::vector<int> ok;
    for (int i = 0; i < nq; i++) {
        out[i].status = spec_from_query(ctx, t, &qs[i], &specs[i]);
        if (out[i].status == RJ_OK) ok.push_back(i);
    }
    std::vector<QuerySpec> run(ok.size());
    for (size_t k = 0; k < ok.size(); k++) run[k] = specs[ok[k]];
    std::vector<QueryOut> outs(ok.size());
    std::vector<int> st(ok.size(), RJ_OK);
    int rc = ctx->execute_batch(t, run.data(), (int)run.size(), outs.data(), st.data());
    if (rc) return rc;
    for (size_t k = 0; k < ok.size(); k++) {
        rj_result* r = &out[ok[k]];
        r->status = st[k];
        if (st[k] == RJ_OK) fill_result(outs[k], run[k].P->ps.np, run[k].fn, run[k].agg_type, r);
    }
    return RJ_OK;
}

int rj_autotune(rj_ctx* ctx, rj_table* t, int32_t nq, const rj_query* qs, int32_t* chosen)
{
    RJ_GUARD(ctx);
    if (!t || nq <= 0 || !qs) return ctx->fail(RJ_ERR_ARG, "bad batch");
    std::vector<QuerySpec> specs(nq);
    std::vector<int> autoq;
    for (int i = 0; i < nq; i++) {
        int rc = spec_from_query(ctx, t, &qs[i], &specs[i]);
        if (rc) return rc;
        if (qs[i].index_res == 0 && qs[i].fn != RJ_RASTER && qs[i].fn != RJ_ERRORBOUNDS) autoq.push_back(i);
    }
    return ctx->autotune(t, specs.data(), nq, autoq, chosen);
}

int rj_time_stats(rj_ctx* ctx, int32_t fn, char* buf, size_t len)
{
    RJ_GUARD(ctx);
    if (fn < 0 || fn > 3 || !buf) return ctx->fail(RJ_ERR_ARG, "bad function");
    const rj_ctx::Times& T = ctx->times[fn];
    if (T.execute.empty()) return ctx->fail(RJ_ERR_STATE, "function not executed yet");
    auto mn = [](const std::vector<double>& v) { return *std::min_element(v.begin(), v.end()); };
    const int fn_id = (fn == RJ_ERRORBOUNDS) ? 5 : fn;
    char line[512];
    int k = std::snprintf(line, sizeof(line), "%lld\t\t%d\t%d\t%d\t%d", (long long)T.pts, T.np, fn_id, T.passes, T.ncons);
    const std::vector<double>* cols[9] = {&T.execute, &T.ptMem, &T.ptRender, &T.polyMem, &T.polyRender,
                                          &T.setup, &T.tri, &T.polyIndex, &T.backend};
    for (auto* v : cols) k += std::snprintf(line + k, sizeof(line) - k, "\t%.3f", mn(*v));
    if (fn == RJ_RASTER) std::snprintf(line + k, sizeof(line) - k, "\t%g", ctx->eps);
    else std::snprintf(line + k, sizeof(line) - k, "\t0");
    std::snprintf(buf, len, "%s", line);
    return RJ_OK;
}

int rj_polyset_grid(rj_ctx* ctx, rj_polyset* ps, int32_t res, int32_t* cell_off, int32_t* cand, int64_t* ncand,
                    int32_t* deficient, int64_t* ndef)
{
    RJ_GUARD(ctx);
    if (!ps || ps->ctx != ctx || res < 1 || res > 16384) return ctx->fail(RJ_ERR_ARG, "bad collection or resolution");
    Grid* G = nullptr;
    int rc = ctx->grid_for(ps, res, &G);
    if (rc) return rc;
    if (ncand) *ncand = G->ncand;
    if (ndef) *ndef = (int64_t)G->g.deficient.size();
    hipError_t e = hipSuccess;
    if (cell_off)
        e = hipMemcpyAsync(cell_off, G->d_cell_off.p, ((size_t)res * res + 1) * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (cand && G->ncand && e == hipSuccess)
        e = hipMemcpyAsync(cand, G->d_cand.p, (size_t)G->ncand * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "grid export");
    if (deficient) std::copy(G->g.deficient.begin(), G->g.deficient.end(), deficient);
    return RJ_OK;
}

int rj_canvas_info(rj_ctx* ctx, int32_t* W, int32_t* H)
{
    RJ_GUARD(ctx);
    if (!ctx->polys) return ctx->fail(RJ_ERR_STATE, "no polygons set");
    CanvasPlan p;
    int rc = make_plan(ctx->polys->ps.bbox, ctx->eps, ctx->ground_res, &p);
    if (rc) return ctx->fail(rc, "bad canvas");
    if (W) *W = p.W;
    if (H) *H = p.H;
    return RJ_OK;
}

int rj_kernel_times(rj_ctx* ctx, int32_t i, char* name, size_t name_len, double* ms, int32_t* launches)
{
    RJ_GUARD(ctx);
    if (i < 0 || i >= (int)ctx->kstats.size()) return (int)ctx->kstats.size();
    if (name && name_len) std::snprintf(name, name_len, "%s", ctx->kstats[i].name.c_str());
    if (ms) *ms = ctx->kstats[i].ms;
    if (launches) *launches = ctx->kstats[i].launches;
    return (int)ctx->kstats.size();
}

int64_t rj_exact_points(rj_ctx* ctx)
{
    RJ_GUARD(ctx);
    return ctx->last_exact;
}

int rj_sync(rj_ctx* ctx)
{
    RJ_GUARD(ctx);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? RJ_OK : ctx->hip_fail(e, "sync");
}

}  // extern "C"
