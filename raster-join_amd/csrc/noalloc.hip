// noalloc.hip -- rj_execute_query_noalloc: one query with the single-query
// kernels and no device or pinned-host allocation, for callers that cannot
// allocate (HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).
//
// Its kernels take their per-call buffers -- the per-polygon results (count,
// Sum hi, Sum lo), the border-queue counter and the border queue of a hybrid
// pass, and the double sums handed back -- from `scratch`, a file-scope device
// array of kNoAllocScratchBytes. Each GPU holds its own copy until the process
// ends. Its device address is looked up once per device with
// hipGetSymbolAddress at context creation (rj_create), with that device
// current, and kept per device; host code uses only that address.
//
// Every call on a device shares that device's array and leaves its contents
// behind: each call clears what must start at zero (with a fill kernel, not a
// memset), and calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "engine.hpp"

namespace rj {

constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr int kNoAllocMaxPolygons = 65536;
constexpr int kNoAllocMaxDevices = 64;

}  // namespace rj

static __device__ __align__(256) unsigned char scratch[rj::kNoAllocScratchBytes];

namespace rj {
namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};
hipError_t g_scratch_err[kNoAllocMaxDevices] = {};
bool g_scratch_tried[kNoAllocMaxDevices] = {};

// sums[p] = hi[p] + lo[p] * 2^-32, the host's fixed_to_double (a product by a
// power of two is exact, so one rounding as on the host).
__global__ void k_fixed_to_double(const long long* __restrict__ hi, const unsigned long long* __restrict__ lo,
                                  int np, double* __restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) out[p] = (double)hi[p] + (double)lo[p] * (1.0 / 4294967296.0);
}

size_t al256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

// The device address of `scratch` on `device` (which must be current), looked
// up once per device. Returns hipSuccess or the lookup's error.
hipError_t noalloc_scratch_setup(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_tried[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch_tried[device] = true;
        g_scratch_err[device] = e;
        g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
        if (e != hipSuccess) (void)hipGetLastError();
    }
    return g_scratch_err[device];
}

namespace {

void* scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kNoAllocMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

// rj_execute_query's single-query path (rj_ctx::run_single) on structures that
// are already resident, with every per-call buffer in `base`.
int run_noalloc(rj_ctx* ctx, rj_table* t, const QuerySpec& s, char* base, rj_result* out)
{
    const int fn = s.exec;
    rj_polyset* P = s.P;
    const int np = P->ps.np;
    if (t->host) return ctx->fail(RJ_ERR_UNSUPPORTED, "no-alloc execution needs a device-resident table");
    if (!ctx->cache) return ctx->fail(RJ_ERR_UNSUPPORTED, "no-alloc execution needs cached polygon structures");
    if (fn == RJ_ERRORBOUNDS)
        return ctx->fail(RJ_ERR_UNSUPPORTED, "errorbounds needs a canvas larger than the no-alloc scratch");
    if (fn == RJ_RASTER && ctx->raster_path == 1)
        return ctx->fail(RJ_ERR_UNSUPPORTED, "the canvas path needs a canvas larger than the no-alloc scratch");
    if (np > kNoAllocMaxPolygons)
        return ctx->fail(RJ_ERR_LIMIT, "no-alloc execution takes at most 65536 polygons (" + std::to_string(np) +
                                           " given)");
    DevQuery q;
    int rc = ctx->build_query(t, s, &q);
    if (rc) return rc;
    const bool avg = q.agg_type != RJ_AGG_COUNT;

    // the structures the pass reads must be resident already
    Canvas* C = nullptr;
    Grid* G = nullptr;
    const char* missing = "the query's polygon structures are not resident: run it once with rj_execute_query";
    if (fn == RJ_RASTER || fn == RJ_HYBRID) {
        CanvasPlan plan;
        if (make_plan(P->ps.bbox, canvas_eps(s), s.ground_res, &plan))
            return ctx->fail(RJ_ERR_ARG, "canvas too large or bad accuracy for this polygon bbox");
        auto it = P->canvases.find(plan.cell);
        if (it == P->canvases.end()) return ctx->fail(RJ_ERR_STATE, missing);
        C = it->second.get();
    }
    if (fn == RJ_HYBRID || fn == RJ_INDEX) {
        auto it = P->grids.find(s.index_res);
        if (it == P->grids.end()) return ctx->fail(RJ_ERR_STATE, missing);
        G = it->second.get();
    }
    if (fn == RJ_RASTER) {
        if (!C->lab_built || C->lab_tshift != ctx->raster_tshift() || C->lab_state == 0)
            return ctx->fail(RJ_ERR_STATE, missing);
        if (C->lab_state == 2)
            return ctx->fail(RJ_ERR_UNSUPPORTED, "polygons overlap: their raster join needs the canvas path");
    }
    if (fn == RJ_HYBRID && (C->hyb_res != G->g.xs || C->hyb_tshift != ctx->hybrid_tshift() || !C->d_hlab.p))
        return ctx->fail(RJ_ERR_STATE, missing);

    // scratch layout: results (count | hi | lo, then the border-queue count),
    // the double sums, then the border queue in what is left
    const size_t res_words = (size_t)np * 3;
    unsigned long long* res_base = reinterpret_cast<unsigned long long*>(base);
    uint32_t* qcount = reinterpret_cast<uint32_t*>(res_base + res_words);
    double* dsum = reinterpret_cast<double*>(base + al256((res_words + 1) * 8));
    char* qbase = reinterpret_cast<char*>(dsum) + al256((size_t)np * 8);
    const size_t room = (size_t)(base + kNoAllocScratchBytes - qbase) / 16;
    DevResult res{res_base, (long long*)(res_base + np), res_base + 2 * (size_t)np};

    struct Pin {
        rj_polyset* P;
        explicit Pin(rj_polyset* p) : P(p) { P->pins++; }
        ~Pin() { P->pins--; }
    } pin(P);
    int64_t rb0 = s.rb, re0 = s.re;
    ctx->narrow_rows(t, s, &rb0, &re0);
    hipStream_t st = ctx->stream;
    auto check = [&](const char* what) -> int {
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? RJ_OK : ctx->hip_fail(e, what);
    };
    auto clear_results = [&] {
        launch_fill_u32(st, reinterpret_cast<uint32_t*>(res_base), (res_words + 1) * 2, 0);
        return check("no-alloc clear");
    };
    if ((rc = clear_results())) return rc;

    PointArgs pa;
    std::memset(&pa, 0, sizeof(pa));
    pa.xy = static_cast<const float*>(t->xy);
    pa.begin = rb0;
    pa.end = re0;
    pa.cols = ctx->dev_cols(t);
    pa.q = q;
    pa.res = res;
    pa.lds_np = lds_direct_cap(avg);
    pa.cap = ctx->opt_cap;
    pa.wg_rows = t->organized ? kPassRowsPerWorkgroup : 0;
    bool queued = false;
    if (fn == RJ_INDEX) {
        pa.g = ctx->dev_grid(G);
        pa.P = ctx->dev_polys(P, nullptr);
        launch_points(st, PT_INDEX, pa);
    } else if (fn == RJ_HYBRID) {
        pa.cv = dev_canvas(C->plan);
        pa.g = ctx->dev_grid(G);
        pa.P = ctx->dev_polys(P, C);
        pa.label = C->d_hlab.p;
        pa.label16 = C->hyb16;
        pa.tshift = C->hyb_tshift;
        pa.tiles = C->d_htiles.as<uint32_t>();
        pa.tiles_w = pa.cv.LW >> pa.tshift;
        ctx->hybrid_args(C, &pa);
        const int64_t rows = std::max<int64_t>(0, re0 - rb0);
        pa.queue = reinterpret_cast<float4*>(qbase);
        pa.qcount = qcount;
        pa.qcap = (uint32_t)std::min<size_t>(room, rj_ctx::queue_cap(C, rows));
        launch_points(st, PT_HYBRID_QUEUE, pa);
        if ((rc = check("no-alloc points_hybrid"))) return rc;
        launch_exact(st, pa);
        queued = true;
    } else {
        pa.cv = dev_canvas(C->plan);
        pa.P = ctx->dev_polys(P, C);
        pa.label = C->d_lab.p;
        pa.label16 = C->lab16;
        pa.tshift = C->lab_tshift;
        pa.tiles = C->d_tiles.as<uint32_t>();
        pa.tiles_w = pa.cv.LW >> pa.tshift;
        launch_points(st, PT_LABEL, pa);
    }
    if ((rc = check("no-alloc point pass"))) return rc;
    if (queued) {
        // a border queue that ran out of room: the pass again, exact tests inline
        uint32_t n = 0;
        hipError_t e = hipMemcpyAsync(&n, qcount, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return ctx->hip_fail(e, "no-alloc queue count");
        if (n > pa.qcap) {
            PointArgs ia = pa;
            ia.queue = nullptr;
            ia.qcount = nullptr;
            if ((rc = clear_results())) return rc;
            launch_points(st, PT_LABEL_HYBRID, ia);
            if ((rc = check("no-alloc points_hybrid inline"))) return rc;
        }
    }
    if (out->sums) {
        hipLaunchKernelGGL(k_fixed_to_double, dim3((np + 255) / 256), dim3(256), 0, st, res.hi, res.lo, np, dsum);
        if ((rc = check("no-alloc sums"))) return rc;
    }
    hipError_t e = hipSuccess;
    auto copy = [&](void* dst, const void* src) {
        if (dst && e == hipSuccess) e = hipMemcpyAsync(dst, src, (size_t)np * 8, hipMemcpyDeviceToHost, st);
    };
    copy(out->counts, res.cnt);
    copy(out->sum_hi, res.hi);
    copy(out->sum_lo, res.lo);
    copy(out->sums, dsum);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return ctx->hip_fail(e, "no-alloc readback");
    out->path = RJ_PATH_SINGLE;
    return RJ_OK;
}

}  // namespace
}  // namespace rj

extern "C" int rj_execute_query_noalloc(rj_ctx* ctx, rj_table* t, const rj_query* q, rj_result* out)
{
    if (!ctx) return RJ_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (!t || !q || !out) return ctx->fail(RJ_ERR_ARG, "null table, query or result");
    if (t->device != ctx->device) return ctx->fail(RJ_ERR_ARG, "table of another device");
    const hipError_t se = rj::noalloc_scratch_setup(ctx->device);
    if (se != hipSuccess) return ctx->hip_fail(se, "no-alloc scratch address");
    char* base = static_cast<char*>(rj::scratch_ptr(ctx->device));
    if (!base) return ctx->fail(RJ_ERR_DEVICE, "no-alloc scratch not available on this device");
    rj::QuerySpec s;
    int rc = spec_from_query(ctx, t, q, &s);
    if (!rc) rc = rj::run_noalloc(ctx, t, s, base, out);
    out->status = rc;
    return rc;
}
