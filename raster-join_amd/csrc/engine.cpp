// engine.cpp -- device memory, resident polygon structures, the single-query
// join (GLFunction::execute per function type) and fused batches.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>

namespace rj {
size_t organize_scratch_bytes(int64_t n);
int organize_rows(hipStream_t s, int64_t n, float* xy, void** cols, const int32_t* types, int ncols,
                  const float bbox[4], const uint32_t* time, uint32_t t0, uint32_t span, uint32_t nb,
                  int64_t* off_dev, void* scratch);
}  // namespace rj

namespace rj {

// Largest collection whose per-polygon sums a workgroup keeps in LDS indexed
// by polygon (16 KiB); larger ones use the hashed slots.
int lds_direct_cap(bool avg) { return avg ? 16384 / 24 : 16384 / 8; }

// Accuracy that gives a query its canvas. The hybrid join is exact on any
// canvas (border pixels take the exact test), so like the reference's outline
// pass (HybridJoin::updateBuffers, HybridJoin.cpp:60-75) it uses a fixed
// resolution, 8192 pixels on the long side of the collection's bbox, whatever
// the accuracy: coarse accuracies would otherwise send most points to the
// exact test.
double canvas_eps(const QuerySpec& s)
{
    if (s.exec != RJ_HYBRID) return s.eps;
    const double* b = s.P->ps.bbox;
    const double ext = std::max(b[2] - b[0], b[3] - b[1]);
    return ext / 8191.999 * std::sqrt(2.0) * s.ground_res;
}

}  // namespace rj

using namespace rj;

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

size_t al256(size_t b) { return (b + 255) / 256 * 256; }


// label-map tiles: columns LW >> tshift, rows ceil(H / 2^tshift); the tile
// buffer ends with two u64 overlap counters
size_t ntiles(const DevCanvas& c, int tshift)
{
    return (size_t)(c.LW >> tshift) * (size_t)((c.H + (1 << tshift) - 1) >> tshift);
}
size_t tiles_bytes(const DevCanvas& c, int tshift) { return (ntiles(c, tshift) * 4 + 7) / 8 * 8 + 16; }
unsigned long long* tiles_aux(const DevBuf& tiles, const DevCanvas& c, int tshift)
{
    return reinterpret_cast<unsigned long long*>(tiles.as<char>() + (ntiles(c, tshift) * 4 + 7) / 8 * 8);
}

}  // namespace

// ---------------------------------------------------------------------------
// buffers, tables, collections
// ---------------------------------------------------------------------------

size_t rj::size_class(size_t b)
{
    if (b <= 256) return 256;
    if (b <= ((size_t)1 << 20)) {
        size_t c = 512;
        while (c < b) c <<= 1;
        return c;
    }
    const int lg = 63 - __builtin_clzll((unsigned long long)(b - 1));  // 2^lg < b <= 2^(lg+1)
    const size_t step = (size_t)1 << (lg - 3);
    return (b + step - 1) / step * step;
}

namespace {
constexpr size_t kPoolBlockMax = (size_t)1 << 30;  // larger blocks are freed, not pooled
constexpr int64_t kPoolUnlimited = (int64_t)4 << 30;  // pool bound of a context without a budget
}  // namespace

void* rj::Budget::take(size_t cls)
{
    auto it = pool.find(cls);
    if (it == pool.end()) return nullptr;
    void* p = it->second;
    pool.erase(it);
    cached -= (int64_t)cls;
    return p;
}

bool rj::Budget::put(void* p, size_t cls)
{
    if (cls > kPoolBlockMax) return false;
    const int64_t room = limit > 0 ? limit - used : kPoolUnlimited;
    if (cached + (int64_t)cls > std::min<int64_t>(room, kPoolUnlimited)) return false;
    pool.emplace(cls, p);
    cached += (int64_t)cls;
    return true;
}

void rj::Budget::trim(int64_t keep)
{
    while (cached > keep && !pool.empty()) {
        auto it = std::prev(pool.end());  // largest first
        (void)hipFree(it->second);
        cached -= (int64_t)it->first;
        pool.erase(it);
    }
}

void rj::DevBuf::release()
{
    if (p) {
        if (budget) {
            budget->used -= (int64_t)bytes;
            if (!budget->put(p, bytes)) (void)hipFree(p);
        } else {
            (void)hipFree(p);
        }
    }
    p = nullptr;
    bytes = 0;
    budget.reset();
}

size_t rj::Canvas::device_bytes() const
{
    return d_pc.bytes + d_XY.bytes + d_lab.bytes + d_tiles.bytes + d_hlab.bytes + d_htiles.bytes + d_mask.bytes +
           d_rects.bytes + d_blab.bytes + d_btiles.bytes + d_bmask.bytes + d_bb.bytes;
}

size_t rj_polyset::device_bytes() const
{
    size_t b = d_xy.bytes + d_off.bytes + d_prev.bytes;
    for (auto& c : canvases) b += c.second->device_bytes();
    for (auto& g : grids) b += g.second->device_bytes();
    return b;
}

static size_t type_bytes(int t) { return t == RJ_ATTR_UINT8 ? 1 : 4; }

rj_table::~rj_table()
{
    if (hdata && hpinned) (void)hipHostFree(hdata);
    if (hdata && !hpinned) std::free(hdata);
    if (xy) (void)hipFree(xy);
    for (int c = 0; c < ncols; c++)
        if (cols[c]) (void)hipFree(cols[c]);
}

size_t rj_table::row_bytes() const
{
    size_t b = 8;
    for (int c = 0; c < ncols; c++) b += type_bytes(types[c]);
    return b;
}

size_t rj_table::bytes() const
{
    size_t b = (size_t)n * 8;
    for (int c = 0; c < ncols; c++) b += (size_t)n * type_bytes(types[c]);
    return b;
}

void rj_table::time_rows(uint32_t lo, uint32_t hi, int64_t* b, int64_t* e) const
{
    *b = 0;
    *e = n;
    if (time_col < 0 || bucket_off.size() < 2) return;
    if (lo > hi) {
        *e = 0;
        return;
    }
    const int64_t nb = (int64_t)bucket_off.size() - 1;
    auto bucket = [&](uint32_t t) -> int64_t {
        if (t < t0) return 0;
        return std::min<int64_t>((t - t0) / span, nb - 1);
    };
    *b = bucket_off[bucket(lo)];
    *e = bucket_off[bucket(hi) + 1];
}

// ---------------------------------------------------------------------------
// context basics
// ---------------------------------------------------------------------------

rj_ctx::~rj_ctx()
{
    if (stream) (void)hipStreamSynchronize(stream);
    if (own_table) delete table;
    if (own_polys) polysets.erase(own_polys.get());
    own_polys.reset();
    for (auto* P : polysets) P->ctx = nullptr;  // caller-owned collections may outlive the context
    for (auto& S : slots) {
        if (S.graph) (void)hipGraphExecDestroy(S.graph);
        if (S.h_block) (void)hipHostFree(S.h_block);
        if (S.h_res) (void)hipHostFree(S.h_res);
        if (S.done) (void)hipEventDestroy(S.done);
        if (S.blk) (void)hipEventDestroy(S.blk);
    }
    if (h_res) (void)hipHostFree(h_res);
    if (h_stage) (void)hipHostFree(h_stage);
    if (stage_ev) (void)hipEventDestroy(stage_ev);
    for (int k = 0; k < 2; k++) {
        if (h_chunk[k]) (void)hipHostFree(h_chunk[k]);
        if (ev_h2d[k]) (void)hipEventDestroy(ev_h2d[k]);
        if (ev_pass[k]) (void)hipEventDestroy(ev_pass[k]);
    }
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    for (auto e : pool) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

int rj_ctx::fail(int code, const std::string& m)
{
    err = m;
    return code;
}

int rj_ctx::hip_fail(hipError_t e, const char* what)
{
    err = std::string(what) + ": " + hipGetErrorString(e);
    return (e == hipErrorOutOfMemory) ? RJ_ERR_NOMEM : RJ_ERR_DEVICE;
}

// Device allocation counted against the budget; reuses a large-enough buffer.
// When the budget is short, cached polygon structures not in use are evicted
// (least recently used first).
int rj_ctx::alloc(DevBuf& b, size_t bytes, const char* what)
{
    if (bytes == 0) bytes = 8;
    if (b.p && b.bytes >= bytes) return RJ_OK;
    b.release();
    bytes = size_class(bytes);
    void* p = budget->take(bytes);
    if (!p) {
        Budget& B = *budget;
        auto fits = [&] { return B.limit <= 0 || B.used + (int64_t)bytes <= B.limit; };
        if (!fits()) make_room((int64_t)bytes);  // evicted structures go to the pool
        p = B.take(bytes);
        if (!p) {
            if (!fits())
                return fail(RJ_ERR_NOMEM, std::string("memory budget exhausted allocating ") + what + " (" +
                                              std::to_string(bytes) + " bytes, " + std::to_string(B.used) + " of " +
                                              std::to_string(B.limit) + " in use)");
            if (B.limit > 0) B.trim(B.limit - B.used - (int64_t)bytes);
            hipError_t e = hipMalloc(&p, bytes);
            if (e != hipSuccess && B.cached > 0) {
                (void)hipGetLastError();
                B.trim(0);
                e = hipMalloc(&p, bytes);
            }
            if (e != hipSuccess) {
                (void)hipGetLastError();
                return hip_fail(e, what);
            }
        }
    }
    b.p = p;
    b.bytes = bytes;
    b.budget = budget;
    budget->used += (int64_t)bytes;
    return RJ_OK;
}

int rj_ctx::upload(DevBuf& b, const void* src, size_t bytes, const char* what)
{
    int rc = alloc(b, bytes, what);
    if (rc) return rc;
    if (bytes == 0) return RJ_OK;
    hipError_t e = hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, what);
    return RJ_OK;
}

// The device address of pinned host memory from pinned() (nullptr if the
// runtime gives none: callers then use a copy-engine transfer).
static void* host_dev(void* h)
{
    void* d = nullptr;
    return hipHostGetDevicePointer(&d, h, 0) == hipSuccess ? d : nullptr;
}

int rj_ctx::pinned(void** p, size_t* cur, size_t bytes, const char* what)
{
    if (*p && *cur >= bytes) return RJ_OK;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *cur = 0;
    // mapped and coherent: kernels read and write it directly (k_copy_words)
    hipError_t e = hipHostMalloc(p, std::max<size_t>(bytes, 64), hipHostMallocMapped | hipHostMallocCoherent);
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_fail(e, what);
    }
    *cur = std::max<size_t>(bytes, 64);
    return RJ_OK;
}

hipEvent_t rj_ctx::event()
{
    if (pool_used == pool.size()) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        pool.push_back(e);
    }
    return pool[pool_used++];
}

void rj_ctx::collect_times(double* pt_ms, double* poly_ms)
{
    if (!timing_pass_only) kstats.clear();
    *pt_ms = *poly_ms = 0;
    for (auto& l : launches) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, l.a, l.b);
        bool found = false;
        for (auto& k : kstats)
            if (k.name == l.name) {
                k.ms += ms;
                k.launches++;
                found = true;
            }
        if (!found) kstats.push_back({l.name, ms, 1});
        if (pass_kernel(l.name))
            *pt_ms += ms;
        else
            *poly_ms += ms;
    }
    launches.clear();
    pool_used = 0;
}

// Evict cached canvases and grids of collections that no batch in flight
// uses, least recently used first, until `bytes` more fit the budget.
int rj_ctx::make_room(int64_t bytes)
{
    if (budget->limit <= 0) return RJ_OK;
    while (budget->used + bytes > budget->limit) {
        rj_polyset* bestP = nullptr;
        double bestKey = 0;
        int bestRes = 0;
        bool bestCanvas = false;
        uint64_t bestUse = UINT64_MAX;
        for (rj_polyset* P : polysets) {
            if (P->pins > 0) continue;
            for (auto& c : P->canvases)
                if (c.second->last_use < bestUse && c.second->device_bytes() > 0) {
                    bestUse = c.second->last_use;
                    bestP = P;
                    bestKey = c.first;
                    bestCanvas = true;
                }
            for (auto& g : P->grids)
                if (g.second->last_use < bestUse && g.second->device_bytes() > 0) {
                    bestUse = g.second->last_use;
                    bestP = P;
                    bestRes = g.first;
                    bestCanvas = false;
                }
        }
        if (!bestP) return RJ_ERR_NOMEM;
        (void)hipStreamSynchronize(stream);  // nothing queued may still read it
        if (bestCanvas) bestP->canvases.erase(bestKey);
        else bestP->grids.erase(bestRes);
    }
    return RJ_OK;
}

void rj_ctx::forget_polyset(rj_polyset* P)
{
    polysets.erase(P);
    if (polys == P) polys = nullptr;
}

// ---------------------------------------------------------------------------
// polygon collections
// ---------------------------------------------------------------------------

int rj_ctx::make_polyset(const float* xy, const int32_t* off, int32_t np, const double* bbox,
                         std::unique_ptr<rj_polyset>* out)
{
    if (np <= 0 || !xy || !off) return fail(RJ_ERR_ARG, "empty polygon collection");
    if (off[0] != 0) return fail(RJ_ERR_ARG, "offsets must start at 0");
    for (int p = 0; p < np; p++)
        if (off[p + 1] - off[p] < 3) return fail(RJ_ERR_ARG, "polygon " + std::to_string(p) + " has < 3 vertices");
    auto P = std::make_unique<rj_polyset>();
    PolySet& ps = P->ps;
    ps.np = np;
    ps.off.assign(off, off + np + 1);
    const int64_t nv = off[np];
    if (!ps.derive(xy)) return fail(RJ_ERR_ARG, "non-finite polygon vertex");
    if (bbox) {
        for (int d = 0; d < 4; d++) ps.bbox[d] = bbox[d];
    } else {
        float mn[2] = {ps.pbox[0], ps.pbox[1]}, mx[2] = {ps.pbox[2], ps.pbox[3]};
        for (int p = 1; p < np; p++) {
            mn[0] = std::min(mn[0], ps.pbox[4 * p]);
            mn[1] = std::min(mn[1], ps.pbox[4 * p + 1]);
            mx[0] = std::max(mx[0], ps.pbox[4 * p + 2]);
            mx[1] = std::max(mx[1], ps.pbox[4 * p + 3]);
        }
        ps.bbox[0] = mn[0]; ps.bbox[1] = mn[1]; ps.bbox[2] = mx[0]; ps.bbox[3] = mx[1];
    }
    for (int d = 0; d < 4; d++) ps.bboxf[d] = (float)ps.bbox[d];
    if (!(ps.bbox[2] > ps.bbox[0]) || !(ps.bbox[3] > ps.bbox[1])) return fail(RJ_ERR_ARG, "degenerate bbox");
    // the caller's vertices straight to the device (the copy is done when
    // this returns); the edge list is derived there
    int rc = upload(P->d_xy, xy, (size_t)nv * 8, "polygons");
    if (!rc) rc = upload(P->d_off, ps.off.data(), ps.off.size() * 4, "polygon offsets");
    if (!rc) rc = alloc(P->d_prev, (size_t)nv * 4, "edge list");
    if (!rc)
        rc = launch("prev", [&] { launch_prev(stream, P->d_off.as<int32_t>(), np, nv, P->d_prev.as<int32_t>()); });
    if (!rc) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) rc = hip_fail(e, "polygon upload");
    }
    if (rc) return rc;
    P->ctx = this;
    polysets.insert(P.get());
    *out = std::move(P);
    return RJ_OK;
}

// Plan, snapped polygons and scan-conversion bands of a collection on the
// canvas of one accuracy (GLFunction::setMaxCellSize, RasterJoin::updateBuffers).
int rj_ctx::canvas_for(rj_polyset* P, double eps_m, double gres, Canvas** out)
{
    CanvasPlan plan;
    int rc = make_plan(P->ps.bbox, eps_m, gres, &plan);
    if (rc) return fail(rc, "canvas too large or bad accuracy for this polygon bbox");
    auto it = P->canvases.find(plan.cell);
    Canvas* C;
    if (it != P->canvases.end()) {
        C = it->second.get();
        C->last_use = ++clock;
        if (cache) {
            *out = C;
            return RJ_OK;
        }
    } else {
        auto nc = std::make_unique<Canvas>();
        C = nc.get();
        P->canvases[plan.cell] = std::move(nc);
    }
    C->last_use = ++clock;
    C->plan = plan;
    C->lab_built = false;
    C->lab_state = 0;
    C->hyb_res = -1;
    double t0 = now_ms();
    rasterize_polys(P->ps, plan, &C->pr);
    cur_setup += now_ms() - t0;
    t0 = now_ms();
    // regions and bands: one pinned staging buffer, one copy (the
    // error-bounds bands follow on the first error-bounds join, bounds_for)
    C->bb_built = false;
    C->bands_bounds = nullptr;
    const PolyRaster& pr = C->pr;
    const size_t parts[2] = {pr.region.size() * 4, pr.bands_sum.size() * sizeof(Band)};
    const void* srcs[2] = {pr.region.data(), pr.bands_sum.data()};
    size_t offs[2], total = 0;
    for (int k = 0; k < 2; k++) {
        offs[k] = total;
        total += al256(parts[k]);
    }
    if (stage_busy) {
        (void)hipEventSynchronize(stage_ev);
        stage_busy = false;
    }
    if ((rc = pinned(&h_stage, &h_stage_bytes, total, "pinned staging"))) return rc;
    for (int k = 0; k < 2; k++)
        if (parts[k]) std::memcpy((char*)h_stage + offs[k], srcs[k], parts[k]);
    if ((rc = alloc(C->d_pc, total, "polygon regions"))) {
        P->canvases.erase(plan.cell);
        return rc;
    }
    // parts are 256-byte aligned, so total is whole 64-bit words
    hipError_t e = hipSuccess;
    if (void* ds = host_dev(h_stage)) {
        if ((rc = launch("upload", [&] {
                 launch_copy_words(stream, CopyWords{(const unsigned long long*)ds, C->d_pc.as<unsigned long long>(),
                                                     total / 8});
             })))
            return rc;
    } else {
        e = hipMemcpyAsync(C->d_pc.p, h_stage, total, hipMemcpyHostToDevice, stream);
    }
    if (e != hipSuccess) return hip_fail(e, "polygon regions");
    if (!stage_ev) (void)hipEventCreateWithFlags(&stage_ev, hipEventDisableTiming);
    (void)hipEventRecord(stage_ev, stream);
    stage_busy = true;
    char* base = C->d_pc.as<char>();
    C->region = (int32_t*)(base + offs[0]);
    C->bands_sum = (Band*)(base + offs[1]);
    const int64_t nv = P->ps.off[P->ps.np];
    if ((rc = alloc(C->d_XY, (size_t)nv * 8, "snapped vertices"))) {
        P->canvases.erase(plan.cell);
        return rc;
    }
    C->X = C->d_XY.as<int32_t>();
    C->Y = C->X + nv;
    const float* pxy = P->d_xy.as<float>();
    int32_t *X = C->X, *Y = C->Y;
    if ((rc = launch("snap", [&] { launch_snap(stream, pxy, nv, plan.minx, plan.miny, plan.fxs, plan.fys, X, Y); })))
        return rc;
    cur_polymem += now_ms() - t0;
    *out = C;
    return RJ_OK;
}

// CPUJoin grid index of a collection (GridIndexF::buildGrid) at one resolution.
int rj_ctx::grid_for(rj_polyset* P, int res, Grid** out)
{
    auto it = P->grids.find(res);
    if (it != P->grids.end() && cache) {
        it->second->last_use = ++clock;
        *out = it->second.get();
        return RJ_OK;
    }
    Grid* G;
    if (it != P->grids.end()) {
        G = it->second.get();
    } else {
        auto ng = std::make_unique<Grid>();
        G = ng.get();
        P->grids[res] = std::move(ng);
    }
    G->last_use = ++clock;
    double t0 = now_ms();
    int rc = grid_device(P, res, G);
    cur_polyindex += now_ms() - t0;
    if (rc) {
        P->grids.erase(res);
        return rc;
    }
    for (auto& c : P->canvases)
        if (c.second->hyb_res == res) c.second->hyb_res = -1;
    *out = G;
    return RJ_OK;
}

// Label map with plain stores and its tile map; with `count` the two overlap
// counters at the end of the tile buffer receive the coverage summed over
// polygons and the number of labeled pixels (equal iff no pixel is in two
// polygons). Returns without waiting.
int rj_ctx::build_labels(rj_polyset* P, Canvas* C, void* lab, int label16, DevBuf& tiles, int tshift, bool count)
{
    const DevCanvas cv = dev_canvas(C->plan);
    const size_t npx = (size_t)cv.LW * cv.H;
    const size_t words = label16 ? npx / 2 : npx;
    int rc = alloc(tiles, tiles_bytes(cv, tshift), "label tiles");
    if (rc) return rc;
    unsigned long long* aux = tiles_aux(tiles, cv, tshift);
    const DevPolys dp = dev_polys(P, C);
    if ((rc = launch("fill_label", [&] {
             launch_fill_u32(stream, static_cast<uint32_t*>(lab), words, 0xFFFFFFFFu);
             launch_fill_u32(stream, (uint32_t*)aux, 4, 0);
         })))
        return rc;
    if ((rc = launch("band_label", [&] {
             launch_band(stream, BAND_LABEL, C->bands_sum, (int)C->pr.bands_sum.size(), dp, cv, nullptr, nullptr,
                         nullptr, DevResult{}, nullptr, lab, label16, count ? aux : nullptr, C->pr.lds_sum, 0);
         })))
        return rc;
    uint32_t* t = tiles.as<uint32_t>();
    return launch("tile_labels",
                  [&] { launch_tile_labels(stream, lab, label16, cv, tshift, t, count ? aux + 1 : nullptr); });
}

int rj_ctx::raster_labels(rj_polyset* P, Canvas* C, bool* fresh)
{
    *fresh = false;
    const int tshift = raster_tshift();
    if (C->lab_built && cache && C->lab_tshift == tshift) return RJ_OK;
    const DevCanvas cv = dev_canvas(C->plan);
    const size_t npx = (size_t)cv.LW * cv.H;
    const int l16 = P->ps.np < (int)0xFFF0 ? 1 : 0;
    int rc = alloc(C->d_lab, npx * (l16 ? 2 : 4) + 16, "label map");
    if (rc) return rc;
    if ((rc = build_labels(P, C, C->d_lab.p, l16, C->d_tiles, tshift, true))) return rc;
    C->lab_built = true;
    C->lab16 = l16;
    C->lab_tshift = tshift;
    C->lab_state = 0;  // confirmed when the counters are read
    *fresh = true;
    return RJ_OK;
}

// The error-bounds bands of a canvas, planned and uploaded on its first
// error-bounds join.
int rj_ctx::bounds_for(rj_polyset* P, Canvas* C)
{
    if (C->bb_built) return RJ_OK;
    bounds_bands(P->ps, &C->pr);
    int rc = upload(C->d_bb, C->pr.bands_bounds.data(), C->pr.bands_bounds.size() * sizeof(Band), "error-bounds bands");
    if (rc) return rc;
    C->bands_bounds = C->d_bb.as<Band>();
    C->bb_built = true;
    return RJ_OK;
}

// The error-bounds label map: the raster label map (with its overlap
// counters) whose boundary pixels, every pixel within one of a pixel a
// snapped edge touches (the boundary rule of the error bounds), carry the
// halo flag (kLabelNoneHalo outside every polygon), then its tile map. The
// point pass counts the points of flagged pixels per pixel for the bounds
// kernel and sums the rest per polygon like a raster pass.
int rj_ctx::bounds_labels(rj_polyset* P, Canvas* C)
{
    const int tshift = raster_tshift();
    if (C->blab_built && cache && C->blab_tshift == tshift) return RJ_OK;
    if (P->ps.np >= (int)kHaloFlag32) return fail(RJ_ERR_LIMIT, "too many polygons for an error-bounds label map");
    const DevCanvas cv = dev_canvas(C->plan);
    const size_t npx = (size_t)cv.LW * cv.H;
    const size_t wpr = ((size_t)cv.W + 31) / 32;
    const int l16 = P->ps.np < (int)(kHaloFlag16 - 16) ? 1 : 0;  // 16-bit ids keep the flag bit free
    int rc = alloc(C->d_blab, npx * (l16 ? 2 : 4) + 16, "error-bounds label map");
    if (!rc) rc = alloc(C->d_bmask, wpr * cv.H * 4, "boundary mask");
    if (rc) return rc;
    if ((rc = build_labels(P, C, C->d_blab.p, l16, C->d_btiles, tshift, true))) return rc;
    uint32_t* mask = C->d_bmask.as<uint32_t>();
    void* lab = C->d_blab.p;
    const int32_t* prev = P->d_prev.as<int32_t>();
    const int64_t nv = P->ps.off[P->ps.np];
    if ((rc = launch("fill_mask", [&] { launch_fill_u32(stream, mask, wpr * cv.H, 0); }))) return rc;
    if ((rc = launch("touch", [&] { launch_touch(stream, C->X, C->Y, prev, nv, cv, mask); }))) return rc;
    if ((rc = launch("halo_flags", [&] { launch_halo_flags(stream, mask, cv, lab, l16); }))) return rc;
    uint32_t* t = C->d_btiles.as<uint32_t>();
    if ((rc = launch("tile_labels", [&] { launch_tile_labels(stream, lab, l16, cv, tshift, t, nullptr); })))
        return rc;
    C->blab_built = true;
    C->blab16 = l16;
    C->blab_tshift = tshift;
    C->blab_state = 0;  // confirmed when the counters are read
    return RJ_OK;
}

int rj_ctx::hybrid_radius(const CanvasPlan& plan) const
{
    if (hybrid_dilation > 0) return hybrid_dilation;
    double m = std::max(std::max(std::fabs(plan.minx), std::fabs(plan.maxx)),
                        std::max(std::fabs(plan.miny), std::fabs(plan.maxy)));
    double err = 16.0 * m * std::ldexp(1.0, -24);
    double pw = std::min((plan.maxx - plan.minx) / plan.W, (plan.maxy - plan.miny) / plan.H);
    return std::max(1, (int)std::ceil(2.0 * err / pw));
}

// Width in pixels of the zone around a touched pixel where a point may be
// misclassified by its pixel center: CPUJoin's float PIP error
// (16 |coord| 2^-24), the snapping of vertices to 1/256 pixel and the pixel
// formula's float error (each below 0.003 pixel), doubled.
double rj_ctx::hybrid_delta(const CanvasPlan& plan) const
{
    const double m = std::max(std::max(std::fabs(plan.minx), std::fabs(plan.maxx)),
                              std::max(std::fabs(plan.miny), std::fabs(plan.maxy)));
    const double err = 16.0 * m * std::ldexp(1.0, -24);
    const double pw = std::min((plan.maxx - plan.minx) / plan.W, (plan.maxy - plan.miny) / plan.H);
    return 2.0 * (err / pw + 0.006);
}

// Border-queue capacity for a pass over `rows` rows: the canvas's last
// measured border share with half again as margin, at least 65536 entries.
uint32_t rj_ctx::queue_cap(const Canvas* C, int64_t rows)
{
    const double share = std::min(1.0, C->hyb_border * 1.5);
    return (uint32_t)std::min<int64_t>(1LL << 27, std::max<int64_t>(65536, (int64_t)(share * (double)rows) + 4096));
}

void rj_ctx::hybrid_args(const Canvas* C, PointArgs* pa) const
{
    pa->halo = C->hyb_halo ? C->d_mask.as<uint32_t>() : nullptr;
    pa->halo_wpr = (C->plan.W + 31) / 32;
    pa->halo_d = C->hyb_delta;
}

// Hybrid label map: polygon labels, with the pixels a snapped edge touches
// (and every pixel of a grid cell with an incomplete candidate list) marked
// for the exact test (HybridJoin::drawOutline). Points of untouched pixels
// near a touched one are sent to the exact test by the point pass (halo_near)
// when the float margin is a small part of a pixel; otherwise the border is
// dilated by whole pixels, as the 3x3 dilation of hybrid.glsl:174-185.
int rj_ctx::hybrid_labels(rj_polyset* P, Canvas* C, Grid* G)
{
    const int tshift = hybrid_tshift();
    if (C->hyb_res == G->g.xs && cache && C->hyb_tshift == tshift) return RJ_OK;
    const double delta = hybrid_delta(C->plan);
    const bool halo = hybrid_dilation == 0 && delta <= 0.25;
    const int r = halo ? 0 : hybrid_radius(C->plan);
    if (r > 31) return fail(RJ_ERR_UNSUPPORTED, "accuracy too fine for the float margin of the exact test");
    const DevCanvas cv = dev_canvas(C->plan);
    const size_t npx = (size_t)cv.LW * cv.H;
    const size_t wpr = ((size_t)cv.W + 31) / 32;
    int rc = alloc(C->d_mask, wpr * cv.H * 4, "border mask");
    if (rc) return rc;
    std::vector<int32_t> rects;
    deficient_rects(G->g, C->plan, &rects);
    if (!rects.empty() && (rc = upload(C->d_rects, rects.data(), rects.size() * 4, "deficient cells"))) return rc;
    const DevPolys dp = dev_polys(P, C);
    uint32_t* mask = C->d_mask.as<uint32_t>();
    // 16-bit maps keep one bit for the halo flag
    int l16 = P->ps.np < (int)(kHaloFlag16 - 16) ? 1 : 0;
    if (P->ps.np >= (int)kHaloFlag32) return fail(RJ_ERR_LIMIT, "too many polygons for a hybrid label map");
    if ((rc = alloc(C->d_hlab, npx * (l16 ? 2 : 4) + 16, "hybrid label map"))) return rc;
    if ((rc = build_labels(P, C, C->d_hlab.p, l16, C->d_htiles, tshift, true))) return rc;
    unsigned long long a[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(a, tiles_aux(C->d_htiles, cv, tshift), 16, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "overlap test");
    if (a[0] != a[1]) {
        // pixels in two or more polygons go to the exact test (32-bit map, CAS)
        l16 = 0;
        if ((rc = alloc(C->d_hlab, npx * 4 + 16, "hybrid label map"))) return rc;
        uint32_t* lab32 = C->d_hlab.as<uint32_t>();
        if ((rc = launch("fill_label", [&] { launch_fill_u32(stream, lab32, npx, kLabelNone); }))) return rc;
        if ((rc = launch("band_label_atomic", [&] {
                 launch_band(stream, BAND_LABEL_ATOMIC, C->bands_sum, (int)C->pr.bands_sum.size(), dp, cv, nullptr,
                             nullptr, nullptr, DevResult{}, nullptr, lab32, 0, nullptr, C->pr.lds_sum, 0);
             })))
            return rc;
    }
    void* lab = C->d_hlab.p;
    const int32_t* prev = P->d_prev.as<int32_t>();
    const int64_t nv = P->ps.off[P->ps.np];
    if ((rc = launch("fill_mask", [&] { launch_fill_u32(stream, mask, wpr * cv.H, 0); }))) return rc;
    if ((rc = launch("touch", [&] { launch_touch(stream, C->X, C->Y, prev, nv, cv, mask); }))) return rc;
    if (!rects.empty()) {
        const int32_t* dr = C->d_rects.as<int32_t>();
        if ((rc = launch("mark_cells", [&] { launch_mark_rects(stream, dr, (int)(rects.size() / 4), cv, mask); })))
            return rc;
    }
    if ((rc = launch("dilate", [&] { launch_dilate_to_label(stream, mask, cv, r, lab, l16); }))) return rc;
    if (halo && (rc = launch("halo_flags", [&] { launch_halo_flags(stream, mask, cv, lab, l16); }))) return rc;
    uint32_t* t = C->d_htiles.as<uint32_t>();
    if ((rc = launch("tile_labels", [&] { launch_tile_labels(stream, lab, l16, cv, tshift, t, nullptr); })))
        return rc;
    C->hyb16 = l16;
    C->hyb_res = G->g.xs;
    C->hyb_tshift = tshift;
    C->hyb_halo = halo;
    C->hyb_delta = (float)delta;
    return RJ_OK;
}

// Device bytes the polygon structures of a query would add (nothing when
// they are resident): the memory model of admission.
int64_t rj_ctx::estimate_bytes(const QuerySpec& s) const
{
    const int np = s.P->ps.np;
    int64_t b = (int64_t)np * 7 * 8 + 64;
    if (s.exec == RJ_INDEX || s.exec == RJ_HYBRID) {
        if (!s.P->grids.count(s.index_res)) {
            const int64_t cells = (int64_t)s.index_res * s.index_res;
            b += cells * 4 + (int64_t)np * 64 * 4;
        }
    }
    if (s.exec != RJ_INDEX) {
        CanvasPlan plan;
        if (make_plan(s.P->ps.bbox, canvas_eps(s), s.ground_res, &plan)) return b;
        auto it = s.P->canvases.find(plan.cell);
        const Canvas* C = it == s.P->canvases.end() ? nullptr : it->second.get();
        const int64_t LW = (plan.W + kLabelRowAlign - 1) & ~(kLabelRowAlign - 1);
        const int64_t npx = LW * plan.H;
        const int64_t lw = np < 0xFFF0 ? 2 : 4;
        if (!C) b += (int64_t)s.P->ps.off[np] * 8 + (int64_t)np * 1024;
        if (s.exec == RJ_RASTER && (!C || !C->lab_built)) b += npx * lw + npx / 64;
        if (s.exec == RJ_HYBRID && (!C || C->hyb_res != s.index_res)) b += npx * lw + npx / 64 + npx / 8;
        if (s.exec == RJ_ERRORBOUNDS) {
            b += (int64_t)plan.W * plan.H * 4;  // count canvas
            if (!C || !C->blab_built) b += npx * lw + npx / 64 + npx / 8;
        }
    }
    return b;
}

// ---------------------------------------------------------------------------
// queries
// ---------------------------------------------------------------------------

DevCols rj_ctx::dev_cols(const rj_table* t) const
{
    DevCols c;
    std::memset(&c, 0, sizeof(c));
    for (int k = 0; k < t->ncols; k++) {
        c.col[k] = t->cols[k];
        c.type[k] = t->types[k];
    }
    for (int k = 0; k < 4; k++) c.box[k] = t->bbox[k];
    c.box_ok = (t->n > 0 && !t->nonfinite) ? 1 : 0;
    c.in_range = t->float_in_range;
    return c;
}

DevPolys rj_ctx::dev_polys(const rj_polyset* P, const Canvas* C) const
{
    DevPolys d;
    std::memset(&d, 0, sizeof(d));
    d.np = P->ps.np;
    d.xy = P->d_xy.as<float>();
    d.off = P->d_off.as<int32_t>();
    if (C) {
        d.X = C->X;
        d.Y = C->Y;
        d.region = C->region;
    }
    return d;
}

DevGrid rj_ctx::dev_grid(const Grid* G) const
{
    DevGrid g;
    g.xs = G->g.xs; g.ys = G->g.ys;
    g.minx = G->g.minx; g.miny = G->g.miny; g.width = G->g.width; g.height = G->g.height;
    g.cell_off = G->d_cell_off.as<int32_t>();
    g.cand = G->d_cand.as<int32_t>();
    g.band_h = G->g.band_h;
    g.band_rows = G->g.band_rows;
    const int32_t* b = G->d_bands.as<int32_t>();
    const size_t np = (size_t)G->np;
    g.band_lo = b;
    g.band_hi = b + np;
    g.pband = b + 2 * np;
    g.band_off = b + 3 * np + 1;
    g.edges = G->d_edges.as<float4>();
    return g;
}

// Key range (col_key) of the values of a column type that satisfy `value op
// c`, the comparison of index.glsl:184-208 / points.vert:138-149 with typed
// operands (Float as IEEE floats: NaN satisfies nothing, -0 == +0). Returns
// false when no value does.
static bool constraint_range(int type, int op, uint32_t c, uint32_t* lo, uint32_t* hi)
{
    uint32_t kmin = 0, kmax = UINT32_MAX, kl = col_key(type, c), ku = kl;
    if (type == RJ_ATTR_FLOAT) {
        kmin = col_key(type, 0xFF800000u);  // -inf
        kmax = col_key(type, 0x7F800000u);  // +inf
        if ((c & 0x7FFFFFFFu) > 0x7F800000u) return false;  // NaN constant
        if ((c & 0x7FFFFFFFu) == 0) {  // either zero: both zeros
            kl = col_key(type, 0x80000000u);
            ku = col_key(type, 0u);
        }
    }
    switch (op) {
    case RJ_EQ: *lo = kl; *hi = ku; return true;
    case RJ_LT: *lo = kmin; *hi = kl - 1; return kl > kmin;
    case RJ_LTE: *lo = kmin; *hi = ku; return true;
    case RJ_GT: *lo = ku + 1; *hi = kmax; return ku < kmax;
    default: *lo = kl; *hi = kmax; return true;  // RJ_GTE
    }
}

// Filters and aggregate of a query on a table's columns (GLFunction::setupPoints,
// GLFunction.cpp:228-323: the columns a query reads, at most 5; at most 10
// constraints).
int rj_ctx::build_query(const rj_table* t, const QuerySpec& s, DevQuery* q)
{
    std::memset(q, 0, sizeof(*q));
    q->agg_slot = -1;
    if ((int)s.cons.size() > RJ_MAX_CONSTRAINTS) return fail(RJ_ERR_LIMIT, "at most 10 constraints");
    auto slot_of = [&](int col) {
        for (int k = 0; k < q->nslots; k++)
            if (q->slot_col[k] == col) return k;
        if (q->nslots == kMaxSlots) return -1;
        q->slot_col[q->nslots] = col;
        return q->nslots++;
    };
    for (int k = 0; k < kMaxSlots; k++) {
        q->slot_lo[k] = 0;
        q->slot_hi[k] = UINT32_MAX;
    }
    for (size_t k = 0; k < s.cons.size(); k++) {
        const rj_constraint& c = s.cons[k];
        if (c.attrib_id < 0 || c.attrib_id >= t->ncols)
            return fail(RJ_ERR_ARG, "constraint column " + std::to_string(c.attrib_id) + " out of range");
        if (c.type < RJ_EQ || c.type > RJ_GTE) return fail(RJ_ERR_ARG, "bad constraint type");
        const int slot = slot_of(c.attrib_id);
        if (slot < 0) return fail(RJ_ERR_LIMIT, "a query may use at most 5 attributes");
        uint32_t lo, hi;
        if (constraint_range(t->types[c.attrib_id], c.type, c.val.uval, &lo, &hi)) {
            q->slot_lo[slot] = std::max(q->slot_lo[slot], lo);
            q->slot_hi[slot] = std::min(q->slot_hi[slot], hi);
        } else {
            q->slot_lo[slot] = 1;  // no value passes
            q->slot_hi[slot] = 0;
        }
    }
    q->plain = 1;
    for (int k = 0; k < q->nslots; k++)
        if (q->slot_lo[k] != 0 || q->slot_hi[k] != UINT32_MAX) q->plain = 0;
    q->agg_type = s.agg_type;
    if (s.agg_type != RJ_AGG_COUNT && s.fn != RJ_ERRORBOUNDS) {
        if (s.agg_col < 0 || s.agg_col >= t->ncols) return fail(RJ_ERR_ARG, "aggregation column out of range");
        q->agg_slot = slot_of(s.agg_col);
        if (q->agg_slot < 0) return fail(RJ_ERR_LIMIT, "a query may use at most 5 attributes");
    } else {
        q->agg_type = RJ_AGG_COUNT;  // errorbounds counts only (RasterJoinBounds.cpp:124)
    }
    return RJ_OK;
}

// Rows a query has to read: its row range, cut to the time buckets its
// constraints on the table's time column allow (the constraints themselves
// are still applied, so the result is the same as over the whole range).
void rj_ctx::narrow_rows(const rj_table* t, const QuerySpec& s, int64_t* b, int64_t* e) const
{
    if (t->time_col < 0 || t->bucket_off.size() < 2) return;
    uint32_t lo = 0, hi = UINT32_MAX;
    bool none = false;
    for (const rj_constraint& c : s.cons) {
        if (c.attrib_id != t->time_col) continue;
        const uint32_t v = c.val.uval;
        switch (c.type) {
        case RJ_EQ: lo = std::max(lo, v); hi = std::min(hi, v); break;
        case RJ_LT: if (v == 0) none = true; else hi = std::min(hi, v - 1); break;
        case RJ_LTE: hi = std::min(hi, v); break;
        case RJ_GT: if (v == UINT32_MAX) none = true; else lo = std::max(lo, v + 1); break;
        case RJ_GTE: lo = std::max(lo, v); break;
        default: break;
        }
    }
    if (none || lo > hi) {
        *e = *b;
        return;
    }
    int64_t r0, r1;
    t->time_rows(lo, hi, &r0, &r1);
    *b = std::max(*b, r0);
    *e = std::min(*e, r1);
    if (*e < *b) *e = *b;
}

// Rows where every constraint of the query holds by the table's time
// buckets alone: the whole buckets inside its time window, when every
// constraint is on the (Uint) time column. *b == *e when there are none.
// Bucket k holds the times [first(k), last(k)]: bucket 0 also every time
// below t0, the last bucket every time from its start on (time_rows).
void rj_ctx::interior_rows(const rj_table* t, const QuerySpec& s, int64_t* b, int64_t* e) const
{
    *b = *e = 0;
    if (t->time_col < 0 || t->bucket_off.size() < 2 || s.cons.empty() || t->span == 0) return;
    if (t->types[t->time_col] != RJ_ATTR_UINT) return;
    uint64_t lo = 0, hi = UINT32_MAX;
    for (const rj_constraint& c : s.cons) {
        if (c.attrib_id != t->time_col) return;
        const uint64_t v = c.val.uval;
        switch (c.type) {
        case RJ_EQ: lo = std::max(lo, v); hi = std::min(hi, v); break;
        case RJ_LT: if (v == 0) return; hi = std::min(hi, v - 1); break;
        case RJ_LTE: hi = std::min(hi, v); break;
        case RJ_GT: if (v == UINT32_MAX) return; lo = std::max(lo, v + 1); break;
        case RJ_GTE: lo = std::max(lo, v); break;
        default: return;
        }
    }
    if (lo > hi) return;
    const int64_t nb = (int64_t)t->bucket_off.size() - 1;
    const uint64_t t0 = t->t0, span = t->span;
    // the first bucket whose times start at or after lo
    int64_t k0;
    if (lo == 0) k0 = 0;
    else if (lo <= t0) k0 = 1;
    else k0 = std::max<int64_t>(1, (int64_t)((lo - t0 + span - 1) / span));
    // the last bucket whose times end at or before hi
    int64_t k1;
    if (hi == UINT32_MAX) k1 = nb - 1;
    else if (hi + 1 < t0 + span) k1 = -1;
    else k1 = std::min<int64_t>(nb - 2, (int64_t)((hi + 1 - t0) / span) - 1);
    if (k0 > k1 || k0 >= nb) return;
    *b = t->bucket_off[k0];
    *e = t->bucket_off[k1 + 1];
}

int rj_ctx::ensure_canvas_buffers(const CanvasPlan& plan, bool sums)
{
    const size_t npx = (size_t)plan.W * plan.H;
    int rc = alloc(d_cnt, npx * 4, "canvas count");
    if (rc) return rc;
    if (sums) {
        if ((rc = alloc(d_chi, npx * 8, "canvas sum hi"))) return rc;
        if ((rc = alloc(d_clo, npx * 8, "canvas sum lo"))) return rc;
    }
    return launch("fill_canvas", [&] {
        launch_fill_u32(stream, d_cnt.as<uint32_t>(), npx, 0);
        if (sums) {
            launch_fill_u32(stream, d_chi.as<uint32_t>(), 2 * npx, 0);
            launch_fill_u32(stream, d_clo.as<uint32_t>(), 2 * npx, 0);
        }
    });
}

// Streaming of host tables. The chunk layout holds x,y and the columns a
// pass reads (need[c + 1]), each part 256-byte aligned; the others are not
// copied. Two device chunk buffers alternate: the copy into a buffer waits
// for the pass that last read it (ev_pass), the pass over it waits for its
// copy (ev_h2d).
int rj_ctx::stream_setup(rj_table* t, const bool* need, int64_t rows, StreamLayout* L)
{
    const size_t rowb = t->row_bytes();
    int64_t chunk = opt_stream_rows;
    if (chunk <= 0) {
        chunk = 32LL << 20;
        if (budget->limit > 0) {
            const int64_t avail = budget->limit - budget->used + (int64_t)(d_chunk[0].bytes + d_chunk[1].bytes);
            chunk = std::min<int64_t>(chunk, avail / 2 / (int64_t)rowb - 256);
        }
    }
    chunk = std::min<int64_t>(std::max<int64_t>(chunk, 1024), std::max<int64_t>(rows, 1));
    L->chunk = (chunk + 255) / 256 * 256;
    L->nparts = 0;
    L->total = 0;
    for (int part = 0; part <= t->ncols; part++) {
        if (!need[part]) continue;
        const int i = L->nparts++;
        L->part_of[i] = part;
        L->width[i] = part == 0 ? 8 : type_bytes(t->types[part - 1]);
        L->off[i] = L->total;
        L->total += al256((size_t)L->chunk * L->width[i]);
    }
    int rc;
    hipError_t e = hipSuccess;
    if (!copy_stream) {
        e = hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking);
        if (e != hipSuccess) return hip_fail(e, "copy stream");
    }
    if ((d_chunk[0].p && d_chunk[0].bytes < L->total) ||
        (h_chunk[0] && !t->hpinned && h_chunk_bytes[0] < L->total)) {
        // buffers about to be replaced may still be read by earlier copies
        e = hipStreamSynchronize(copy_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "streamed chunk");
    }
    for (int k = 0; k < 2; k++) {
        if ((rc = alloc(d_chunk[k], L->total, "streamed chunk"))) return rc;
        if (!t->hpinned && (rc = pinned(&h_chunk[k], &h_chunk_bytes[k], L->total, "pinned chunk"))) return rc;
        if (!ev_h2d[k]) {
            e = hipEventCreateWithFlags(&ev_h2d[k], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_pass[k], hipEventDisableTiming);
            // recorded once so that the first waits on them return at once
            if (e == hipSuccess) e = hipEventRecord(ev_h2d[k], copy_stream);
            if (e == hipSuccess) e = hipEventRecord(ev_pass[k], stream);
            if (e != hipSuccess) return hip_fail(e, "stream events");
        }
    }
    return RJ_OK;
}

// Copy rows [r, r + n) of t into device chunk k on the copy stream and make
// the context's stream wait for it. Pinned table memory is copied from
// directly; pageable memory goes through pinned staging buffer k, filled by
// host threads.
int rj_ctx::stream_copy(rj_table* t, const StreamLayout& L, int k, int64_t r, int64_t n)
{
    char* db = d_chunk[k].as<char>();
    hipError_t e = hipStreamWaitEvent(copy_stream, ev_pass[k], 0);
    if (t->hpinned) {
        for (int i = 0; i < L.nparts && e == hipSuccess; i++)
            e = hipMemcpyAsync(db + L.off[i], t->hdata + t->hoff[L.part_of[i]] + (size_t)r * L.width[i],
                               (size_t)n * L.width[i], hipMemcpyHostToDevice, copy_stream);
    } else {
        // staging buffer k is free once its previous copy is done
        if (e == hipSuccess) e = hipEventSynchronize(ev_h2d[k]);
        if (e != hipSuccess) return hip_fail(e, "streamed chunk");
        char* hb = static_cast<char*>(h_chunk[k]);
        const int nth = 8;
        auto copy_rows = [&](int i) {
            const int64_t a = n * i / nth, b = n * (i + 1) / nth;
            for (int j = 0; j < L.nparts; j++)
                std::memcpy(hb + L.off[j] + (size_t)a * L.width[j],
                            t->hdata + t->hoff[L.part_of[j]] + (size_t)(r + a) * L.width[j],
                            (size_t)(b - a) * L.width[j]);
        };
        std::vector<std::thread> th;
        for (int i = 0; i < nth; i++) th.emplace_back(copy_rows, i);
        for (auto& x : th) x.join();
        e = hipMemcpyAsync(db, hb, L.total, hipMemcpyHostToDevice, copy_stream);
    }
    if (e == hipSuccess) e = hipEventRecord(ev_h2d[k], copy_stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(stream, ev_h2d[k], 0);
    if (e != hipSuccess) return hip_fail(e, "streamed chunk");
    return RJ_OK;
}

// Point-pass arguments over device chunk k holding the rows from r0 on:
// pointers rebased so that row indices stay those of the table (rows outside
// the chunk are never read: the pass's row range and end lie inside it).
void rj_ctx::stream_args(const rj_table* t, const StreamLayout& L, int k, int64_t r0, PointArgs* a) const
{
    char* db = d_chunk[k].as<char>();
    a->xy = reinterpret_cast<const float*>(db + L.off[0]) - 2 * r0;
    for (int col = 0; col < t->ncols; col++) a->cols.col[col] = nullptr;
    for (int i = 1; i < L.nparts; i++) a->cols.col[L.part_of[i] - 1] = db + L.off[i] - (size_t)r0 * L.width[i];
}

// A point pass. For a host-resident table the rows [rb, re) stream through
// the device in chunks, as the reference's point passes do
// (GLFunction.cpp:329-337 sizes them, RasterJoin.cpp:214-236 loops over them):
// the copy of chunk k + 1 runs on its own stream while the pass over chunk k
// runs, and the per-polygon sums (or the canvas) accumulate across chunks.
int rj_ctx::point_pass(const char* name, int mode, PointArgs pa, rj_table* t, int64_t rb, int64_t re)
{
    if (!t->host) {
        pa.begin = rb;
        pa.end = re;
        return launch(name, [&] { launch_points(stream, mode, pa); });
    }
    if (re <= rb) return RJ_OK;
    bool need[RJ_MAX_COLUMNS + 1] = {true};
    for (int k = 0; k < pa.q.nslots; k++) need[pa.q.slot_col[k] + 1] = true;
    StreamLayout L;
    int rc = stream_setup(t, need, re - rb, &L);
    if (rc) return rc;
    // chunks start at multiples of 256 rows (aligned vector loads)
    int k = 0;
    for (int64_t c0 = rb - rb % 256; c0 < re; c0 += L.chunk, k ^= 1) {
        const int64_t c1 = std::min<int64_t>(c0 + L.chunk, re);
        if ((rc = stream_copy(t, L, k, c0, c1 - c0))) return rc;
        PointArgs c = pa;
        stream_args(t, L, k, c0, &c);
        c.begin = std::max(rb, c0);
        c.end = c1;
        if ((rc = launch(name, [&] { launch_points(stream, mode, c); }))) return rc;
        const hipError_t e = hipEventRecord(ev_pass[k], stream);
        if (e != hipSuccess) return hip_fail(e, "streamed chunk");
        stream_passes++;
    }
    return RJ_OK;
}

// One query with the single-query kernels: GLFunction::execute of the
// function type (RasterJoin, IndexJoin, HybridJoin, RasterJoinBounds).
int rj_ctx::run_single(rj_table* t, const QuerySpec& s, QueryOut* out)
{
    int fn = s.exec;
    if (fn < RJ_RASTER || fn > RJ_ERRORBOUNDS) return fail(RJ_ERR_UNSUPPORTED, "unknown function type");
    if (!s.P) return fail(RJ_ERR_STATE, "no polygons set");
    if (!t) return fail(RJ_ERR_STATE, "no points set");
    rj_polyset* P = s.P;
    DevQuery q;
    int rc = build_query(t, s, &q);
    if (rc) return rc;
    // the collection's structures may not be evicted while this query runs
    struct Pin {
        rj_polyset* P;
        explicit Pin(rj_polyset* p) : P(p) { P->pins++; }
        ~Pin() { P->pins--; }
    } pin(P);
    const bool avg = q.agg_type != RJ_AGG_COUNT;
    const int np = P->ps.np;
    int64_t rb0 = s.rb, re0 = s.re;
    narrow_rows(t, s, &rb0, &re0);
    const size_t res_words = (size_t)np * 7;
    if ((rc = alloc(d_res, (res_words + 2) * 8, "results"))) return rc;
    unsigned long long* res_base = d_res.as<unsigned long long>();
    DevResult res{res_base, (long long*)(res_base + np), res_base + 2 * (size_t)np};
    long long* bnd = (long long*)(res_base + 3 * (size_t)np);
    // the results and, in the word after them, the border-queue count of a
    // hybrid pass (read back with them)
    auto clear_results = [&] {
        return launch("fill_results", [&] { launch_fill_u32(stream, (uint32_t*)res_base, (res_words + 1) * 2, 0); });
    };
    if ((rc = clear_results())) return rc;

    PointArgs pa;
    std::memset(&pa, 0, sizeof(pa));
    pa.xy = static_cast<const float*>(t->xy);
    pa.begin = rb0;
    pa.end = re0;
    pa.cols = dev_cols(t);
    pa.q = q;
    pa.res = res;
    pa.lds_np = lds_direct_cap(avg);
    pa.cap = opt_cap;
    pa.wg_rows = t->organized ? kPassRowsPerWorkgroup : 0;
    interior_rows(t, s, &pa.in_b, &pa.in_e);
    Canvas* C = nullptr;
    Grid* G = nullptr;
    const unsigned long long* check_aux = nullptr;  // overlap counters of a fresh label map
    int* check_state = nullptr;                     // ... and the overlap state they settle
    const uint32_t* check_queue = nullptr;          // border-queue count of a hybrid pass
    out->path = RJ_PATH_SINGLE;
    last_exact = -1;

    auto canvas_pass = [&](bool sums) -> int {
        int r2 = ensure_canvas_buffers(C->plan, sums);
        if (r2) return r2;
        PointArgs sp = pa;
        sp.cv = dev_canvas(C->plan);
        sp.cnt = d_cnt.as<uint32_t>();
        sp.chi = sums ? d_chi.as<unsigned long long>() : nullptr;
        sp.clo = sums ? d_clo.as<unsigned long long>() : nullptr;
        return point_pass("splat", PT_SPLAT, sp, t, rb0, re0);
    };
    auto canvas_join = [&]() -> int {
        int r2 = canvas_pass(avg);
        if (r2) return r2;
        const DevCanvas cv = dev_canvas(C->plan);
        const DevPolys dp = dev_polys(P, C);
        return launch("band_sum", [&] {
            launch_band(stream, BAND_SUM, C->bands_sum, (int)C->pr.bands_sum.size(), dp, cv, d_cnt.as<uint32_t>(),
                        d_chi.as<long long>(), d_clo.as<unsigned long long>(), res, nullptr, nullptr, 0, nullptr,
                        C->pr.lds_sum, avg ? kBandSums : 0);
        });
    };

    // the error-bounds kernel over the count canvas; counts: also the
    // polygons' counts from the canvas (else a label pass wrote them)
    auto bounds_pass = [&](bool counts) -> int {
        const DevCanvas cv = dev_canvas(C->plan);
        const DevPolys dp = dev_polys(P, C);
        // narrow bands with their small LDS (more workgroups per CU), then the wide ones
        const int ns = C->pr.nbounds_small, nall = (int)C->pr.bands_bounds.size();
        return launch("band_bounds", [&] {
            launch_band(stream, BAND_BOUNDS, C->bands_bounds, ns, dp, cv, d_cnt.as<uint32_t>(), nullptr, nullptr, res,
                        bnd, nullptr, 0, nullptr, C->pr.lds_bounds_small, counts ? 0 : kBandNoCounts);
            launch_band(stream, BAND_BOUNDS, C->bands_bounds + ns, nall - ns, dp, cv, d_cnt.as<uint32_t>(), nullptr,
                        nullptr, res, bnd, nullptr, 0, nullptr, C->pr.lds_bounds, counts ? 0 : kBandNoCounts);
        });
    };

    auto index_join = [&]() -> int {
        int r2 = grid_for(P, s.index_res, &G);
        if (r2) return r2;
        pa.g = dev_grid(G);
        pa.P = dev_polys(P, nullptr);
        return point_pass("points_index", PT_INDEX, pa, t, rb0, re0);
    };

    if (fn == RJ_HYBRID) {
        rc = canvas_for(P, canvas_eps(s), s.ground_res, &C);
        if (!rc) rc = grid_for(P, s.index_res, &G);
        if (!rc) rc = hybrid_labels(P, C, G);
        // an index join whose label map does not fit the budget runs the
        // per-point grid test instead (the same results)
        if (rc == RJ_ERR_NOMEM && s.fn == RJ_INDEX) {
            C = nullptr;
            fn = RJ_INDEX;
        } else if (rc) {
            return rc;
        }
    }
    if (fn == RJ_INDEX) {
        if ((rc = index_join())) return rc;
    } else if (fn == RJ_HYBRID) {
        pa.cv = dev_canvas(C->plan);
        pa.g = dev_grid(G);
        pa.P = dev_polys(P, C);
        pa.label = C->d_hlab.p;
        pa.label16 = C->hyb16;
        pa.tshift = C->hyb_tshift;
        pa.tiles = C->d_htiles.as<uint32_t>();
        pa.tiles_w = pa.cv.LW >> pa.tshift;
        hybrid_args(C, &pa);
        // border points queued for a separate exact pass; a queue that runs
        // out of room is seen at the readback and the pass is rerun inline
        const int64_t rows = std::max<int64_t>(0, re0 - rb0);
        const uint32_t qcap = queue_cap(C, rows);
        if (opt_queue && !t->host && alloc(d_queue, (size_t)qcap * 16 + 256, "border queue") == RJ_OK) {
            pa.queue = d_queue.as<float4>();
            pa.qcount = reinterpret_cast<uint32_t*>(res_base + res_words);  // zeroed by clear_results
            pa.qcap = qcap;
            if ((rc = launch("points_hybrid", [&] { launch_points(stream, PT_HYBRID_QUEUE, pa); }))) return rc;
            if ((rc = launch("exact_queue", [&] { launch_exact(stream, pa); }))) return rc;
            check_queue = pa.qcount;
        } else {
            if ((rc = point_pass("points_hybrid", PT_LABEL_HYBRID, pa, t, rb0, re0))) return rc;
        }
    } else if (fn == RJ_RASTER) {
        if ((rc = canvas_for(P, canvas_eps(s), s.ground_res, &C))) return rc;
        const bool known_overlap = C->lab_built && cache && C->lab_state == 2;
        if (raster_path == 2 && known_overlap)
            return fail(RJ_ERR_UNSUPPORTED, "polygons overlap: the label-map path needs disjoint polygons");
        bool use_label = raster_path != 1 && !known_overlap;
        if (use_label) {
            bool fresh = false;
            if ((rc = raster_labels(P, C, &fresh))) return rc;
            const DevCanvas cv = dev_canvas(C->plan);
            pa.cv = cv;
            pa.P = dev_polys(P, C);
            pa.label = C->d_lab.p;
            pa.label16 = C->lab16;
            pa.tshift = C->lab_tshift;
            pa.tiles = C->d_tiles.as<uint32_t>();
            pa.tiles_w = cv.LW >> pa.tshift;
            if ((rc = point_pass("points_label", PT_LABEL, pa, t, rb0, re0))) return rc;
            if (C->lab_state == 0) {
                check_aux = tiles_aux(C->d_tiles, cv, C->lab_tshift);
                check_state = &C->lab_state;
            }
        } else {
            if ((rc = canvas_join())) return rc;
        }
    } else {
        if ((rc = canvas_for(P, canvas_eps(s), s.ground_res, &C))) return rc;
        if ((rc = bounds_for(P, C))) return rc;
        if (!C->pr.bounds_ok) return fail(RJ_ERR_UNSUPPORTED, "a polygon is too wide for the error-bounds kernel");
        const bool known_overlap = cache && ((C->blab_built && C->blab_state == 2) || (C->lab_built && C->lab_state == 2));
        if (raster_path == 2 && known_overlap)
            return fail(RJ_ERR_UNSUPPORTED, "polygons overlap: the label-map path needs disjoint polygons");
        if (raster_path != 1 && !known_overlap) {
            // the error-bounds label map: one pass sums the polygons' counts
            // and counts the points of boundary pixels per pixel
            if ((rc = bounds_labels(P, C))) return rc;
            if ((rc = ensure_canvas_buffers(C->plan, false))) return rc;
            const DevCanvas cv = dev_canvas(C->plan);
            pa.cv = cv;
            pa.P = dev_polys(P, C);
            pa.label = C->d_blab.p;
            pa.label16 = C->blab16;
            pa.tshift = C->blab_tshift;
            pa.tiles = C->d_btiles.as<uint32_t>();
            pa.tiles_w = cv.LW >> pa.tshift;
            pa.cnt = d_cnt.as<uint32_t>();
            if ((rc = point_pass("points_bounds", PT_LABEL_BOUNDS, pa, t, rb0, re0))) return rc;
            if ((rc = bounds_pass(false))) return rc;
            if (C->blab_state == 0) {
                check_aux = tiles_aux(C->d_btiles, cv, C->blab_tshift);
                check_state = &C->blab_state;
            }
        } else {
            if ((rc = canvas_pass(false))) return rc;
            if ((rc = bounds_pass(true))) return rc;
        }
    }

    // one pinned copy of [results | bounds] (+ the overlap counters of a
    // raster pass, or the border-queue count of a hybrid pass in the word
    // after the results: never both)
    const size_t res_bytes = (res_words + 3) * 8;
    if ((rc = pinned((void**)&h_res, &h_res_bytes, res_bytes, "pinned results"))) return rc;
    hipError_t e = hipSuccess;
    const size_t main_words = res_words + (check_queue ? 1 : 0);
    if (auto* dr = (unsigned long long*)host_dev(h_res)) {
        // one kernel writes the results (and counters) into pinned memory
        const CopyWords aux = check_aux ? CopyWords{check_aux, dr + res_words, 2} : CopyWords{nullptr, nullptr, 0};
        if ((rc = launch("read_results", [&] { launch_copy_words(stream, CopyWords{res_base, dr, main_words}, aux); })))
            return rc;
    } else {
        e = hipMemcpyAsync(h_res, res_base, main_words * 8, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess && check_aux)
            e = hipMemcpyAsync(h_res + res_words, check_aux, 16, hipMemcpyDeviceToHost, stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "execute");
    const uint32_t queued = check_queue ? (uint32_t)h_res[res_words] : 0;
    if (check_queue) last_exact = queued;
    if (check_queue && re0 - rb0 >= 65536)  // remember the border share for the next queue
        C->hyb_border = std::max(0.01, (double)queued / (double)(re0 - rb0));
    if (check_queue && queued > pa.qcap) {
        // the queue overflowed: the whole pass again with the exact test inline
        last_exact = -1;
        PointArgs ia = pa;
        ia.queue = nullptr;
        ia.qcount = nullptr;
        if ((rc = clear_results())) return rc;
        if ((rc = launch("points_hybrid", [&] { launch_points(stream, PT_LABEL_HYBRID, ia); }))) return rc;
        e = hipMemcpyAsync(h_res, res_base, res_words * 8, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "execute");
    }
    if (check_aux) {
        *check_state = (h_res[res_words] != h_res[res_words + 1]) ? 2 : 1;
        if (*check_state == 2) {
            // the fresh label map has pixels in two polygons: redo on the canvas
            if (raster_path == 2)
                return fail(RJ_ERR_UNSUPPORTED, "polygons overlap: the label-map path needs disjoint polygons");
            if ((rc = clear_results())) return rc;
            if (fn == RJ_ERRORBOUNDS) {
                if ((rc = canvas_pass(false))) return rc;
                if ((rc = bounds_pass(true))) return rc;
            } else if ((rc = canvas_join())) {
                return rc;
            }
            e = hipMemcpyAsync(h_res, res_base, res_words * 8, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_fail(e, "execute");
        }
    }
    out->counts.assign((const int64_t*)h_res, (const int64_t*)h_res + np);
    out->hi.assign((const int64_t*)h_res + np, (const int64_t*)h_res + 2 * np);
    out->lo.assign(h_res + 2 * np, h_res + 3 * np);
    if (fn == RJ_ERRORBOUNDS) out->bounds.assign((const int64_t*)h_res + 3 * np, (const int64_t*)h_res + 7 * np);
    else out->bounds.clear();
    if (!avg) {
        std::fill(out->hi.begin(), out->hi.end(), 0);
        std::fill(out->lo.begin(), out->lo.end(), 0);
    }
    return RJ_OK;
}

int rj_ctx::execute(int fn, int64_t* out_counts, double* out_sums, int32_t* legacy)
{
    if (fn < RJ_RASTER || fn > RJ_ERRORBOUNDS) return fail(RJ_ERR_UNSUPPORTED, "unknown function type");
    if (!polys) return fail(RJ_ERR_STATE, "no polygons set");
    if (!table) return fail(RJ_ERR_STATE, "no points set");
    const double t_start = now_ms();
    cur_ptmem = cur_polymem = cur_setup = cur_polyindex = 0;
    launches.clear();
    pool_used = 0;
    QuerySpec s;
    s.P = polys;
    s.fn = fn;
    set_exec(s);
    s.agg_type = agg_type;
    s.agg_col = agg_col;
    s.cons = cons;
    s.eps = eps;
    s.ground_res = ground_res;
    s.index_res = index_res;
    s.rb = rb;
    s.re = re;
    stream_passes = 0;
    int rc = run_single(table, s, &last);
    if (rc) return rc;
    last_fn = fn;
    const int np = polys->ps.np;
    const bool avg = agg_type != RJ_AGG_COUNT && fn != RJ_ERRORBOUNDS;
    for (int p = 0; p < np; p++) {
        if (out_counts) out_counts[p] = last.counts[p];
        const double sm = avg ? fixed_to_double(last.hi[p], last.lo[p]) : 0.0;
        if (out_sums) out_sums[p] = sm;
        if (legacy) {
            legacy[p] = (int32_t)last.counts[p];
            const double hundreds = std::trunc(sm / 100.0);
            legacy[np + p] = (int32_t)hundreds;
            legacy[2 * np + p] = (int32_t)((sm - 100.0 * hundreds) * 10.0);
        }
    }
    double pt_ms = 0, poly_ms = 0;
    if (timing) collect_times(&pt_ms, &poly_ms);
    Times& T = times[fn];
    T.execute.push_back(now_ms() - t_start);
    T.ptMem.push_back(cur_ptmem);
    T.ptRender.push_back(pt_ms);
    T.polyMem.push_back(cur_polymem);
    T.polyRender.push_back(poly_ms);
    T.setup.push_back(cur_setup);
    T.tri.push_back(0.0);
    T.polyIndex.push_back(cur_polyindex);
    T.backend.push_back(0.0);
    T.pts = re - rb;
    T.np = np;
    T.passes = (fn == RJ_INDEX || fn == RJ_HYBRID) ? index_res : std::max(1, stream_passes);
    T.ncons = (int)cons.size();
    return RJ_OK;
}

// ---------------------------------------------------------------------------
// fused batches
// ---------------------------------------------------------------------------

int spec_check(const rj_ctx* ctx, const rj_table* t, const rj_query* q, QuerySpec* s, std::string* why)
{
    auto bad = [&](int code, const char* m) {
        *why = m;
        return code;
    };
    if (!q || !q->polys) return bad(RJ_ERR_ARG, "query without a polygon collection");
    if (q->polys->ctx != ctx) return bad(RJ_ERR_ARG, "polygon collection of another context");
    if (q->fn < RJ_RASTER || q->fn > RJ_ERRORBOUNDS) return bad(RJ_ERR_UNSUPPORTED, "unknown function type");
    if (q->ncons < 0 || q->ncons > RJ_MAX_CONSTRAINTS) return bad(RJ_ERR_LIMIT, "at most 10 constraints");
    if (q->agg_type < RJ_AGG_COUNT || q->agg_type > RJ_AGG_SUM) return bad(RJ_ERR_ARG, "bad aggregation");
    if (q->agg_type != RJ_AGG_COUNT && (q->agg_col < 0 || q->agg_col >= t->ncols))
        return bad(RJ_ERR_ARG, "aggregation column out of range");
    for (int k = 0; k < q->ncons; k++)
        if (q->cons[k].attrib_id < 0 || q->cons[k].attrib_id >= t->ncols || q->cons[k].type < RJ_EQ ||
            q->cons[k].type > RJ_GTE)
            return bad(RJ_ERR_ARG, "bad constraint");
    if (q->fn != RJ_INDEX && !(q->eps_m > 0)) return bad(RJ_ERR_ARG, "accuracy must be > 0");
    if (q->ground_res < 0 || q->index_res < 0 || q->index_res > 16384) return bad(RJ_ERR_ARG, "bad query field");
    s->P = q->polys;
    s->fn = q->fn;
    ctx->set_exec(*s);
    s->agg_type = q->agg_type;
    s->agg_col = q->agg_col;
    s->cons.assign(q->cons, q->cons + q->ncons);
    s->eps = q->eps_m > 0 ? q->eps_m : 100.0;
    s->ground_res = q->ground_res > 0 ? q->ground_res : 1.0;
    s->index_res = q->index_res > 0 ? q->index_res : ctx->opt_auto_res;
    s->rb = q->row_begin;
    s->re = q->row_end < 0 ? t->n : q->row_end;
    if (s->rb < 0 || s->re < s->rb || s->re > t->n) return bad(RJ_ERR_ARG, "bad row range");
    return RJ_OK;
}

int spec_from_query(rj_ctx* ctx, const rj_table* t, const rj_query* q, QuerySpec* s)
{
    std::string why;
    const int rc = spec_check(ctx, t, q, s, &why);
    if (rc) ctx->err = why;
    return rc;
}

// Polygon structures of every query of a batch; sorts the queries into those
// the fused pass runs and those that run alone (errorbounds, the canvas path).
int rj_ctx::batch_prepare(rj_table* t, const QuerySpec* qs, int nq, std::vector<int>* fused,
                          std::vector<int>* alone, std::vector<int>* status)
{
    fused->clear();
    alone->clear();
    status->assign(nq, RJ_OK);
    std::vector<std::pair<int, Canvas*>> checks;  // queries on fresh label maps
    for (int i = 0; i < nq; i++) qs[i].P->pins++;  // released by the caller once the batch is done
    for (int i = 0; i < nq; i++) {
        const QuerySpec& s = qs[i];
        DevQuery dq;
        int rc = build_query(t, s, &dq);
        if (rc) {
            (*status)[i] = rc;
            continue;
        }
        if (s.exec == RJ_ERRORBOUNDS || (s.exec == RJ_RASTER && raster_path == 1)) {
            alone->push_back(i);
            continue;
        }
        Canvas* C = nullptr;
        Grid* G = nullptr;
        if (s.exec == RJ_INDEX) {
            rc = grid_for(s.P, s.index_res, &G);
        } else if (s.exec == RJ_HYBRID) {
            rc = canvas_for(s.P, canvas_eps(s), s.ground_res, &C);
            if (!rc) rc = grid_for(s.P, s.index_res, &G);
            if (!rc) rc = hybrid_labels(s.P, C, G);
            if (rc == RJ_ERR_NOMEM && s.fn == RJ_INDEX) {  // alone, on the grid index
                alone->push_back(i);
                continue;
            }
        } else {
            rc = canvas_for(s.P, canvas_eps(s), s.ground_res, &C);
            if (!rc && C->lab_built && C->lab_state == 2) {
                alone->push_back(i);
                continue;
            }
            bool fresh = false;
            if (!rc) rc = raster_labels(s.P, C, &fresh);
            if (!rc && C->lab_state == 0) checks.push_back({i, C});
        }
        if (rc) {
            (*status)[i] = rc;
            continue;
        }
        fused->push_back(i);
    }
    if (!checks.empty()) {
        // one wait for the overlap counters of every fresh label map
        std::vector<unsigned long long> a(2 * checks.size());
        hipError_t e = hipSuccess;
        for (size_t k = 0; k < checks.size() && e == hipSuccess; k++) {
            Canvas* C = checks[k].second;
            e = hipMemcpyAsync(&a[2 * k], tiles_aux(C->d_tiles, dev_canvas(C->plan), C->lab_tshift), 16,
                               hipMemcpyDeviceToHost, stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            for (int i = 0; i < nq; i++) qs[i].P->pins--;
            return hip_fail(e, "overlap test");
        }
        for (size_t k = 0; k < checks.size(); k++) {
            Canvas* C = checks[k].second;
            if (C->lab_state == 0) C->lab_state = (a[2 * k] != a[2 * k + 1]) ? 2 : 1;
        }
        std::vector<int> keep;
        for (int i : *fused) {
            bool overl = false;
            for (auto& c : checks)
                if (c.first == i && c.second->lab_state == 2) overl = true;
            if (overl) alone->push_back(i);
            else keep.push_back(i);
        }
        *fused = keep;
        if (raster_path == 2)
            for (int i : *alone)
                if (qs[i].exec == RJ_RASTER) (*status)[i] = RJ_ERR_UNSUPPORTED;
    }
    return RJ_OK;
}

// Fill a pass slot's block for `members` and launch the fused pass (and the
// copy of its results) on the stream. Returns without waiting.
// Segments of a fused pass over rows [lo, hi): elementary intervals of the
// members' row ranges with a constant query set, split into chunks of whole
// row groups (opt_batch_chunk_rows rows, fewer when the rows are too few to
// give every workgroup one).
void rj_ctx::batch_segments(BatchBlock& B, const std::vector<int64_t>& qb, const std::vector<int64_t>& qe, int64_t lo,
                            int64_t hi) const
{
    const size_t nm = qb.size();
    std::vector<int64_t> cut;
    for (size_t j = 0; j < nm; j++) {
        const int64_t b = std::max(qb[j], lo), e = std::min(qe[j], hi);
        if (e > b) {
            cut.push_back(b);
            cut.push_back(e);
        }
    }
    std::sort(cut.begin(), cut.end());
    cut.erase(std::unique(cut.begin(), cut.end()), cut.end());
    int64_t union_rows = 0;
    std::vector<BatchSeg> segs;
    for (size_t k = 0; k + 1 < cut.size(); k++) {
        unsigned long long m = 0;
        for (size_t j = 0; j < nm; j++)
            if (qb[j] <= cut[k] && qe[j] >= cut[k + 1] && qe[j] > qb[j]) m |= 1ull << j;
        if (!m) continue;
        if (!segs.empty() && segs.back().mask == m && segs.back().r1 == cut[k]) {
            segs.back().r1 = cut[k + 1];
        } else {
            segs.push_back(BatchSeg{cut[k], cut[k + 1], m, 0});
        }
        union_rows += cut[k + 1] - cut[k];
    }
    const int64_t group = 256LL * 8;  // whole row groups at 4 or 8 rows per thread
    int64_t chunk = std::min<int64_t>(opt_batch_chunk_rows, union_rows / std::max(1, opt_batch_wg));
    chunk = std::max<int64_t>(group, (chunk + group - 1) / group * group);
    int64_t nchunks = 0;
    for (auto& g : segs) {
        g.chunk0 = nchunks;
        nchunks += (g.r1 - g.r0 + chunk - 1) / chunk;
    }
    B.nseg = (int)segs.size();
    for (size_t k = 0; k < segs.size(); k++) B.seg[k] = segs[k];
    B.nchunks = nchunks;
    B.chunk_rows = chunk;
}

int rj_ctx::batch_launch(int slot, rj_table* t, const QuerySpec* qs, const std::vector<int>& members)
{
    PassSlot& S = slots[slot];
    int rc;
    if (!S.h_block) {
        size_t cur = 0;
        if ((rc = pinned((void**)&S.h_block, &cur, sizeof(BatchBlock), "pinned batch block"))) return rc;
    }
    // the block, then the pass's counters
    const size_t ctr_off = al256(sizeof(BatchBlock));
    if ((rc = alloc(S.d_block, ctr_off + al256(kBatchCounters * 4), "batch block"))) return rc;
    uint32_t* counter = reinterpret_cast<uint32_t*>(S.d_block.as<char>() + ctr_off);
    if (!S.done) (void)hipEventCreateWithFlags(&S.done, hipEventDisableTiming);
    BatchBlock& B = *S.h_block;
    std::memset(&B, 0, sizeof(B));
    // results: 3 np words per member
    S.members = members;
    S.res_off.assign(members.size(), 0);
    int64_t words = 0;
    for (size_t j = 0; j < members.size(); j++) {
        S.res_off[j] = words;
        words += 3 * (int64_t)qs[members[j]].P->ps.np;
    }
    S.res_words = words;
    if ((rc = alloc(S.d_res, (size_t)words * 8, "batch results"))) return rc;
    unsigned long long* dres = S.d_res.as<unsigned long long>();
    B.res = dres;
    B.res_words = words;
    B.nq = (int)members.size();
    // per query: the arguments of its single-query pass over the table
    std::vector<int64_t> qb(members.size()), qe(members.size());
    for (size_t j = 0; j < members.size(); j++) {
        const QuerySpec& s = qs[members[j]];
        rj_polyset* P = s.P;
        const int np = P->ps.np;
        PointArgs& a = B.q[j];
        if ((rc = build_query(t, s, &a.q))) return rc;
        a.xy = static_cast<const float*>(t->xy);
        a.begin = 0;
        a.end = t->n;
        a.cols = dev_cols(t);
        a.P = dev_polys(P, nullptr);
        Grid* G = nullptr;
        if (s.exec != RJ_RASTER) {
            auto gi = P->grids.find(s.index_res);
            if (gi == P->grids.end()) return fail(RJ_ERR_STATE, "grid of a fused query is not resident");
            G = gi->second.get();
            a.g = dev_grid(G);
        }
        if (s.exec == RJ_INDEX) {
            B.mode[j] = PT_INDEX;
        } else {
            CanvasPlan plan;
            make_plan(P->ps.bbox, canvas_eps(s), s.ground_res, &plan);
            auto ci = P->canvases.find(plan.cell);
            if (ci == P->canvases.end()) return fail(RJ_ERR_STATE, "canvas of a fused query is not resident");
            Canvas* C = ci->second.get();
            a.cv = dev_canvas(C->plan);
            a.P = dev_polys(P, C);
            if (s.exec == RJ_HYBRID) {
                B.mode[j] = PT_LABEL_HYBRID;
                a.label = C->d_hlab.p;
                a.label16 = C->hyb16;
                a.tshift = C->hyb_tshift;
                a.tiles = C->d_htiles.as<uint32_t>();
                hybrid_args(C, &a);
            } else {
                B.mode[j] = PT_LABEL;
                a.label = C->d_lab.p;
                a.label16 = C->lab16;
                a.tshift = C->lab_tshift;
                a.tiles = C->d_tiles.as<uint32_t>();
            }
            a.tiles_w = a.cv.LW >> a.tshift;
        }
        unsigned long long* r = dres + S.res_off[j];
        a.res = DevResult{r, (long long*)(r + np), r + 2 * (size_t)np};
        qb[j] = s.rb;
        qe[j] = s.re;
        narrow_rows(t, s, &qb[j], &qe[j]);
        interior_rows(t, s, &a.in_b, &a.in_e);
    }
    // border queues of the hybrid members: one region each, 1/8 of its rows
    // (at least 65536); members beyond kQueueEntriesMax queued points in all test inline.
    // Passes over a host table test inline (a queue per streamed chunk would
    // need its own exact pass).
    S.qcap.assign(members.size(), 0);
    S.qrows.assign(members.size(), 0);
    S.qcanvas.assign(members.size(), nullptr);
    if (opt_queue && !t->host) {
        int64_t total = 0;
        for (size_t j = 0; j < members.size(); j++) {
            if (B.mode[j] != PT_LABEL_HYBRID) continue;
            const QuerySpec& s = qs[members[j]];
            CanvasPlan plan;
            make_plan(s.P->ps.bbox, canvas_eps(s), s.ground_res, &plan);
            Canvas* C = s.P->canvases.at(plan.cell).get();
            S.qcanvas[j] = C;
            S.qrows[j] = qe[j] - qb[j];
            const int64_t cap = queue_cap(C, qe[j] - qb[j]);
            if (total + cap > kQueueEntriesMax) continue;
            S.qcap[j] = (uint32_t)cap;
            total += cap;
        }
        if (total > 0 && alloc(S.d_queue, (size_t)total * 16, "batch border queues") != RJ_OK) {
            std::fill(S.qcap.begin(), S.qcap.end(), 0u);
            total = 0;
        }
        int64_t off = 0;
        for (size_t j = 0; j < members.size(); j++) {
            if (!S.qcap[j]) continue;
            PointArgs& a = B.q[j];
            B.mode[j] = PT_HYBRID_QUEUE;
            a.queue = S.d_queue.as<float4>() + off;
            a.qcount = counter + 2 + j;
            a.qcap = S.qcap[j];
            off += S.qcap[j];
        }
    }

    // one persistent kernel per class of members
    unsigned classes = 0;
    for (size_t j = 0; j < members.size(); j++) {
        const int c = B.mode[j] == PT_LABEL    ? BC_LABEL
                      : B.mode[j] == PT_INDEX  ? BC_INDEX
                      : B.mode[j] == PT_HYBRID_QUEUE ? BC_HYBRID
                                                     : BC_HYBRID_INLINE;
        B.cls[c] |= 1ull << j;
        classes |= 1u << c;
    }
    // groups: raster members that read the same 16-bit label map, and
    // hybrid members (queue mode) that read the same 16-bit hybrid map, share
    // each chunk's labels (computed once into LDS)
    const int64_t chunk_max = (opt_batch_chunk_rows + 2047) / 2048 * 2048;
    B.ngroups = 0;
    for (size_t j = 0; j < members.size(); j++) B.grp[j] = -1;
    if (opt_batch_share && chunk_max <= kGroupChunkMax) {
        for (size_t j = 0; j < members.size(); j++) {
            if ((B.mode[j] != PT_LABEL && B.mode[j] != PT_HYBRID_QUEUE) || !B.q[j].label16 || B.grp[j] >= 0)
                continue;
            unsigned long long m = 1ull << j;
            for (size_t i = j + 1; i < members.size(); i++)
                if (B.mode[i] == B.mode[j] && B.q[i].label == B.q[j].label) m |= 1ull << i;
            if (__builtin_popcountll(m) < 2) continue;
            const int g = B.ngroups++;
            B.gmask[g] = m;
            for (size_t i = j; i < members.size(); i++)
                if ((m >> i) & 1ull) B.grp[i] = g;
        }
    }
    // LDS sums: direct for small collections, hashed for the others; sized
    // for the largest need among the members; then the group keys
    const int cap = lds_direct_cap(true);
    B.lds_cap = cap;
    size_t lds = 0;
    for (int i : members) lds = std::max(lds, acc_lds_bytes(qs[i].P->ps.np, cap, true));
    lds = (lds + 15) / 16 * 16;
    B.key_off = (int32_t)lds;
    if (B.ngroups) lds += (size_t)chunk_max * 2;
    const int rows = opt_batch_rows;
    const int crow[BC_COUNT] = {rows, opt_batch_hybrid_rows, rows, rows};
    const int nwg = opt_batch_wg;
    const BatchBlock* dblock = S.d_block.as<BatchBlock>();
    if (t->host) {
        if ((rc = batch_stream(S, t, qb, qe, counter, lds, classes))) return rc;
    } else {
        batch_segments(B, qb, qe, 0, t->n);
        const void* xy = t->xy;
        if (opt_graph) {
            // the pass as one captured HIP graph: block upload, result clear and
            // the fused kernel; everything that varies lives in the block
            if (!S.graph || S.graph_key[0] != xy || S.graph_key[1] != dblock || S.graph_rows != rows ||
                S.graph_wg != nwg || S.graph_nrows != t->n || S.graph_lds != lds || S.graph_classes != classes ||
                S.graph_hrows != opt_batch_hybrid_rows) {
                if (S.graph) (void)hipGraphExecDestroy(S.graph);
                S.graph = nullptr;
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) return hip_fail(e, "graph capture");
                (void)hipMemcpyAsync(S.d_block.p, S.h_block, sizeof(BatchBlock), hipMemcpyHostToDevice, stream);
                launch_batch(stream, dblock, counter, crow, nwg, lds, classes);
                e = hipStreamEndCapture(stream, &g);
                if (e != hipSuccess) return hip_fail(e, "graph capture");
                e = hipGraphInstantiate(&S.graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) {
                    S.graph = nullptr;
                    return hip_fail(e, "graph instantiate");
                }
                S.graph_key[0] = xy;
                S.graph_key[1] = dblock;
                S.graph_rows = rows;
                S.graph_wg = nwg;
                S.graph_nrows = t->n;
                S.graph_lds = lds;
                S.graph_classes = classes;
                S.graph_hrows = opt_batch_hybrid_rows;
            }
            if ((rc = launch("batch", [&] { (void)hipGraphLaunch(S.graph, stream); }))) return rc;
        } else {
            hipError_t e = hipMemcpyAsync(S.d_block.p, S.h_block, sizeof(BatchBlock), hipMemcpyHostToDevice, stream);
            if (e != hipSuccess) return hip_fail(e, "batch block");
            if ((rc = launch("batch", [&] { launch_batch(stream, dblock, counter, crow, nwg, lds, classes); })))
                return rc;
        }
    }
    if ((rc = pinned((void**)&S.h_res, &S.h_res_cap, ((size_t)words + kBatchCounters) * 8, "pinned batch results")))
        return rc;
    if (words) {
        hipError_t e = hipMemcpyAsync(S.h_res, S.d_res.p, (size_t)words * 8, hipMemcpyDeviceToHost, stream);
        if (e != hipSuccess) return hip_fail(e, "batch results");
    }
    {
        hipError_t e = hipMemcpyAsync(S.h_res + words, counter, kBatchCounters * 4, hipMemcpyDeviceToHost, stream);
        if (e != hipSuccess) return hip_fail(e, "batch counters");
    }
    (void)hipEventRecord(S.done, stream);
    S.busy = true;
    return RJ_OK;
}

// A fused pass over a host table: the union of the members' rows streams
// through the device chunk buffers once for all members (the columns any
// member reads), and the fused kernel runs over each chunk with the members'
// arguments rebased onto it; results accumulate across chunks (cleared by the
// first chunk's launch only).
int rj_ctx::batch_stream(PassSlot& S, rj_table* t, const std::vector<int64_t>& qb, const std::vector<int64_t>& qe,
                         uint32_t* counter, size_t lds, unsigned classes)
{
    BatchBlock& B = *S.h_block;
    const int nm = B.nq;
    int64_t lo = t->n, hi = 0;
    bool need[RJ_MAX_COLUMNS + 1] = {true};
    for (int j = 0; j < nm; j++) {
        if (qe[j] > qb[j]) {
            lo = std::min(lo, qb[j]);
            hi = std::max(hi, qe[j]);
        }
        for (int k = 0; k < B.q[j].q.nslots; k++) need[B.q[j].q.slot_col[k] + 1] = true;
    }
    if (hi <= lo) {
        lo = hi = 0;
    }
    StreamLayout L;
    int rc = stream_setup(t, need, hi - lo, &L);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    if (!S.blk) e = hipEventCreateWithFlags(&S.blk, hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(e, "stream events");
    std::vector<PointArgs> base(B.q, B.q + nm);
    const int64_t words = B.res_words;
    const BatchBlock* dblock = S.d_block.as<BatchBlock>();
    int k = 0;
    bool first = true;
    for (int64_t c0 = lo - lo % 256; first || c0 < hi; c0 += L.chunk, k ^= 1) {
        const int64_t c1 = std::min<int64_t>(c0 + L.chunk, hi);
        if (c1 > c0 && (rc = stream_copy(t, L, k, c0, c1 - c0))) return rc;
        // the previous chunk's block has reached the device
        if (!first && (e = hipEventSynchronize(S.blk)) != hipSuccess) return hip_fail(e, "batch block");
        for (int j = 0; j < nm; j++) {
            B.q[j] = base[j];
            stream_args(t, L, k, c0, &B.q[j]);
            B.q[j].begin = c0;
            B.q[j].end = c1;
        }
        batch_segments(B, qb, qe, c0, c1);
        B.res_words = first ? words : 0;
        e = hipMemcpyAsync(S.d_block.p, S.h_block, sizeof(BatchBlock), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipEventRecord(S.blk, stream);
        if (e != hipSuccess) return hip_fail(e, "batch block");
        const int crow[BC_COUNT] = {opt_batch_rows, opt_batch_hybrid_rows, opt_batch_rows, opt_batch_rows};
        if ((rc = launch("batch", [&] { launch_batch(stream, dblock, counter, crow, opt_batch_wg, lds, classes); })))
            return rc;
        if ((e = hipEventRecord(ev_pass[k], stream)) != hipSuccess) return hip_fail(e, "streamed chunk");
        stream_passes++;
        first = false;
        if (c1 >= hi) break;
    }
    // the host block must not change while its last upload is pending
    if ((e = hipEventSynchronize(S.blk)) != hipSuccess) return hip_fail(e, "batch block");
    return RJ_OK;
}

// Wait for a pass slot and decode its results.
int rj_ctx::batch_finish(int slot, rj_table* t, const QuerySpec* qs, QueryOut* outs)
{
    PassSlot& S = slots[slot];
    if (!S.busy) return RJ_OK;
    hipError_t e = hipEventSynchronize(S.done);
    S.busy = false;
    if (e != hipSuccess) return hip_fail(e, "fused pass");
    const uint32_t* qcount = reinterpret_cast<const uint32_t*>(S.h_res + S.res_words) + 2;
    for (size_t j = 0; j < S.members.size(); j++) {
        const QuerySpec& s = qs[S.members[j]];
        QueryOut& o = outs[S.members[j]];
        if (S.qcap[j] && S.qrows[j] >= 65536)
            S.qcanvas[j]->hyb_border = std::max(0.01, (double)qcount[j] / (double)S.qrows[j]);
        if (S.qcap[j] && qcount[j] > S.qcap[j]) {
            // its border queue overflowed: the query again on its own
            const int rc = run_single(t, s, &o);
            if (rc) return rc;
            continue;
        }
        const int np = s.P->ps.np;
        const unsigned long long* r = S.h_res + S.res_off[j];
        o.counts.assign((const int64_t*)r, (const int64_t*)r + np);
        o.hi.assign((const int64_t*)r + np, (const int64_t*)r + 2 * np);
        o.lo.assign(r + 2 * np, r + 3 * np);
        o.bounds.clear();
        o.path = RJ_PATH_FUSED;
    }
    return RJ_OK;
}

// Fused queries in passes of at most max_batch queries, in query order; a
// pass also ends where its hybrid members' border queues would pass
// kQueueEntriesMax entries in all (members beyond it would test inline, which
// is slower than another pass).
void rj_ctx::group_passes(rj_table* t, const QuerySpec* qs, const std::vector<int>& fused, int max_batch,
                          std::vector<std::vector<int>>* passes)
{
    passes->clear();
    int64_t entries = 0;
    for (int i : fused) {
        const QuerySpec& s = qs[i];
        int64_t cap = 0;
        if (s.exec == RJ_HYBRID && opt_queue && !t->host) {
            CanvasPlan plan;
            make_plan(s.P->ps.bbox, canvas_eps(s), s.ground_res, &plan);
            auto ci = s.P->canvases.find(plan.cell);
            int64_t rb = s.rb, re = s.re;
            narrow_rows(t, s, &rb, &re);
            if (ci != s.P->canvases.end()) cap = queue_cap(ci->second.get(), re - rb);
        }
        if (passes->empty() || (int)passes->back().size() >= max_batch ||
            (entries + cap > kQueueEntriesMax && entries > 0)) {
            passes->emplace_back();
            entries = 0;
        }
        passes->back().push_back(i);
        entries += cap;
    }
}

int rj_ctx::execute_batch(rj_table* t, const QuerySpec* qs, int nq, QueryOut* outs, int* status)
{
    launches.clear();
    pool_used = 0;
    std::vector<int> fused, alone, st;
    int rc = batch_prepare(t, qs, nq, &fused, &alone, &st);
    if (rc) return rc;
    std::vector<std::vector<int>> passes;
    group_passes(t, qs, fused, kBatchMax, &passes);
    for (size_t p = 0; p < passes.size() && !rc; p++) {
        const int slot = (int)(p & 1);
        rc = batch_finish(slot, t, qs, outs);
        if (!rc) rc = batch_launch(slot, t, qs, passes[p]);
    }
    for (int slot = 0; slot < 2; slot++) {
        const int r2 = batch_finish(slot, t, qs, outs);
        if (!rc) rc = r2;
    }
    for (size_t k = 0; k < alone.size() && !rc; k++) {
        const int i = alone[k];
        if (st[i] != RJ_OK) continue;
        st[i] = run_single(t, qs[i], &outs[i]);
    }
    for (int i = 0; i < nq; i++) qs[i].P->pins--;
    if (rc) return rc;
    for (int i = 0; i < nq; i++) status[i] = st[i];
    if (timing) {
        double pt = 0, poly = 0;
        collect_times(&pt, &poly);
    }
    return RJ_OK;
}

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------

// Times the fused pass of a batch over a grid of launch shapes and label tile
// sizes (and, for queries with the automatic grid resolution, grid
// resolutions), checks that every candidate returns the results of the first
// one exactly, and keeps the fastest. chosen: tile shift, rows per thread,
// workgroups, chunks per workgroup, automatic grid resolution.
int rj_ctx::autotune(rj_table* t, const QuerySpec* qs_in, int nq, const std::vector<int>& auto_res_q,
                     int32_t* chosen)
{
    if (nq <= 0) return fail(RJ_ERR_ARG, "nothing to tune");
    std::vector<QuerySpec> qs(qs_in, qs_in + nq);
    const bool saved_timing = timing, saved_pass_only = timing_pass_only;
    timing = true;
    timing_pass_only = false;
    const int base_tshift = opt_tshift, base_rows = opt_batch_rows, base_wg = opt_batch_wg,
              base_chunks = opt_batch_chunk_rows, base_res = opt_auto_res;
    std::vector<QueryOut> ref(nq), cur(nq);
    std::vector<int> st(nq);
    bool have_ref = false;
    double best = 1e30;
    int32_t bestc[5] = {raster_tshift(), base_rows, base_wg, base_chunks, base_res};
    auto run = [&](double* ms) -> int {
        double m = 1e30;
        for (int k = 0; k < 3; k++) {
            int rc = execute_batch(t, qs.data(), nq, cur.data(), st.data());
            if (rc) return rc;
            double sum = 0;
            for (auto& ks : kstats)
                if (ks.name == "batch") sum += ks.ms;
            m = std::min(m, sum);
        }
        for (int i = 0; i < nq; i++)
            if (st[i]) return fail(st[i], "autotune: a query failed");
        *ms = m;
        return RJ_OK;
    };
    auto same = [&]() {
        for (int i = 0; i < nq; i++)
            if (cur[i].counts != ref[i].counts || cur[i].hi != ref[i].hi || cur[i].lo != ref[i].lo) return false;
        return true;
    };
    int rc = RJ_OK;
    const int tsh[3] = {3, 4, 5};
    const int rws[2] = {4, 8};
    const int wgs[3] = {1024, 2048, 4096};
    const int chs[3] = {4096, 8192, 16384};
    for (int a = 0; a < 3 && !rc; a++) {
        opt_tshift = tsh[a];
        for (int b = 0; b < 2 && !rc; b++)
            for (int c = 0; c < 3 && !rc; c++)
                for (int d = 0; d < 3 && !rc; d++) {
                    opt_batch_rows = rws[b];
                    opt_batch_wg = wgs[c];
                    opt_batch_chunk_rows = chs[d];
                    double ms = 0;
                    if ((rc = run(&ms))) break;
                    if (!have_ref) {
                        ref = cur;
                        have_ref = true;
                    } else if (!same()) {
                        rc = fail(RJ_ERR_DEVICE, "autotune: results differ between launch shapes");
                        break;
                    }
                    if (ms < best) {
                        best = ms;
                        bestc[0] = tsh[a]; bestc[1] = rws[b]; bestc[2] = wgs[c]; bestc[3] = chs[d];
                    }
                }
    }
    opt_tshift = bestc[0];
    opt_batch_rows = bestc[1];
    opt_batch_wg = bestc[2];
    opt_batch_chunk_rows = bestc[3];
    // rows per thread of the hybrid kernel, with the rest fixed
    const int base_hrows = opt_batch_hybrid_rows;
    for (int hr : {4, 8}) {
        if (rc || hr == base_hrows) continue;
        opt_batch_hybrid_rows = hr;
        double ms = 0;
        if ((rc = run(&ms))) break;
        if (!same()) {
            rc = fail(RJ_ERR_DEVICE, "autotune: results differ between launch shapes");
            break;
        }
        if (ms >= best) opt_batch_hybrid_rows = base_hrows;
        else best = ms;
    }
    if (!rc && !auto_res_q.empty()) {
        // grid resolution of queries that leave it to the engine: a
        // resolution is kept only where the results equal those at 256
        const int res_c[4] = {128, 256, 512, 1024};
        std::vector<double> ms_r(4, 1e30);
        std::vector<std::vector<QueryOut>> outs_r(4);
        for (int k = 0; k < 4 && !rc; k++) {
            for (int i : auto_res_q) qs[i].index_res = res_c[k];
            if ((rc = run(&ms_r[k]))) break;
            outs_r[k] = cur;
        }
        int pick = 1;
        for (int k = 0; k < 4 && !rc; k++) {
            bool eq = true;
            for (int i = 0; i < nq; i++)
                if (outs_r[k][i].counts != outs_r[1][i].counts || outs_r[k][i].hi != outs_r[1][i].hi ||
                    outs_r[k][i].lo != outs_r[1][i].lo)
                    eq = false;
            if (eq && ms_r[k] < ms_r[pick]) pick = k;
        }
        bestc[4] = res_c[pick];
        opt_auto_res = bestc[4];
    }
    timing = saved_timing;
    timing_pass_only = saved_pass_only;
    if (rc) {
        opt_tshift = base_tshift;
        opt_batch_hybrid_rows = base_hrows;
        opt_batch_rows = base_rows;
        opt_batch_wg = base_wg;
        opt_batch_chunk_rows = base_chunks;
        opt_auto_res = base_res;
        return rc;
    }
    for (int k = 0; k < 5; k++)
        if (chosen) chosen[k] = bestc[k];
    return RJ_OK;
}
