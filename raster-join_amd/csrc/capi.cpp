// capi.cpp -- the join context behind the C ABI (include/rasterjoin.h).
//
// One rj_ctx plays the role of the reference's GLHandler + GLFunction objects
// (RasterJoin/join/GLHandler.cpp, GLFunction.cpp): it owns the device
// resources, the memory budget and the per-function timing vectors, and runs
// the four join types as sequences of HIP kernels on its own stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rasterjoin.h"
#include "rj_impl.hpp"

namespace rj {
size_t organize_scratch_bytes(int64_t n);
int organize_rows(hipStream_t s, int64_t n, float* xy, void** cols, const int32_t* types, int ncols,
                  const float bbox[4], void* scratch);
}  // namespace rj

using namespace rj;

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

size_t type_bytes(int t) { return t == RJ_ATTR_UINT8 ? 1 : 4; }

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int64_t* used = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) {
            (void)hipFree(p);
            if (used) *used -= (int64_t)bytes;
        }
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct rj_table {
    int device = 0;
    int64_t n = 0;
    int32_t ncols = 0;
    int32_t types[RJ_MAX_COLUMNS] = {0};
    void* xy = nullptr;
    void* cols[RJ_MAX_COLUMNS] = {nullptr};
    float bbox[4] = {0, 0, 0, 0};
    bool organized = false;
    ~rj_table()
    {
        if (xy) (void)hipFree(xy);
        for (int c = 0; c < ncols; c++)
            if (cols[c]) (void)hipFree(cols[c]);
    }
    size_t bytes() const
    {
        size_t b = (size_t)n * 8;
        for (int c = 0; c < ncols; c++) b += (size_t)n * type_bytes(types[c]);
        return b;
    }
};

struct rj_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t budget = 0;  // bytes; <= 0: unlimited
    int64_t used = 0;
    int inmem = 0;
    std::string err;

    // polygons
    PolySet polys;
    bool has_polys = false;
    DevBuf d_pxy, d_poff, d_prev;

    // points
    rj_table* table = nullptr;
    bool own_table = false;
    int64_t rb = 0, re = 0;

    // query
    std::vector<rj_constraint> cons;
    int agg_type = RJ_AGG_COUNT, agg_col = -1;
    double eps = 100.0, ground_res = 1.0;
    int index_res = 256;

    // options
    int raster_path = 0;
    bool timing = false;
    int hybrid_dilation = 0;
    bool cache = true;
    bool organize = true;

    // canvas-dependent polygon data
    bool raster_valid = false;
    CanvasPlan plan;
    PolyRaster pr;
    DevBuf d_pc;                      // regions + bands (one allocation)
    DevBuf d_XY;                      // snapped vertices: X then Y
    void* h_stage = nullptr;          // pinned staging buffer for its upload
    size_t h_stage_bytes = 0;
    int32_t *p_X = nullptr, *p_Y = nullptr, *p_region = nullptr;
    Band *p_bands_sum = nullptr, *p_bands_bounds = nullptr;

    // CPUJoin grid
    bool grid_valid = false;
    GridIndex grid;
    DevBuf d_cell_off, d_cand;

    // label maps
    bool label_raster_valid = false;
    bool label_raster_conflict = false;
    DevBuf d_label_raster, d_tiles_raster;
    bool label_hybrid_valid = false;
    int label_hybrid16 = 0;
    DevBuf d_label_hybrid, d_tiles_hybrid, d_mask, d_rects;

    // canvas
    DevBuf d_cnt, d_chi, d_clo;

    // results
    DevBuf d_res;
    unsigned long long* h_res = nullptr;  // pinned copy of d_res
    size_t h_res_bytes = 0;
    std::vector<int64_t> counts, hi, bounds;
    std::vector<uint64_t> lo;
    int last_fn = -1;

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

    ~rj_ctx()
    {
        if (own_table) delete table;
        if (h_stage) (void)hipHostFree(h_stage);
        if (h_res) (void)hipHostFree(h_res);
        for (auto e : pool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int fail(int code, const std::string& m)
    {
        err = m;
        return code;
    }

    int hip_fail(hipError_t e, const char* what)
    {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return (e == hipErrorOutOfMemory) ? RJ_ERR_NOMEM : RJ_ERR_DEVICE;
    }

    // Device allocation counted against the budget; reuses a large-enough buffer.
    int alloc(DevBuf& b, size_t bytes, const char* what)
    {
        if (bytes == 0) bytes = 8;
        if (b.p && b.bytes >= bytes) return RJ_OK;
        b.release();
        if (budget > 0 && used + (int64_t)bytes > budget)
            return fail(RJ_ERR_NOMEM, std::string("memory budget exhausted allocating ") + what + " (" +
                                          std::to_string(bytes) + " bytes, " + std::to_string(used) +
                                          " of " + std::to_string(budget) + " in use)");
        hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) {
            b.p = nullptr;
            (void)hipGetLastError();
            return hip_fail(e, what);
        }
        b.bytes = bytes;
        b.used = &used;
        used += (int64_t)bytes;
        return RJ_OK;
    }

    int upload(DevBuf& b, const void* src, size_t bytes, const char* what)
    {
        int rc = alloc(b, bytes, what);
        if (rc) return rc;
        if (bytes == 0) return RJ_OK;
        hipError_t e = hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return hip_fail(e, what);
        return RJ_OK;
    }

    hipEvent_t event()
    {
        if (pool_used == pool.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[pool_used++];
    }

    // Launch bracket for per-kernel device timing.
    template <class F>
    int launch(const char* name, F f)
    {
        hipEvent_t a = nullptr, b = nullptr;
        if (timing) {
            a = event();
            b = event();
            (void)hipEventRecord(a, stream);
        }
        f();
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, name);
        if (timing) {
            (void)hipEventRecord(b, stream);
            launches.push_back({name, a, b});
        }
        return RJ_OK;
    }

    void collect_times(double* pt_ms, double* poly_ms)
    {
        kstats.clear();
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
            if (l.name.rfind("points", 0) == 0 || l.name.rfind("splat", 0) == 0) *pt_ms += ms;
            else *poly_ms += ms;
        }
        launches.clear();
        pool_used = 0;
    }

    void invalidate_polys()
    {
        raster_valid = grid_valid = label_raster_valid = label_hybrid_valid = false;
    }

    DevCols dev_cols() const
    {
        DevCols c;
        std::memset(&c, 0, sizeof(c));
        for (int k = 0; k < table->ncols; k++) {
            c.col[k] = table->cols[k];
            c.type[k] = table->types[k];
        }
        return c;
    }

    DevPolys dev_polys() const
    {
        DevPolys P;
        P.np = polys.np;
        P.xy = d_pxy.as<float>();
        P.off = d_poff.as<int32_t>();
        P.X = p_X;
        P.Y = p_Y;
        P.region = p_region;
        return P;
    }

    DevGrid dev_grid() const
    {
        DevGrid g;
        g.xs = grid.xs; g.ys = grid.ys;
        g.minx = grid.minx; g.miny = grid.miny; g.width = grid.width; g.height = grid.height;
        g.cell_off = d_cell_off.as<int32_t>();
        g.cand = d_cand.as<int32_t>();
        return g;
    }

    int build_query(DevQuery* q)
    {
        std::memset(q, 0, sizeof(*q));
        q->agg_slot = -1;
        if ((int)cons.size() > RJ_MAX_CONSTRAINTS) return fail(RJ_ERR_LIMIT, "at most 10 constraints");
        auto slot_of = [&](int col) {
            for (int s = 0; s < q->nslots; s++)
                if (q->slot_col[s] == col) return s;
            if (q->nslots == kMaxSlots) return -1;
            q->slot_col[q->nslots] = col;
            return q->nslots++;
        };
        for (size_t k = 0; k < cons.size(); k++) {
            const rj_constraint& c = cons[k];
            if (c.attrib_id < 0 || c.attrib_id >= table->ncols)
                return fail(RJ_ERR_ARG, "constraint column " + std::to_string(c.attrib_id) + " out of range");
            if (c.type < RJ_EQ || c.type > RJ_GTE) return fail(RJ_ERR_ARG, "bad constraint type");
            q->con_col[k] = c.attrib_id;
            q->con_op[k] = c.type;
            q->con_val[k] = c.val.uval;
            q->con_slot[k] = slot_of(c.attrib_id);
            if (q->con_slot[k] < 0) return fail(RJ_ERR_LIMIT, "a query may use at most 5 attributes");
        }
        q->ncons = (int)cons.size();
        q->agg_type = agg_type;
        q->agg_col = 0;
        if (agg_type != RJ_AGG_COUNT) {
            if (agg_col < 0 || agg_col >= table->ncols)
                return fail(RJ_ERR_ARG, "aggregation column out of range");
            q->agg_col = agg_col;
            q->agg_slot = slot_of(agg_col);
            if (q->agg_slot < 0) return fail(RJ_ERR_LIMIT, "a query may use at most 5 attributes");
        }
        return RJ_OK;
    }

    // plan + snapped polygons + bands for the current eps
    int ensure_raster()
    {
        if (raster_valid && cache) return RJ_OK;
        double t0 = now_ms();
        int rc = make_plan(polys.bbox, eps, ground_res, &plan);
        if (rc) return fail(rc, "canvas too large or bad accuracy for this polygon bbox");
        rasterize_polys(polys, plan, &pr);
        cur_setup += now_ms() - t0;
        t0 = now_ms();
        // regions and bands: one pinned staging buffer, one copy
        auto al = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t parts[3] = {pr.region.size() * 4, pr.bands_sum.size() * sizeof(Band),
                                 pr.bands_bounds.size() * sizeof(Band)};
        const void* srcs[3] = {pr.region.data(), pr.bands_sum.data(), pr.bands_bounds.data()};
        size_t offs[3], total = 0;
        for (int k = 0; k < 3; k++) {
            offs[k] = total;
            total += al(parts[k]);
        }
        if (total > h_stage_bytes) {
            if (h_stage) (void)hipHostFree(h_stage);
            h_stage = nullptr;
            hipError_t e = hipHostMalloc(&h_stage, total);
            if (e != hipSuccess) {
                h_stage = nullptr;
                h_stage_bytes = 0;
                return hip_fail(e, "pinned staging");
            }
            h_stage_bytes = total;
        }
        for (int k = 0; k < 3; k++)
            if (parts[k]) std::memcpy((char*)h_stage + offs[k], srcs[k], parts[k]);
        if ((rc = alloc(d_pc, total, "polygon regions"))) return rc;
        hipError_t e = hipMemcpyAsync(d_pc.p, h_stage, total, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return hip_fail(e, "polygon regions");
        char* base = d_pc.as<char>();
        p_region = (int32_t*)(base + offs[0]);
        p_bands_sum = (Band*)(base + offs[1]);
        p_bands_bounds = (Band*)(base + offs[2]);
        // snapped vertices on the device
        const int64_t nv = polys.off[polys.np];
        if ((rc = alloc(d_XY, (size_t)nv * 8, "snapped vertices"))) return rc;
        p_X = d_XY.as<int32_t>();
        p_Y = p_X + nv;
        if ((rc = launch("snap", [&] {
                 launch_snap(stream, d_pxy.as<float>(), nv, plan.minx, plan.miny, plan.fxs, plan.fys, p_X, p_Y);
             })))
            return rc;
        cur_polymem += now_ms() - t0;
        raster_valid = true;
        label_raster_valid = label_hybrid_valid = false;
        return RJ_OK;
    }

    int ensure_grid()
    {
        if (grid_valid && grid.xs == index_res && cache) return RJ_OK;
        double t0 = now_ms();
        build_grid(polys, index_res, &grid);
        cur_polyindex += now_ms() - t0;
        int rc;
        if ((rc = upload(d_cell_off, grid.cell_off.data(), grid.cell_off.size() * 4, "grid cells"))) return rc;
        if ((rc = upload(d_cand, grid.cand.data(), grid.cand.size() * 4, "grid candidates"))) return rc;
        grid_valid = true;
        label_hybrid_valid = false;
        return RJ_OK;
    }

    // d_res: [count | sum hi | sum lo] (3 np) [error bounds] (4 np) [aux] (4):
    // aux[0] = pixels covered summed over polygons, aux[1] = labeled pixels.
    unsigned long long* aux_ptr() const { return d_res.as<unsigned long long>() + 7 * (size_t)polys.np; }

    // Label map of the polygons on the canvas with plain stores. Returns
    // without waiting; the overlap test (aux[0] == aux[1]) is read later.
    // Label map (plain stores) and its tile map. Returns without waiting; the
    // overlap test aux[0] == aux[1] (coverage summed over polygons == labeled
    // pixels) is read later.
    int build_labels(void* lab, int label16, DevBuf& tiles)
    {
        const size_t npx = label_pixels();
        const size_t words = label16 ? npx / 2 : npx;
        DevCanvas cv = dev_canvas(plan);
        DevPolys P = dev_polys();
        unsigned long long* aux = aux_ptr();
        int rc;
        if ((rc = launch("fill_label", [&] {
                 launch_fill_u32(stream, static_cast<uint32_t*>(lab), words, 0xFFFFFFFFu);
                 launch_fill_u32(stream, (uint32_t*)aux, 8, 0);
             })))
            return rc;
        if ((rc = launch("band_label", [&] {
                 launch_band(stream, BAND_LABEL, p_bands_sum, (int)pr.bands_sum.size(), P, cv, nullptr, nullptr,
                             nullptr, DevResult{}, nullptr, lab, label16, aux, pr.lds_sum, 0);
             })))
            return rc;
        return build_tiles(lab, label16, tiles, aux + 1);
    }

    int read_overlap(bool* overlap)
    {
        unsigned long long a[2] = {0, 0};
        hipError_t e = hipMemcpyAsync(a, aux_ptr(), 16, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "overlap test");
        *overlap = a[0] != a[1];
        return RJ_OK;
    }

    int label16_ok() const { return polys.np < (int)0xFFF0 ? 1 : 0; }
    // label maps are stored with rows of LW = W rounded up to 16 pixels
    size_t label_pixels() const { return (size_t)dev_canvas(plan).LW * plan.H; }
    int tiles_w() const { return (plan.W + (1 << kTileShift) - 1) >> kTileShift; }
    size_t ntiles() const
    {
        return (size_t)tiles_w() * (size_t)((plan.H + (1 << kTileShift) - 1) >> kTileShift);
    }

    int build_tiles(const void* lab, int label16, DevBuf& tiles, unsigned long long* labeled)
    {
        int rc = alloc(tiles, ntiles() * 4, "label tiles");
        if (rc) return rc;
        return launch("tile_labels", [&] {
            launch_tile_labels(stream, lab, label16, dev_canvas(plan), tiles.as<uint32_t>(), labeled);
        });
    }

    int ensure_label_raster(bool* fresh)
    {
        *fresh = false;
        if (label_raster_valid && cache) return RJ_OK;
        const size_t npx = label_pixels();
        const int l16 = label16_ok();
        int rc = alloc(d_label_raster, npx * (l16 ? 2 : 4) + 16, "label map");
        if (rc) return rc;
        if ((rc = build_labels(d_label_raster.p, l16, d_tiles_raster))) return rc;
        label_raster_valid = true;
        label_raster_conflict = false;  // confirmed after the point pass
        *fresh = true;
        return RJ_OK;
    }

    int hybrid_radius() const
    {
        if (hybrid_dilation > 0) return hybrid_dilation;
        double m = std::max(std::max(std::fabs(plan.minx), std::fabs(plan.maxx)),
                            std::max(std::fabs(plan.miny), std::fabs(plan.maxy)));
        double err = 16.0 * m * std::ldexp(1.0, -24);
        double pw = std::min((plan.maxx - plan.minx) / plan.W, (plan.maxy - plan.miny) / plan.H);
        return std::max(1, (int)std::ceil(2.0 * err / pw));
    }

    int ensure_label_hybrid()
    {
        if (label_hybrid_valid && cache) return RJ_OK;
        const size_t npx = label_pixels();
        const size_t wpr = ((size_t)plan.W + 31) / 32;
        int rc = alloc(d_mask, wpr * plan.H * 4, "border mask");
        if (rc) return rc;
        std::vector<int32_t> rects;
        deficient_rects(grid, plan, &rects);
        if (!rects.empty() && (rc = upload(d_rects, rects.data(), rects.size() * 4, "deficient cells")))
            return rc;
        DevCanvas cv = dev_canvas(plan);
        DevPolys P = dev_polys();
        uint32_t* mask = d_mask.as<uint32_t>();
        label_hybrid16 = label16_ok();
        if ((rc = alloc(d_label_hybrid, npx * (label_hybrid16 ? 2 : 4) + 16, "hybrid label map"))) return rc;
        if ((rc = build_labels(d_label_hybrid.p, label_hybrid16, d_tiles_hybrid))) return rc;
        bool overlap = false;
        if ((rc = read_overlap(&overlap))) return rc;
        if (overlap) {
            // pixels in two or more polygons go to the exact test (32-bit map, CAS)
            label_hybrid16 = 0;
            if ((rc = alloc(d_label_hybrid, npx * 4 + 16, "hybrid label map"))) return rc;
            uint32_t* lab32 = d_label_hybrid.as<uint32_t>();
            if ((rc = launch("fill_label", [&] { launch_fill_u32(stream, lab32, npx, kLabelNone); }))) return rc;
            if ((rc = launch("band_label_atomic", [&] {
                     launch_band(stream, BAND_LABEL_ATOMIC, p_bands_sum, (int)pr.bands_sum.size(), P, cv, nullptr,
                                 nullptr, nullptr, DevResult{}, nullptr, lab32, 0, nullptr, pr.lds_sum, 0);
                 })))
                return rc;
        }
        void* lab = d_label_hybrid.p;
        if ((rc = launch("fill_mask", [&] { launch_fill_u32(stream, mask, wpr * plan.H, 0); }))) return rc;
        if ((rc = launch("touch", [&] {
                 launch_touch(stream, p_X, p_Y, d_prev.as<int32_t>(), (int64_t)polys.off[polys.np], cv, mask);
             })))
            return rc;
        if (!rects.empty())
            if ((rc = launch("mark_cells", [&] {
                     launch_mark_rects(stream, d_rects.as<int32_t>(), (int)(rects.size() / 4), cv, mask);
                 })))
                return rc;
        const int r = hybrid_radius();
        if (r > 31) return fail(RJ_ERR_UNSUPPORTED, "accuracy too fine for the float margin of the exact test");
        if ((rc = launch("dilate", [&] { launch_dilate_to_label(stream, mask, cv, r, lab, label_hybrid16); })))
            return rc;
        if ((rc = build_tiles(lab, label_hybrid16, d_tiles_hybrid, nullptr))) return rc;
        label_hybrid_valid = true;
        return RJ_OK;
    }

    int ensure_canvas(bool sums)
    {
        const size_t npx = (size_t)plan.W * plan.H;
        int rc = alloc(d_cnt, npx * 4, "canvas count");
        if (rc) return rc;
        if (sums) {
            if ((rc = alloc(d_chi, npx * 8, "canvas sum hi"))) return rc;
            if ((rc = alloc(d_clo, npx * 8, "canvas sum lo"))) return rc;
        }
        if ((rc = launch("fill_canvas", [&] {
                 launch_fill_u32(stream, d_cnt.as<uint32_t>(), npx, 0);
                 if (sums) {
                     launch_fill_u32(stream, d_chi.as<uint32_t>(), 2 * npx, 0);
                     launch_fill_u32(stream, d_clo.as<uint32_t>(), 2 * npx, 0);
                 }
             })))
            return rc;
        return RJ_OK;
    }

    int execute(int fn, int64_t* out_counts, double* out_sums, int32_t* legacy);
};

int rj_ctx::execute(int fn, int64_t* out_counts, double* out_sums, int32_t* legacy)
{
    if (fn < RJ_RASTER || fn > RJ_ERRORBOUNDS) return fail(RJ_ERR_UNSUPPORTED, "unknown function type");
    if (!has_polys) return fail(RJ_ERR_STATE, "no polygons set");
    if (!table) return fail(RJ_ERR_STATE, "no points set");
    const double t_start = now_ms();
    cur_ptmem = cur_polymem = cur_setup = cur_polyindex = 0;
    launches.clear();
    pool_used = 0;
    DevQuery q;
    int rc = build_query(&q);
    if (rc) return rc;
    if (fn == RJ_ERRORBOUNDS) {  // counts only (RasterJoinBounds.cpp:124)
        q.agg_type = RJ_AGG_COUNT;
        q.agg_slot = -1;
    }
    const bool avg = q.agg_type != RJ_AGG_COUNT;
    const int np = polys.np;
    if ((rc = alloc(d_res, (size_t)np * 7 * 8 + 32, "results"))) return rc;
    unsigned long long* res_base = d_res.as<unsigned long long>();
    DevResult res{res_base, (long long*)(res_base + np), res_base + 2 * (size_t)np};
    long long* bnd = (long long*)(res_base + 3 * (size_t)np);
    auto clear_results = [&] {
        return launch("fill_results", [&] { launch_fill_u32(stream, (uint32_t*)res_base, (size_t)np * 14, 0); });
    };
    if ((rc = clear_results())) return rc;

    bool check_overlap = false;
    PointArgs pa;
    std::memset(&pa, 0, sizeof(pa));
    pa.xy = static_cast<const float*>(table->xy);
    pa.begin = rb;
    pa.end = re;
    pa.cols = dev_cols();
    pa.q = q;
    pa.res = res;
    pa.lds_np = (np <= 2048) ? np : 0;
    auto canvas_pass = [&](bool sums) -> int {
        int r2 = ensure_canvas(sums);
        if (r2) return r2;
        PointArgs sp = pa;
        sp.cv = dev_canvas(plan);
        sp.cnt = d_cnt.as<uint32_t>();
        sp.chi = sums ? d_chi.as<unsigned long long>() : nullptr;
        sp.clo = sums ? d_clo.as<unsigned long long>() : nullptr;
        return launch("splat", [&] { launch_points(stream, PT_SPLAT, sp); });
    };

    if (fn == RJ_INDEX) {
        if ((rc = ensure_grid())) return rc;
        pa.g = dev_grid();
        pa.P = dev_polys();
        if ((rc = launch("points_index", [&] { launch_points(stream, PT_INDEX, pa); }))) return rc;
    } else if (fn == RJ_HYBRID) {
        if ((rc = ensure_raster())) return rc;
        if ((rc = ensure_grid())) return rc;
        if ((rc = ensure_label_hybrid())) return rc;
        pa.cv = dev_canvas(plan);
        pa.g = dev_grid();
        pa.P = dev_polys();
        pa.label = d_label_hybrid.p;
        pa.label16 = label_hybrid16;
        pa.tiles = d_tiles_hybrid.as<uint32_t>();
        pa.tiles_w = tiles_w();
        if ((rc = launch("points_hybrid", [&] { launch_points(stream, PT_LABEL_HYBRID, pa); }))) return rc;
    } else if (fn == RJ_RASTER) {
        if ((rc = ensure_raster())) return rc;
        const bool known_overlap = label_raster_valid && cache && label_raster_conflict;
        if (raster_path == 2 && known_overlap)
            return fail(RJ_ERR_UNSUPPORTED, "polygons overlap: the label-map path needs disjoint polygons");
        bool use_label = raster_path != 1 && !known_overlap;
        bool fresh = false;
        if (use_label) {
            if ((rc = ensure_label_raster(&fresh))) return rc;
            if (label_raster_conflict) {
                if (raster_path == 2)
                    return fail(RJ_ERR_UNSUPPORTED, "polygons overlap: the label-map path needs disjoint polygons");
                use_label = false;
            }
        }
        if (use_label) {
            pa.cv = dev_canvas(plan);
            pa.P = dev_polys();
            pa.label = d_label_raster.p;
            pa.label16 = label16_ok();
            pa.tiles = d_tiles_raster.as<uint32_t>();
            pa.tiles_w = tiles_w();
            if ((rc = launch("points_label", [&] { launch_points(stream, PT_LABEL, pa); }))) return rc;
            check_overlap = fresh;
        }
        if (!use_label) {
            if ((rc = canvas_pass(avg))) return rc;
            DevCanvas cv = dev_canvas(plan);
            if ((rc = launch("band_sum", [&] {
                     launch_band(stream, BAND_SUM, p_bands_sum, (int)pr.bands_sum.size(), dev_polys(),
                                 cv, d_cnt.as<uint32_t>(), d_chi.as<long long>(), d_clo.as<unsigned long long>(),
                                 res, nullptr, nullptr, 0, nullptr, pr.lds_sum, avg ? 1 : 0);
                 })))
                return rc;
        }
    } else {
        if ((rc = ensure_raster())) return rc;
        if (!pr.bounds_ok) return fail(RJ_ERR_UNSUPPORTED, "a polygon is too wide for the error-bounds kernel");
        DevCanvas cv = dev_canvas(plan);
        if ((rc = canvas_pass(false))) return rc;
        if ((rc = launch("band_bounds", [&] {
                 launch_band(stream, BAND_BOUNDS, p_bands_bounds, (int)pr.bands_bounds.size(),
                             dev_polys(), cv, d_cnt.as<uint32_t>(), nullptr, nullptr, res, bnd, nullptr, 0,
                             nullptr, pr.lds_bounds, 0);
             })))
            return rc;
    }

    // one pinned copy of [results | bounds | aux]
    const size_t res_bytes = ((size_t)np * 7 + 4) * 8;
    if (res_bytes > h_res_bytes) {
        if (h_res) (void)hipHostFree(h_res);
        h_res = nullptr;
        h_res_bytes = 0;
        hipError_t e = hipHostMalloc((void**)&h_res, res_bytes);
        if (e != hipSuccess) {
            h_res = nullptr;
            return hip_fail(e, "pinned results");
        }
        h_res_bytes = res_bytes;
    }
    hipError_t e = hipMemcpyAsync(h_res, res_base, res_bytes, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "execute");
    if (check_overlap && h_res[7 * (size_t)np] != h_res[7 * (size_t)np + 1]) {
        // the fresh label map has pixels in two polygons: redo on the canvas
        label_raster_conflict = true;
        if (raster_path == 2)
            return fail(RJ_ERR_UNSUPPORTED, "polygons overlap: the label-map path needs disjoint polygons");
        if ((rc = clear_results())) return rc;
        if ((rc = canvas_pass(avg))) return rc;
        DevCanvas cv = dev_canvas(plan);
        if ((rc = launch("band_sum", [&] {
                 launch_band(stream, BAND_SUM, p_bands_sum, (int)pr.bands_sum.size(), dev_polys(), cv,
                             d_cnt.as<uint32_t>(), d_chi.as<long long>(), d_clo.as<unsigned long long>(), res,
                             nullptr, nullptr, 0, nullptr, pr.lds_sum, avg ? 1 : 0);
             })))
            return rc;
        e = hipMemcpyAsync(h_res, res_base, res_bytes, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "execute");
    }
    counts.assign((const int64_t*)h_res, (const int64_t*)h_res + np);
    hi.assign((const int64_t*)h_res + np, (const int64_t*)h_res + 2 * np);
    lo.assign(h_res + 2 * np, h_res + 3 * np);
    bounds.assign((const int64_t*)h_res + 3 * np, (const int64_t*)h_res + 7 * np);
    last_fn = fn;

    for (int p = 0; p < np; p++) {
        if (out_counts) out_counts[p] = counts[p];
        const double s = avg ? fixed_to_double(hi[p], lo[p]) : 0.0;
        if (out_sums) out_sums[p] = s;
        if (legacy) {
            legacy[p] = (int32_t)counts[p];
            const double hundreds = std::trunc(s / 100.0);
            legacy[np + p] = (int32_t)hundreds;
            legacy[2 * np + p] = (int32_t)((s - 100.0 * hundreds) * 10.0);
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
    T.passes = (fn == RJ_INDEX || fn == RJ_HYBRID) ? index_res : 1;
    T.ncons = (int)cons.size();
    return RJ_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

#define RJ_GUARD(ctx)                      \
    if (!(ctx)) return RJ_ERR_ARG;         \
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
    // GLHandler keeps 100 MB of the budget for other data (GLHandler.cpp:128-132).
    c->budget = gpu_mem_mb > 0 ? (gpu_mem_mb - 100) * 1024LL * 1024LL : 0;
    if (gpu_mem_mb > 0 && c->budget <= 0) c->budget = 1;
    c->inmem = inmem;
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
    if (np <= 0 || !xy || !off) return ctx->fail(RJ_ERR_ARG, "empty polygon collection");
    if (off[0] != 0) return ctx->fail(RJ_ERR_ARG, "offsets must start at 0");
    for (int p = 0; p < np; p++)
        if (off[p + 1] - off[p] < 3) return ctx->fail(RJ_ERR_ARG, "polygon " + std::to_string(p) + " has < 3 vertices");
    PolySet ps;
    ps.np = np;
    ps.off.assign(off, off + np + 1);
    ps.xy.assign(xy, xy + 2 * (size_t)off[np]);
    float mn[2] = {xy[0], xy[1]}, mx[2] = {xy[0], xy[1]};
    for (int64_t k = 0; k < off[np]; k++) {
        for (int d = 0; d < 2; d++) {
            float v = xy[2 * k + d];
            if (!std::isfinite(v)) return ctx->fail(RJ_ERR_ARG, "non-finite polygon vertex");
            mn[d] = std::min(mn[d], v);
            mx[d] = std::max(mx[d], v);
        }
    }
    if (bbox) {
        for (int d = 0; d < 4; d++) ps.bbox[d] = bbox[d];
    } else {
        ps.bbox[0] = mn[0]; ps.bbox[1] = mn[1]; ps.bbox[2] = mx[0]; ps.bbox[3] = mx[1];
    }
    for (int d = 0; d < 4; d++) ps.bboxf[d] = (float)ps.bbox[d];
    if (!(ps.bbox[2] > ps.bbox[0]) || !(ps.bbox[3] > ps.bbox[1])) return ctx->fail(RJ_ERR_ARG, "degenerate bbox");
    double t0 = now_ms();
    ps.derive();
    ctx->polys = std::move(ps);
    int rc = ctx->upload(ctx->d_pxy, ctx->polys.xy.data(), ctx->polys.xy.size() * 4, "polygons");
    if (!rc) rc = ctx->upload(ctx->d_poff, ctx->polys.off.data(), ctx->polys.off.size() * 4, "polygon offsets");
    if (!rc) rc = ctx->upload(ctx->d_prev, ctx->polys.prev.data(), ctx->polys.prev.size() * 4, "edge list");
    if (!rc) {
        hipError_t e = hipStreamSynchronize(ctx->stream);  // the host vectors may move later
        if (e != hipSuccess) rc = ctx->hip_fail(e, "polygon upload");
    }
    if (rc) return rc;
    ctx->cur_polymem = now_ms() - t0;
    ctx->has_polys = true;
    ctx->invalidate_polys();
    return RJ_OK;
}

static int table_fill(rj_ctx* ctx, rj_table* t, int64_t n, const float* xy, int32_t ncols, const void* const* cols,
                      const int32_t* types)
{
    t->n = n;
    t->ncols = ncols;
    hipError_t e = hipMalloc(&t->xy, std::max<size_t>(8, (size_t)n * 8));
    if (e != hipSuccess) return ctx->hip_fail(e, "point table");
    if (n) e = hipMemcpyAsync(t->xy, xy, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream);
    for (int c = 0; c < ncols && e == hipSuccess; c++) {
        t->types[c] = types[c];
        size_t b = (size_t)n * type_bytes(types[c]);
        e = hipMalloc(&t->cols[c], std::max<size_t>(8, b));
        if (e == hipSuccess && n) e = hipMemcpyAsync(t->cols[c], cols[c], b, hipMemcpyHostToDevice, ctx->stream);
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "point table");
    float mn[2] = {0, 0}, mx[2] = {1, 1};
    bool first = true;
    for (int64_t i = 0; i < n; i++) {
        float x = xy[2 * i], y = xy[2 * i + 1];
        if (!std::isfinite(x) || !std::isfinite(y)) continue;
        if (first) {
            mn[0] = mx[0] = x; mn[1] = mx[1] = y;
            first = false;
        }
        mn[0] = std::min(mn[0], x); mx[0] = std::max(mx[0], x);
        mn[1] = std::min(mn[1], y); mx[1] = std::max(mx[1], y);
    }
    t->bbox[0] = mn[0]; t->bbox[1] = mn[1]; t->bbox[2] = mx[0]; t->bbox[3] = mx[1];
    return RJ_OK;
}

static int table_organize(rj_ctx* ctx, rj_table* t)
{
    // the radix sort counts items in an int: larger tables keep their order
    if (t->n < 2 || t->n >= ((int64_t)1 << 31)) return RJ_OK;
    void* scratch = nullptr;
    hipError_t e = hipMalloc(&scratch, organize_scratch_bytes(t->n));
    if (e != hipSuccess) return ctx->hip_fail(e, "organize scratch");
    int r = organize_rows(ctx->stream, t->n, (float*)t->xy, t->cols, t->types, t->ncols, t->bbox, scratch);
    e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(scratch);
    if (r != 0) return ctx->hip_fail((hipError_t)r, "organize");
    if (e != hipSuccess) return ctx->hip_fail(e, "organize");
    t->organized = true;
    return RJ_OK;
}

static int check_columns(rj_ctx* ctx, int64_t n, const float* xy, int32_t ncols, const void* const* cols,
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

int rj_set_points(rj_ctx* ctx, int64_t n, const float* xy, int32_t ncols, const void* const* cols, const int32_t* types)
{
    RJ_GUARD(ctx);
    int rc = check_columns(ctx, n, xy, ncols, cols, types);
    if (rc) return rc;
    size_t bytes = (size_t)n * 8;
    for (int c = 0; c < ncols; c++) bytes += (size_t)n * type_bytes(types[c]);
    if (ctx->budget > 0 && ctx->used + (int64_t)bytes > ctx->budget)
        return ctx->fail(RJ_ERR_NOMEM, "points do not fit the memory budget (" + std::to_string(bytes) + " bytes)");
    double t0 = now_ms();
    if (ctx->own_table) delete ctx->table;
    ctx->table = nullptr;
    ctx->own_table = false;
    rj_table* t = new rj_table();
    (void)hipGetDevice(&t->device);
    rc = table_fill(ctx, t, n, xy, ncols, cols, types);
    if (!rc && ctx->organize) rc = table_organize(ctx, t);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = ctx->hip_fail(e, "point upload");
    }
    if (rc) {
        delete t;
        return rc;
    }
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
    int rc = check_columns(ctx, n, xy, ncols, cols, types);
    if (rc) return rc;
    rj_table* t = new rj_table();
    (void)hipGetDevice(&t->device);
    rc = table_fill(ctx, t, n, xy, ncols, cols, types);
    if (!rc && organize) rc = table_organize(ctx, t);
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

int rj_table_rows(const rj_table* t, int64_t* n)
{
    if (!t || !n) return RJ_ERR_ARG;
    *n = t->n;
    return RJ_OK;
}

int rj_table_read(const rj_table* t, int64_t row0, int64_t n, float* xy, void* const* cols)
{
    if (!t || row0 < 0 || n < 0 || row0 + n > t->n) return RJ_ERR_ARG;
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
    if (eps_m != ctx->eps || ground_res != ctx->ground_res) {
        ctx->eps = eps_m;
        ctx->ground_res = ground_res;
        ctx->raster_valid = ctx->label_raster_valid = ctx->label_hybrid_valid = false;
    }
    return RJ_OK;
}

int rj_set_index_res(rj_ctx* ctx, int32_t rx, int32_t ry)
{
    RJ_GUARD(ctx);
    (void)ry;
    if (rx <= 0 || rx > 16384) return ctx->fail(RJ_ERR_ARG, "index resolution must be in 1..16384");
    if (rx != ctx->index_res) {
        ctx->index_res = rx;
        ctx->grid_valid = ctx->label_hybrid_valid = false;
    }
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
    case RJ_OPT_TIMING: ctx->timing = value != 0; return RJ_OK;
    case RJ_OPT_HYBRID_DILATION:
        if (value < 0 || value > 31) return ctx->fail(RJ_ERR_ARG, "dilation radius is 0 (auto) .. 31");
        ctx->hybrid_dilation = (int)value;
        ctx->label_hybrid_valid = false;
        return RJ_OK;
    case RJ_OPT_CACHE: ctx->cache = value != 0; return RJ_OK;
    case RJ_OPT_ORGANIZE: ctx->organize = value != 0; return RJ_OK;
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
    if (hi) std::copy(ctx->hi.begin(), ctx->hi.end(), hi);
    if (lo) std::copy(ctx->lo.begin(), ctx->lo.end(), lo);
    return RJ_OK;
}

int rj_error_bounds(rj_ctx* ctx, int64_t* out4np)
{
    RJ_GUARD(ctx);
    if (ctx->last_fn != RJ_ERRORBOUNDS) return ctx->fail(RJ_ERR_STATE, "last execute was not RJ_ERRORBOUNDS");
    if (out4np) std::copy(ctx->bounds.begin(), ctx->bounds.end(), out4np);
    return RJ_OK;
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

int rj_canvas_info(rj_ctx* ctx, int32_t* W, int32_t* H)
{
    RJ_GUARD(ctx);
    if (!ctx->has_polys) return ctx->fail(RJ_ERR_STATE, "no polygons set");
    CanvasPlan p;
    int rc = make_plan(ctx->polys.bbox, ctx->eps, ctx->ground_res, &p);
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

int rj_sync(rj_ctx* ctx)
{
    RJ_GUARD(ctx);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? RJ_OK : ctx->hip_fail(e, "sync");
}

}  // extern "C"
