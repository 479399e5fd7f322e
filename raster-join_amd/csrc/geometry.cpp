// geometry.cpp -- host preprocessing of the polygon side of the join:
// canvas plan, fixed-point snapping and scan-conversion bands, and the pixel
// rectangles of the grid cells whose candidate lists are incomplete.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <cstring>

#include "rj_impl.hpp"

namespace rj {

// ---------------------------------------------------------------------------
// canvas plan
// ---------------------------------------------------------------------------

int make_plan(const double bbox[4], double eps_m, double ground_res, CanvasPlan* p)
{
    *p = CanvasPlan();
    const double diffx = bbox[2] - bbox[0], diffy = bbox[3] - bbox[1];
    if (!(diffx > 0) || !(diffy > 0) || !(eps_m > 0) || !(ground_res > 0)) return RJ_ERR_ARG;
    // GLFunction::setMaxCellSize: the pixel diagonal is eps.
    const double cell = (eps_m / std::sqrt(2.0)) / ground_res;
    const double ax = std::ceil(diffx / cell), ay = std::ceil(diffy / cell);
    if (ax > 1e9 || ay > 1e9) return RJ_ERR_LIMIT;
    // RasterJoin::updateBuffers: tiles of at most 8192 pixels per side; the
    // effective resolution is tiles x tile size.
    const int actx = (int)ax, acty = (int)ay;
    const int splitx = (int)std::ceil((double)actx / 8192.0);
    const int splity = (int)std::ceil((double)acty / 8192.0);
    const int resx = (int)std::ceil((double)actx / splitx);
    const int resy = (int)std::ceil((double)acty / splity);
    p->W = splitx * resx;
    p->H = splity * resy;
    if ((int64_t)p->W * kSub >= ((int64_t)1 << 31) || (int64_t)p->H * kSub >= ((int64_t)1 << 31) ||
        (int64_t)p->W * p->H >= ((int64_t)1 << 31))
        return RJ_ERR_LIMIT;
    p->splitx = splitx;
    p->splity = splity;
    p->minx = bbox[0]; p->miny = bbox[1]; p->maxx = bbox[2]; p->maxy = bbox[3];
    p->minx_f = (float)bbox[0]; p->miny_f = (float)bbox[1];
    p->maxx_f = (float)bbox[2]; p->maxy_f = (float)bbox[3];
    p->sx_f = (float)((double)p->W / diffx);
    p->sy_f = (float)((double)p->H / diffy);
    p->fxs = (double)p->W * 256.0 / diffx;
    p->fys = (double)p->H * 256.0 / diffy;
    p->cell = cell;
    return RJ_OK;
}

DevCanvas dev_canvas(const CanvasPlan& p)
{
    DevCanvas c;
    c.W = p.W; c.H = p.H;
    c.LW = (p.W + kLabelRowAlign - 1) & ~(kLabelRowAlign - 1);
    c.minx = p.minx_f; c.miny = p.miny_f; c.maxx = p.maxx_f; c.maxy = p.maxy_f;
    c.sx = p.sx_f; c.sy = p.sy_f;
    return c;
}

// ---------------------------------------------------------------------------
// polygons on the canvas
// ---------------------------------------------------------------------------

static int64_t floor_div(int64_t a, int64_t b)
{
    int64_t q = a / b;
    if ((a % b != 0) && (a < 0)) q--;
    return q;
}

bool PolySet::derive(const float* xy)
{
    pbox.assign(4 * (size_t)np, 0.0f);
    const int64_t nv = off.empty() ? 0 : off[np];
    std::atomic<bool> finite{true};
    auto run = [&](int pa, int pb) {
        bool ok = true;
        for (int p = pa; p < pb; p++) {
            const int v0 = off[p], v1 = off[p + 1];
            float mnx = xy[2 * v0], mxx = mnx, mny = xy[2 * v0 + 1], mxy = mny;
            for (int k = v0; k < v1; k++) {
                const float x = xy[2 * k], y = xy[2 * k + 1];
                ok = ok && std::isfinite(x) && std::isfinite(y);
                mnx = std::min(mnx, x);
                mxx = std::max(mxx, x);
                mny = std::min(mny, y);
                mxy = std::max(mxy, y);
            }
            pbox[4 * p] = mnx; pbox[4 * p + 1] = mny; pbox[4 * p + 2] = mxx; pbox[4 * p + 3] = mxy;
        }
        if (!ok) finite = false;
    };
    // large collections on a few threads, split by vertex count
    const int nth = nv >= (1 << 18) ? 8 : 1;
    if (nth == 1) {
        run(0, np);
    } else {
        std::vector<std::thread> th;
        int pa = 0;
        for (int t = 0; t < nth; t++) {
            int pb = pa;
            const int64_t goal = nv * (t + 1) / nth;
            while (pb < np && (t == nth - 1 || off[pb] < goal)) pb++;
            th.emplace_back(run, pa, pb);
            pa = pb;
        }
        for (auto& x : th) x.join();
    }
    return finite;
}

int64_t snap_x(const CanvasPlan& plan, float x) { return std::llrint(((double)x - plan.minx) * plan.fxs); }
int64_t snap_y(const CanvasPlan& plan, float y) { return std::llrint(((double)y - plan.miny) * plan.fys); }

void rasterize_polys(const PolySet& ps, const CanvasPlan& plan, PolyRaster* out)
{
    out->region.assign(4 * (size_t)ps.np, 0);
    out->bands_sum.clear();
    out->bands_bounds.clear();
    out->lds_sum = out->lds_bounds = out->lds_bounds_small = 0;
    out->nbounds_small = 0;
    out->bounds_ok = true;
    const int sum_words = 12288;    // LDS words per band, sum / label kernels
    for (int p = 0; p < ps.np; p++) {
        const int64_t mnx = snap_x(plan, ps.pbox[4 * p]), mny = snap_y(plan, ps.pbox[4 * p + 1]);
        const int64_t mxx = snap_x(plan, ps.pbox[4 * p + 2]), mxy = snap_y(plan, ps.pbox[4 * p + 3]);
        // Cells whose center can be inside, plus one ring for the dilation of
        // the error bounds.
        int64_t c0 = std::max<int64_t>(0, floor_div(mnx, kSub) - 2);
        int64_t c1 = std::min<int64_t>(plan.W - 1, floor_div(mxx, kSub) + 2);
        int64_t r0 = std::max<int64_t>(0, floor_div(mny, kSub) - 2);
        int64_t r1 = std::min<int64_t>(plan.H - 1, floor_div(mxy, kSub) + 2);
        int32_t* reg = &out->region[4 * (size_t)p];
        reg[0] = (int32_t)c0; reg[1] = (int32_t)c1; reg[2] = (int32_t)r0; reg[3] = (int32_t)r1;
        if (c1 < c0 || r1 < r0) continue;
        const int bw = (int)(c1 - c0 + 1);
        // sum / label bands: column tiles of at most 4096 pixels, rows to fill LDS
        const int tile_w = 4096;
        for (int64_t cs = c0; cs <= c1; cs += tile_w) {
            const int64_t ce = std::min<int64_t>(c1, cs + tile_w - 1);
            const int wpr = (int)((ce - cs + 1 + 31) / 32);
            const int rows = std::max(1, std::min(64, sum_words / wpr));
            for (int64_t r = r0; r <= r1; r += rows) {
                Band b;
                b.poly = p;
                b.r0 = (int32_t)r;
                b.nrows = (int32_t)std::min<int64_t>(rows, r1 - r + 1);
                b.wpr = wpr;
                b.c0 = (int32_t)cs;
                b.c1 = (int32_t)ce;
                out->bands_sum.push_back(b);
                out->lds_sum = std::max(out->lds_sum, band_lds_bytes(BAND_SUM, b.nrows, wpr));
            }
        }
    }
}

void bounds_bands(const PolySet& ps, PolyRaster* out)
{
    out->bands_bounds.clear();
    out->lds_bounds = out->lds_bounds_small = 0;
    out->nbounds_small = 0;
    out->bounds_ok = true;
    // LDS words per band: bands aim at kBoundsSmallWords, so that eight
    // workgroups share a CU (at most 64 rows; 1.19 -> 0.53 ms for 260
    // polygons at 7072^2, although a wide region then takes bands of few
    // rows, each walking all of its polygon's edges). A region wider than
    // one row of that takes up to 16384 words, in a launch of its own.
    const int bound_words = 16384;
    std::vector<Band> wide;
    for (int p = 0; p < ps.np; p++) {
        const int32_t* reg = &out->region[4 * (size_t)p];
        const int64_t c0 = reg[0], c1 = reg[1], r0 = reg[2], r1 = reg[3];
        if (c1 < c0 || r1 < r0) continue;
        const int bw = (int)(c1 - c0 + 1);
        const int wpr = (bw + 31) / 32;
        const int fit = (bound_words - 2 * wpr - kBandScan) / (34 * wpr);
        const int rows = std::min({64, fit, std::max(1, (kBoundsSmallWords - 2 * wpr - kBandScan) / (34 * wpr))});
        if (rows < 1) {
            out->bounds_ok = false;
            continue;
        }
        for (int64_t r = r0; r <= r1; r += rows) {
            Band b;
            b.poly = p;
            b.r0 = (int32_t)r;
            b.nrows = (int32_t)std::min<int64_t>(rows, r1 - r + 1);
            b.wpr = wpr;
            b.c0 = (int32_t)c0;
            b.c1 = (int32_t)c1;
            const size_t lds = band_lds_bytes(BAND_BOUNDS, b.nrows, wpr);
            if (lds <= (size_t)kBoundsSmallWords * 4) {
                out->bands_bounds.push_back(b);
                out->lds_bounds_small = std::max(out->lds_bounds_small, lds);
            } else {
                wide.push_back(b);
                out->lds_bounds = std::max(out->lds_bounds, lds);
            }
        }
    }
    out->nbounds_small = (int)out->bands_bounds.size();
    out->bands_bounds.insert(out->bands_bounds.end(), wide.begin(), wide.end());
}

// ---------------------------------------------------------------------------
// CPUJoin grid index (built on the device, gridbuild.hip)
// ---------------------------------------------------------------------------

void deficient_rects(const GridIndex& g, const CanvasPlan& plan, std::vector<int32_t>* rects)
{
    rects->clear();
    const double sx = plan.W / (plan.maxx - plan.minx), sy = plan.H / (plan.maxy - plan.miny);
    for (int32_t cell : g.deficient) {
        const int x = cell % g.xs, y = cell / g.xs;
        const double x0 = (double)((float)x * g.width + g.minx), x1 = (double)((float)(x + 1) * g.width + g.minx);
        const double y0 = (double)((float)y * g.height + g.miny), y1 = (double)((float)(y + 1) * g.height + g.miny);
        rects->push_back((int32_t)std::floor((x0 - plan.minx) * sx) - 1);
        rects->push_back((int32_t)std::floor((x1 - plan.minx) * sx) + 1);
        rects->push_back((int32_t)std::floor((y0 - plan.miny) * sy) - 1);
        rects->push_back((int32_t)std::floor((y1 - plan.miny) * sy) + 1);
    }
}

}  // namespace rj
