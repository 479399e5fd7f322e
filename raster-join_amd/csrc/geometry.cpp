// geometry.cpp -- host preprocessing of the polygon side of the join:
// canvas plan, fixed-point snapping and scan-conversion bands, and the CPUJoin
// grid index with its cell-candidate test.
//
// Compiled with -ffp-contract=off: the grid index must reproduce the float
// arithmetic of CPUJoin/GridIndexF.cpp exactly (the candidate lists decide
// which polygons a point is tested against, so they are part of the result).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

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

void PolySet::derive()
{
    pbox.assign(4 * (size_t)np, 0.0f);
    prev.resize(off.empty() ? 0 : off[np]);
    for (int p = 0; p < np; p++) {
        const int v0 = off[p], v1 = off[p + 1];
        float mnx = xy[2 * v0], mxx = mnx, mny = xy[2 * v0 + 1], mxy = mny;
        for (int k = v0; k < v1; k++) {
            prev[k] = (k == v0) ? v1 - 1 : k - 1;
            mnx = std::min(mnx, xy[2 * k]);
            mxx = std::max(mxx, xy[2 * k]);
            mny = std::min(mny, xy[2 * k + 1]);
            mxy = std::max(mxy, xy[2 * k + 1]);
        }
        pbox[4 * p] = mnx; pbox[4 * p + 1] = mny; pbox[4 * p + 2] = mxx; pbox[4 * p + 3] = mxy;
    }
}

int64_t snap_x(const CanvasPlan& plan, float x) { return std::llrint(((double)x - plan.minx) * plan.fxs); }
int64_t snap_y(const CanvasPlan& plan, float y) { return std::llrint(((double)y - plan.miny) * plan.fys); }

void rasterize_polys(const PolySet& ps, const CanvasPlan& plan, PolyRaster* out)
{
    out->region.assign(4 * (size_t)ps.np, 0);
    out->bands_sum.clear();
    out->bands_bounds.clear();
    out->lds_sum = out->lds_bounds = 0;
    out->bounds_ok = true;
    const int sum_words = 12288;    // LDS words per band, sum / label kernels
    const int bound_words = 16384;  // LDS words per band, error-bounds kernel
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
        // error-bounds bands span the full region width
        {
            const int wpr = (bw + 31) / 32;
            const int rows = std::min(64, (bound_words - 2 * wpr) / (18 * wpr));
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
                out->bands_bounds.push_back(b);
                out->lds_bounds = std::max(out->lds_bounds, band_lds_bytes(BAND_BOUNDS, b.nrows, wpr));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// CPUJoin grid index
// ---------------------------------------------------------------------------
namespace {

// Crossing number, GridIndexF::isInsidePoly (GridIndexF.cpp:383-402).
bool pip(const float* v, int n, float tx, float ty)
{
    bool c = false;
    for (int i = 0, j = n - 1; i < n; j = i++) {
        const float vix = v[2 * i], viy = v[2 * i + 1], vjx = v[2 * j], vjy = v[2 * j + 1];
        if (((viy > ty) != (vjy > ty)) && (tx < (vjx - vix) * (ty - viy) / (vjy - viy) + vix)) c = !c;
    }
    return c;
}

// Cohen-Sutherland region code, GridIndexF::ComputeOutCode (:499-511).
int region_code(float x, float y, float lbx, float lby, float rtx, float rty)
{
    int code = 0;
    if (x < lbx) code |= 1;
    else if (x > rtx) code |= 2;
    if (y < lby) code |= 4;
    else if (y > rty) code |= 8;
    return code;
}

// Ray from o along d against segment p0-p1, GridIndexF::rayToLineSegmentIntersection
// (:513-549): parallel when the two slopes compare equal, hit when r >= 0
// and 0 <= s <= 1; the hit point is o + d*r.
bool ray_hits(float ox, float oy, float dx, float dy, float p0x, float p0y, float p1x, float p1y,
              float* hx, float* hy)
{
    const float x1 = ox, y1 = oy, x2 = ox + dx, y2 = oy + dy;
    if ((y2 - y1) / (x2 - x1) == (p1y - p0y) / (p1x - p0x)) return false;
    const float d = ((x2 - x1) * (p1y - p0y)) - (y2 - y1) * (p1x - p0x);
    if (d == 0) return false;
    const float r = (((y1 - p0y) * (p1x - p0x)) - (x1 - p0x) * (p1y - p0y)) / d;
    const float s = (((y1 - p0y) * (x2 - x1)) - (x1 - p0x) * (y2 - y1)) / d;
    if (!(r >= 0) || !(s >= 0 && s <= 1)) return false;
    *hx = ox + dx * r;
    *hy = oy + dy * r;
    return true;
}

// GridIndexF::lineRectIntersection (:404-424). The second test compares the
// hit's y with the rectangle's right x exactly as the reference does (:412).
bool segment_meets_cell(float ax, float ay, float bx, float by, float lbx, float lby, float rtx, float rty)
{
    float hx, hy;
    if (ray_hits(lbx, lby, 0.0f, 1.0f, ax, ay, bx, by, &hx, &hy) && hy <= rty) return true;
    if (ray_hits(lbx, lby, 1.0f, 0.0f, ax, ay, bx, by, &hx, &hy) && hy <= rtx) return true;
    if (ray_hits(rtx, rty, 0.0f, -1.0f, ax, ay, bx, by, &hx, &hy) && hy >= lby) return true;
    if (ray_hits(rtx, rty, -1.0f, 0.0f, ax, ay, bx, by, &hx, &hy) && hx >= lbx) return true;
    return false;
}

// GridIndexF::polygonRectIntersectionCPU (:461-496): a vertex in the cell, a
// cell corner in the polygon, or one of the edges (i, i+1), i < n-1, meeting
// the cell (the closing edge is not tested).
bool cell_candidate(const float* v, int n, float lbx, float lby, float rtx, float rty)
{
    for (int i = 0; i < n; i++) {
        const float x = v[2 * i], y = v[2 * i + 1];
        if (x >= lbx && x <= rtx && y >= lby && y <= rty) return true;
    }
    if (pip(v, n, lbx, lby) || pip(v, n, rtx, rty) || pip(v, n, lbx, rty) || pip(v, n, rtx, lby))
        return true;
    int code_b = region_code(v[0], v[1], lbx, lby, rtx, rty);
    for (int i = 0; i + 1 < n; i++) {
        const int code_a = code_b;
        code_b = region_code(v[2 * i + 2], v[2 * i + 3], lbx, lby, rtx, rty);
        if (!(code_a & code_b) &&
            segment_meets_cell(v[2 * i], v[2 * i + 1], v[2 * i + 2], v[2 * i + 3], lbx, lby, rtx, rty))
            return true;
    }
    return false;
}

// Conservative double-precision test: does the closed polygon meet the
// closed rectangle (expanded by a relative margin)?
bool polygon_meets_rect(const float* v, int n, double lbx, double lby, double rtx, double rty)
{
    const double mx = 1e-6 * (std::fabs(rtx) + std::fabs(lbx) + (rtx - lbx));
    const double my = 1e-6 * (std::fabs(rty) + std::fabs(lby) + (rty - lby));
    lbx -= mx; rtx += mx; lby -= my; rty += my;
    for (int i = 0; i < n; i++) {
        const double x = v[2 * i], y = v[2 * i + 1];
        if (x >= lbx && x <= rtx && y >= lby && y <= rty) return true;
    }
    // rectangle center inside (the rectangle may lie in the polygon's interior)
    {
        const double cx = 0.5 * (lbx + rtx), cy = 0.5 * (lby + rty);
        bool c = false;
        for (int i = 0, j = n - 1; i < n; j = i++) {
            const double xi = v[2 * i], yi = v[2 * i + 1], xj = v[2 * j], yj = v[2 * j + 1];
            if (((yi > cy) != (yj > cy)) && (cx < (xj - xi) * (cy - yi) / (yj - yi) + xi)) c = !c;
        }
        if (c) return true;
    }
    // any edge (including the closing one) crossing the rectangle: Liang-Barsky clip
    for (int i = 0, j = n - 1; i < n; j = i++) {
        const double x0 = v[2 * j], y0 = v[2 * j + 1], dx = v[2 * i] - x0, dy = v[2 * i + 1] - y0;
        double t0 = 0, t1 = 1;
        const double pp[4] = {-dx, dx, -dy, dy};
        const double qq[4] = {x0 - lbx, rtx - x0, y0 - lby, rty - y0};
        bool ok = true;
        for (int k = 0; k < 4 && ok; k++) {
            if (pp[k] == 0) {
                if (qq[k] < 0) ok = false;
            } else {
                const double t = qq[k] / pp[k];
                if (pp[k] < 0) t0 = std::max(t0, t);
                else t1 = std::min(t1, t);
                if (t0 > t1) ok = false;
            }
        }
        if (ok) return true;
    }
    return false;
}

template <class F>
void parallel_polys(int np, F f)
{
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (np < 64) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) th.emplace_back([=] { f((int)t, (int)nt); });
    for (auto& x : th) x.join();
}

}  // namespace

// Grid row of a y coordinate, as the point pass computes it (getYIndex,
// GridIndexF.cpp:44-47, with the in-grid test of getRegion): -1 below the
// grid, ys above it. Monotone in t.
// Edge bands of every polygon (see GridIndex). Edge (vi, vj) toggles the
// crossing test of a point at ty iff min(vi.y, vj.y) <= ty < max(vi.y, vj.y)
// ((vi.y > ty) != (vj.y > ty)); the bands of those ty run from band_row(min)
// to band_row(the largest float below max), as band_row is monotone. The
// bands are the grid rows split k ways, k (a power of two up to 16) chosen so
// that a band holds about 8 edges of a polygon on average.
static void bands_for(const PolySet& ps, float miny, float band_h, int rows, GridIndex* g)
{
    const int np = ps.np;
    g->band_h = band_h;
    g->band_rows = rows;
    g->band_lo.assign(np, 0);
    g->band_hi.assign(np, -1);
    g->pband.assign(np + 1, 0);
    std::vector<std::vector<int32_t>> lo_hi(np);  // per polygon: (band0, band1) per edge
    parallel_polys(np, [&](int t, int nt) {
        for (int p = t; p < np; p += nt) {
            const float* v = ps.xy.data() + 2 * (size_t)ps.off[p];
            const int n = ps.off[p + 1] - ps.off[p];
            std::vector<int32_t>& r = lo_hi[p];
            r.resize(2 * (size_t)n);
            int blo = rows, bhi = -1;
            for (int i = 0; i < n; i++) {
                const int j = i ? i - 1 : n - 1;
                const float y0 = std::min(v[2 * i + 1], v[2 * j + 1]);
                const float y1 = std::max(v[2 * i + 1], v[2 * j + 1]);
                int r0 = 0, r1 = -1;
                if (y0 < y1) {
                    r0 = band_row(y0, miny, band_h, rows);
                    r1 = band_row(std::nextafter(y1, -INFINITY), miny, band_h, rows);
                }
                r[2 * i] = r0;
                r[2 * i + 1] = r1;
                if (r0 <= r1) {
                    blo = std::min(blo, r0);
                    bhi = std::max(bhi, r1);
                }
            }
            g->band_lo[p] = bhi >= blo ? blo : 0;
            g->band_hi[p] = bhi >= blo ? bhi : -1;
        }
    });
    for (int p = 0; p < np; p++) g->pband[p + 1] = g->pband[p] + (g->band_hi[p] - g->band_lo[p] + 1);
    const size_t nbands = (size_t)g->pband[np];
    g->band_off.assign(nbands + 1, 0);
    for (int p = 0; p < np; p++) {
        const int n = ps.off[p + 1] - ps.off[p];
        for (int i = 0; i < n; i++)
            for (int y = lo_hi[p][2 * i]; y <= lo_hi[p][2 * i + 1]; y++)
                g->band_off[(size_t)g->pband[p] + (y - g->band_lo[p]) + 1]++;
    }
    for (size_t b = 0; b < nbands; b++) g->band_off[b + 1] += g->band_off[b];
    g->edges.assign(4 * (size_t)g->band_off[nbands], 0.0f);
    std::vector<int32_t> fill(g->band_off.begin(), g->band_off.end() - 1);
    for (int p = 0; p < np; p++) {
        const float* v = ps.xy.data() + 2 * (size_t)ps.off[p];
        const int n = ps.off[p + 1] - ps.off[p];
        for (int i = 0; i < n; i++) {
            const int j = i ? i - 1 : n - 1;
            for (int y = lo_hi[p][2 * i]; y <= lo_hi[p][2 * i + 1]; y++) {
                float* e = &g->edges[4 * (size_t)fill[(size_t)g->pband[p] + (y - g->band_lo[p])]++];
                e[0] = v[2 * i]; e[1] = v[2 * i + 1];   // vi
                e[2] = v[2 * j]; e[3] = v[2 * j + 1];   // vj (the previous vertex)
            }
        }
    }
}

static void build_bands(const PolySet& ps, GridIndex* g)
{
    bands_for(ps, g->miny, g->height, g->ys, g);
    const size_t nb = std::max<size_t>(1, g->band_off.size() - 1);
    const double per = (double)g->band_off.back() / (double)nb;  // edges per band at the grid's rows
    int k = 1;
    while (k < 64 && per / k > 4.0) k *= 2;
    if (k > 1) bands_for(ps, g->miny, g->height / (float)k, g->ys * k, g);
}

void build_grid(const PolySet& ps, int res, GridIndex* g)
{
    g->xs = g->ys = res;
    g->minx = ps.bboxf[0];
    g->miny = ps.bboxf[1];
    g->width = (ps.bboxf[2] - ps.bboxf[0]) / (float)res;   // GridIndexF.cpp:66-67
    g->height = (ps.bboxf[3] - ps.bboxf[1]) / (float)res;
    const float minx = g->minx, miny = g->miny, width = g->width, height = g->height;
    const int xs = res, ys = res;

    struct Pair { int32_t cell, poly; };
    std::vector<std::vector<Pair>> acc(64);
    std::vector<std::vector<int32_t>> defi(64);
    parallel_polys(ps.np, [&](int t, int nt) {
        for (int p = t; p < ps.np; p += nt) {
            const float* v = ps.xy.data() + 2 * (size_t)ps.off[p];
            const int n = ps.off[p + 1] - ps.off[p];
            float bminx = v[0], bmaxx = v[0], bminy = v[1], bmaxy = v[1];
            for (int k = 1; k < n; k++) {
                bminx = std::min(bminx, v[2 * k]);
                bmaxx = std::max(bmaxx, v[2 * k]);
                bminy = std::min(bminy, v[2 * k + 1]);
                bmaxy = std::max(bmaxy, v[2 * k + 1]);
            }
            int stx = (int)((bminx - minx) / width);
            int enx = (int)((bmaxx - minx) / width) + 1;
            int sty = (int)((bminy - miny) / height);
            int eny = (int)((bmaxy - miny) / height) + 1;
            if (enx >= xs) enx = xs - 1;
            if (eny >= ys) eny = ys - 1;
            if (stx < 0) stx = 0;
            if (sty < 0) sty = 0;
            for (int x = stx; x <= enx; x++) {
                for (int y = sty; y <= eny; y++) {
                    const float lbx = (float)x * width + minx, lby = (float)y * height + miny;
                    const float rtx = (float)(x + 1) * width + minx, rty = (float)(y + 1) * height + miny;
                    if (cell_candidate(v, n, lbx, lby, rtx, rty)) acc[t].push_back({y * xs + x, p});
                    else if (polygon_meets_rect(v, n, lbx, lby, rtx, rty)) defi[t].push_back(y * xs + x);
                }
            }
        }
    });
    std::vector<Pair> all;
    for (auto& a : acc) all.insert(all.end(), a.begin(), a.end());
    std::sort(all.begin(), all.end(), [](const Pair& a, const Pair& b) {
        return a.cell != b.cell ? a.cell < b.cell : a.poly < b.poly;
    });
    g->cell_off.assign((size_t)xs * ys + 1, 0);
    g->cand.resize(all.size());
    for (size_t k = 0; k < all.size(); k++) {
        g->cell_off[all[k].cell + 1]++;
        g->cand[k] = all[k].poly;
    }
    for (size_t c = 0; c < (size_t)xs * ys; c++) g->cell_off[c + 1] += g->cell_off[c];
    g->deficient.clear();
    for (auto& d : defi) g->deficient.insert(g->deficient.end(), d.begin(), d.end());
    std::sort(g->deficient.begin(), g->deficient.end());
    g->deficient.erase(std::unique(g->deficient.begin(), g->deficient.end()), g->deficient.end());
    build_bands(ps, g);
}

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
