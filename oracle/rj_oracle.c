/* rj_oracle.c -- CPU restatement of the raster-join hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py. The product never links or calls it.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math -fopenmp -fPIC -shared
 * (no FMA contraction, IEEE float division, SSE arithmetic: the same float
 * semantics as the reference CPUJoin built by g++ for x86-64).
 *
 * Reference paths below are relative to the reference repository root.
 */
#include "rj_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* CPUJoin float geometry                                                    */
/* ------------------------------------------------------------------------- */

/* Crossing-number test; CPUJoin/GridIndexF.cpp:383-402 (same operation order:
 * ((vjx - vix) * (ty - viy)) / (vjy - viy) + vix, compared with tx). */
static int pip_f32(const float* poly, int n, float tx, float ty)
{
    int c = 0;
    for (int i = 0, j = n - 1; i < n; j = i++) {
        float vix = poly[2 * i], viy = poly[2 * i + 1];
        float vjx = poly[2 * j], vjy = poly[2 * j + 1];
        if (((viy > ty) != (vjy > ty)) && (tx < (vjx - vix) * (ty - viy) / (vjy - viy) + vix))
            c = !c;
    }
    return c;
}

enum { OC_INSIDE = 0, OC_LEFT = 1, OC_RIGHT = 2, OC_BOTTOM = 4, OC_TOP = 8 };

/* GridIndexF.cpp:499-511 */
static int outcode(float x, float y, float lbx, float lby, float rtx, float rty)
{
    int code = OC_INSIDE;
    if (x < lbx) code |= OC_LEFT;
    else if (x > rtx) code |= OC_RIGHT;
    if (y < lby) code |= OC_BOTTOM;
    else if (y > rty) code |= OC_TOP;
    return code;
}

/* Ray / segment intersection; GridIndexF.cpp:513-549. Returns the hit point. */
static int ray_seg(float ox, float oy, float dx, float dy, float p0x, float p0y, float p1x,
                   float p1y, float* rx, float* ry)
{
    float x1 = ox, y1 = oy;
    float x2 = ox + dx, y2 = oy + dy;
    float x3 = p0x, y3 = p0y, x4 = p1x, y4 = p1y;
    if ((y2 - y1) / (x2 - x1) != (y4 - y3) / (x4 - x3)) {
        float d = ((x2 - x1) * (y4 - y3)) - (y2 - y1) * (x4 - x3);
        if (d != 0) {
            float r = (((y1 - y3) * (x4 - x3)) - (x1 - x3) * (y4 - y3)) / d;
            float s = (((y1 - y3) * (x2 - x1)) - (x1 - x3) * (y2 - y1)) / d;
            if (r >= 0) {
                if (s >= 0 && s <= 1) {
                    *rx = ox + dx * r;
                    *ry = oy + dy * r;
                    return 1;
                }
            }
        }
    }
    return 0;
}

/* GridIndexF.cpp:404-424, including the comparison of the hit's y with the
 * rectangle's right x on the bottom side (line 412). */
static int line_rect(float v1x, float v1y, float v2x, float v2y, float lbx, float lby, float rtx,
                     float rty)
{
    float rx, ry;
    if (ray_seg(lbx, lby, 0.0f, 1.0f, v1x, v1y, v2x, v2y, &rx, &ry) && ry <= rty) return 1;
    if (ray_seg(lbx, lby, 1.0f, 0.0f, v1x, v1y, v2x, v2y, &rx, &ry) && ry <= rtx) return 1;
    if (ray_seg(rtx, rty, 0.0f, -1.0f, v1x, v1y, v2x, v2y, &rx, &ry) && ry >= lby) return 1;
    if (ray_seg(rtx, rty, -1.0f, 0.0f, v1x, v1y, v2x, v2y, &rx, &ry) && rx >= lbx) return 1;
    return 0;
}

/* GridIndexF.cpp:461-496: vertex in rect, rect corner in polygon, then the
 * open chain of edges (i, i+1) for i < n-1 (the closing edge is not tested). */
static int poly_rect_cpu(const float* poly, int n, float lbx, float lby, float rtx, float rty)
{
    for (int i = 0; i < n; i++) {
        float vx = poly[2 * i], vy = poly[2 * i + 1];
        if (vx >= lbx && vx <= rtx && vy >= lby && vy <= rty) return 1;
    }
    if (pip_f32(poly, n, lbx, lby)) return 1;
    if (pip_f32(poly, n, rtx, rty)) return 1;
    if (pip_f32(poly, n, lbx, rty)) return 1;
    if (pip_f32(poly, n, rtx, lby)) return 1;
    int c2 = outcode(poly[0], poly[1], lbx, lby, rtx, rty);
    for (int i = 0; i < n - 1; i++) {
        int c1 = c2;
        c2 = outcode(poly[2 * (i + 1)], poly[2 * (i + 1) + 1], lbx, lby, rtx, rty);
        if (!(c1 & c2) &&
            line_rect(poly[2 * i], poly[2 * i + 1], poly[2 * (i + 1)], poly[2 * (i + 1) + 1], lbx,
                      lby, rtx, rty))
            return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Grid index: GridIndexF::buildGrid (GridIndexF.cpp:60-155)                 */
/* ------------------------------------------------------------------------- */

struct orc_grid {
    int32_t xs, ys, np;
    float minx, miny, width, height;
    const float* xy;      /* borrowed polygon vertices */
    const int32_t* off;   /* borrowed offsets */
    int32_t* cell_off;    /* xs*ys+1 */
    int32_t* cand;
    int64_t ncand;
};

typedef struct { int32_t cell, poly; } pair_t;

static int cmp_pair(const void* a, const void* b)
{
    const pair_t* x = (const pair_t*)a;
    const pair_t* y = (const pair_t*)b;
    if (x->cell != y->cell) return x->cell < y->cell ? -1 : 1;
    return (x->poly > y->poly) - (x->poly < y->poly);
}

orc_grid* orc_grid_build(const float* xy, const int32_t* off, int32_t np, const float bbox[4],
                         int32_t xs, int32_t ys, int32_t nthreads)
{
    if (np < 0 || xs <= 0 || ys <= 0) return NULL;
    orc_grid* g = (orc_grid*)calloc(1, sizeof(orc_grid));
    g->xs = xs; g->ys = ys; g->np = np; g->xy = xy; g->off = off;
    g->minx = bbox[0]; g->miny = bbox[1];
    g->width = (bbox[2] - bbox[0]) / (float)xs;   /* GridIndexF.cpp:66-67 */
    g->height = (bbox[3] - bbox[1]) / (float)ys;
    if (nthreads <= 0) nthreads = omp_get_max_threads();

    pair_t** tp = (pair_t**)calloc(nthreads, sizeof(pair_t*));
    int64_t* tn = (int64_t*)calloc(nthreads, sizeof(int64_t));
    int64_t* tcap = (int64_t*)calloc(nthreads, sizeof(int64_t));

#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int bc = 0; bc < np; bc++) {
        int th = omp_get_thread_num();
        const float* poly = xy + 2 * (int64_t)off[bc];
        int n = off[bc + 1] - off[bc];
        float bminx = 3.402823466e+38f, bminy = 3.402823466e+38f;
        float bmaxx = -3.402823466e+38f, bmaxy = -3.402823466e+38f;
        for (int k = 0; k < n; k++) {   /* BoundF::updateBound, Common.h:99-104 */
            float x = poly[2 * k], y = poly[2 * k + 1];
            bminx = x < bminx ? x : bminx;  /* std::min(minx, x) */
            bmaxx = bmaxx < x ? x : bmaxx;  /* std::max(maxx, x) */
            bminy = y < bminy ? y : bminy;
            bmaxy = bmaxy < y ? y : bmaxy;
        }
        int stx = (int)((bminx - g->minx) / g->width);
        int enx = (int)((bmaxx - g->minx) / g->width) + 1;
        int sty = (int)((bminy - g->miny) / g->height);
        int eny = (int)((bmaxy - g->miny) / g->height) + 1;
        if (enx >= xs) enx = xs - 1;
        if (eny >= ys) eny = ys - 1;
        if (stx < 0) stx = 0;
        if (sty < 0) sty = 0;
        for (int x = stx; x <= enx; x++) {
            for (int y = sty; y <= eny; y++) {
                float lbx = (float)x * g->width + g->minx;        /* getX, :49-52 */
                float lby = (float)y * g->height + g->miny;
                float rtx = (float)(x + 1) * g->width + g->minx;
                float rty = (float)(y + 1) * g->height + g->miny;
                if (poly_rect_cpu(poly, n, lbx, lby, rtx, rty)) {
                    if (tn[th] == tcap[th]) {
                        tcap[th] = tcap[th] ? 2 * tcap[th] : 1024;
                        tp[th] = (pair_t*)realloc(tp[th], tcap[th] * sizeof(pair_t));
                    }
                    tp[th][tn[th]].cell = y * xs + x;
                    tp[th][tn[th]].poly = bc;
                    tn[th]++;
                }
            }
        }
    }
    int64_t total = 0;
    for (int t = 0; t < nthreads; t++) total += tn[t];
    pair_t* all = (pair_t*)malloc((total ? total : 1) * sizeof(pair_t));
    int64_t pos = 0;
    for (int t = 0; t < nthreads; t++) {
        if (tn[t]) memcpy(all + pos, tp[t], tn[t] * sizeof(pair_t));
        pos += tn[t];
        free(tp[t]);
    }
    free(tp); free(tn); free(tcap);
    qsort(all, total, sizeof(pair_t), cmp_pair);
    int64_t ncell = (int64_t)xs * ys;
    g->cell_off = (int32_t*)calloc(ncell + 1, sizeof(int32_t));
    g->cand = (int32_t*)malloc((total ? total : 1) * sizeof(int32_t));
    for (int64_t k = 0; k < total; k++) {
        g->cell_off[all[k].cell + 1]++;
        g->cand[k] = all[k].poly;
    }
    for (int64_t c = 0; c < ncell; c++) g->cell_off[c + 1] += g->cell_off[c];
    g->ncand = total;
    free(all);
    return g;
}

int64_t orc_grid_ncand(const orc_grid* g) { return g->ncand; }

void orc_grid_export(const orc_grid* g, int32_t* cell_off, int32_t* cand)
{
    int64_t ncell = (int64_t)g->xs * g->ys;
    if (cell_off) memcpy(cell_off, g->cell_off, (ncell + 1) * sizeof(int32_t));
    if (cand && g->ncand) memcpy(cand, g->cand, g->ncand * sizeof(int32_t));
}

void orc_grid_free(orc_grid* g)
{
    if (!g) return;
    free(g->cell_off);
    free(g->cand);
    free(g);
}

/* ------------------------------------------------------------------------- */
/* Filters and aggregate values                                              */
/* ------------------------------------------------------------------------- */

/* Conjunctive filters (index.glsl:184-208 / points.vert:138-149) with integer
 * comparisons for Uint and Int columns and float comparisons for Float. */
static int pass_filters(const orc_points* P, int64_t i)
{
    for (int c = 0; c < P->ncons; c++) {
        const orc_constraint* k = &P->cons[c];
        int col = k->attrib_id;
        int t = P->types[col];
        int ok = 0;
        if (t == 1 || t == 4) {
            uint32_t a = (t == 1) ? ((const uint32_t*)P->cols[col])[i]
                                  : (uint32_t)((const uint8_t*)P->cols[col])[i];
            uint32_t v = k->val.uval;
            switch (k->type) {
            case 0: ok = a == v; break;
            case 1: ok = a < v; break;
            case 2: ok = a <= v; break;
            case 3: ok = a > v; break;
            case 4: ok = a >= v; break;
            }
        } else if (t == 2) {
            int32_t a = ((const int32_t*)P->cols[col])[i];
            int32_t v = k->val.ival;
            switch (k->type) {
            case 0: ok = a == v; break;
            case 1: ok = a < v; break;
            case 2: ok = a <= v; break;
            case 3: ok = a > v; break;
            case 4: ok = a >= v; break;
            }
        } else if (t == 3) {
            float a = ((const float*)P->cols[col])[i];
            float v = k->val.fval;
            switch (k->type) {
            case 0: ok = a == v; break;
            case 1: ok = a < v; break;
            case 2: ok = (a < v) || (a == v); break;
            case 3: ok = a > v; break;
            case 4: ok = (a > v) || (a == v); break;
            }
        }
        if (!ok) return 0;
    }
    return 1;
}

/* Exact fixed-point value: v = hi + lo * 2^-32 with hi = floor(v), lo = the
 * fraction times 2^32 truncated (exact for |v| >= 2^-9). Integer columns give
 * hi = value, lo = 0. Non-finite or |v| >= 2^62 contributes 0. A float in
 * [-2^-25, 0), whose float fraction v - floor(v) rounds to 1, counts as 0. */
static void agg_value(const orc_points* P, int64_t i, int64_t* hi, uint64_t* lo, double* d)
{
    int col = P->agg_col;
    int t = P->types[col];
    *lo = 0;
    if (t == 1) {
        uint32_t a = ((const uint32_t*)P->cols[col])[i];
        *hi = (int64_t)a; *d = (double)a;
    } else if (t == 4) {
        uint8_t a = ((const uint8_t*)P->cols[col])[i];
        *hi = (int64_t)a; *d = (double)a;
    } else if (t == 2) {
        int32_t a = ((const int32_t*)P->cols[col])[i];
        *hi = (int64_t)a; *d = (double)a;
    } else {
        float f = ((const float*)P->cols[col])[i];
        if (!(fabsf(f) < 4.611686018427387904e18f)) { *hi = 0; *d = 0.0; return; }
        float fl = floorf(f);
        float fr = f - fl;
        *hi = (int64_t)fl;
        *d = (double)f;
        if (fr >= 1.0f) { *hi += 1; return; }
        *lo = (uint64_t)(fr * 4294967296.0f);
    }
}

/* ------------------------------------------------------------------------- */
/* Exact join: CPUJoin::execute (CPUJoin.cpp:62-122) + getRegion (:158-175)  */
/* ------------------------------------------------------------------------- */

int orc_join_exact(const orc_grid* g, const orc_points* P, int32_t nthreads, int64_t* counts,
                   int64_t* sum_hi, uint64_t* sum_lo, double* sum_f64)
{
    int np = g->np;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    int want_sum = (P->agg_type != 0);
    int64_t* tc = (int64_t*)calloc((size_t)nthreads * np + 1, sizeof(int64_t));
    int64_t* th = want_sum ? (int64_t*)calloc((size_t)nthreads * np + 1, sizeof(int64_t)) : NULL;
    uint64_t* tl = want_sum ? (uint64_t*)calloc((size_t)nthreads * np + 1, sizeof(uint64_t)) : NULL;
    double* td = want_sum ? (double*)calloc((size_t)nthreads * np + 1, sizeof(double)) : NULL;

#pragma omp parallel num_threads(nthreads)
    {
        int t = omp_get_thread_num();
        int64_t* c = tc + (size_t)t * np;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < P->n; i++) {
            if (P->ncons && !pass_filters(P, i)) continue;
            float x = P->xy[2 * i], y = P->xy[2 * i + 1];
            int stx = (int)((x - g->minx) / g->width);    /* getXIndex, GridIndexF.cpp:39-42 */
            int sty = (int)((y - g->miny) / g->height);
            if (stx < 0 || stx >= g->xs || sty < 0 || sty >= g->ys) continue;
            int cell = sty * g->xs + stx;
            int64_t hi = 0; uint64_t lo = 0; double d = 0;
            int have = 0;
            for (int k = g->cell_off[cell]; k < g->cell_off[cell + 1]; k++) {
                int p = g->cand[k];
                const float* poly = g->xy + 2 * (int64_t)g->off[p];
                int n = g->off[p + 1] - g->off[p];
                if (pip_f32(poly, n, x, y)) {
                    c[p]++;
                    if (want_sum) {
                        if (!have) { agg_value(P, i, &hi, &lo, &d); have = 1; }
                        th[(size_t)t * np + p] += hi;
                        tl[(size_t)t * np + p] += lo;
                        td[(size_t)t * np + p] += d;
                    }
                }
            }
        }
    }
    for (int p = 0; p < np; p++) {
        int64_t cs = 0, hs = 0; uint64_t ls = 0; double ds = 0;
        for (int t = 0; t < nthreads; t++) {
            cs += tc[(size_t)t * np + p];
            if (want_sum) {
                hs += th[(size_t)t * np + p];
                ls += tl[(size_t)t * np + p];
                ds += td[(size_t)t * np + p];
            }
        }
        if (counts) counts[p] = cs;
        if (sum_hi) sum_hi[p] = hs;
        if (sum_lo) sum_lo[p] = ls;
        if (sum_f64) sum_f64[p] = ds;
    }
    free(tc); free(th); free(tl); free(td);
    return 0;
}

int orc_join_brute(const float* xy, const int32_t* off, int32_t np, const orc_points* P,
                   int64_t* counts)
{
    memset(counts, 0, sizeof(int64_t) * np);
    for (int64_t i = 0; i < P->n; i++) {
        if (P->ncons && !pass_filters(P, i)) continue;
        float x = P->xy[2 * i], y = P->xy[2 * i + 1];
        for (int p = 0; p < np; p++) {
            if (pip_f32(xy + 2 * (int64_t)off[p], off[p + 1] - off[p], x, y)) counts[p]++;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Canvas geometry: GLFunction::setMaxCellSize (GLFunction.cpp:26-30) and    */
/* RasterJoin::updateBuffers (RasterJoin.cpp:51-78) with maxRes = 8192.      */
/* ------------------------------------------------------------------------- */

int orc_plan_make(const double bbox[4], double eps_m, double ground_res, orc_plan* p)
{
    memset(p, 0, sizeof(*p));
    double diffx = bbox[2] - bbox[0], diffy = bbox[3] - bbox[1];
    if (!(diffx > 0) || !(diffy > 0) || !(eps_m > 0) || !(ground_res > 0)) return -1;
    double cell = (eps_m / sqrt(2.0)) / ground_res;
    double ax = ceil(diffx / cell), ay = ceil(diffy / cell);
    if (ax > 1e9 || ay > 1e9) return -1;
    int actx = (int)ax, acty = (int)ay;
    int splitx = (int)ceil((double)actx / 8192.0);
    int splity = (int)ceil((double)acty / 8192.0);
    int resx = (int)ceil((double)actx / splitx);
    int resy = (int)ceil((double)acty / splity);
    p->W = splitx * resx;
    p->H = splity * resy;
    if ((int64_t)p->W * 256 >= ((int64_t)1 << 31) || (int64_t)p->H * 256 >= ((int64_t)1 << 31))
        return -1;
    p->splitx = splitx; p->splity = splity;
    p->minx = bbox[0]; p->miny = bbox[1]; p->maxx = bbox[2]; p->maxy = bbox[3];
    p->minx_f = (float)bbox[0]; p->miny_f = (float)bbox[1];
    p->maxx_f = (float)bbox[2]; p->maxy_f = (float)bbox[3];
    p->sx_f = (float)((double)p->W / diffx);
    p->sy_f = (float)((double)p->H / diffy);
    p->fxs = (double)p->W * 256.0 / diffx;
    p->fys = (double)p->H * 256.0 / diffy;
    p->cell = cell;
    return 0;
}

/* The one pixel formula of every kernel: canvas membership by exact float
 * comparison with the bbox ([min, max) per axis), pixel = int((v - min) * s),
 * clamped to the last pixel. */
int64_t orc_pixel(const orc_plan* p, float x, float y)
{
    if (!(x >= p->minx_f && x < p->maxx_f && y >= p->miny_f && y < p->maxy_f)) return -1;
    float fx = (x - p->minx_f) * p->sx_f;
    float fy = (y - p->miny_f) * p->sy_f;
    int ix = (int)fx, iy = (int)fy;
    if (ix > p->W - 1) ix = p->W - 1;
    if (iy > p->H - 1) iy = p->H - 1;
    return (int64_t)iy * p->W + ix;
}

void orc_snap(const orc_plan* p, const float* xy, int64_t nv, int64_t* X, int64_t* Y)
{
    for (int64_t k = 0; k < nv; k++) {
        X[k] = (int64_t)llrint(((double)xy[2 * k] - p->minx) * p->fxs);
        Y[k] = (int64_t)llrint(((double)xy[2 * k + 1] - p->miny) * p->fys);
    }
}

/* ------------------------------------------------------------------------- */
/* Exact fixed-point coverage (pixel-center rule)                            */
/* ------------------------------------------------------------------------- */

static int64_t floor_div(int64_t a, int64_t b)  /* b > 0 */
{
    int64_t q = a / b;
    if ((a % b != 0) && (a < 0)) q--;
    return q;
}

static int64_t ceil_div(int64_t a, int64_t b)   /* b > 0 */
{
    int64_t q = a / b;
    if ((a % b != 0) && (a > 0)) q++;
    return q;
}

/* Crossing number at (cx, cy) with exact integer arithmetic: edge (i, j)
 * straddles when (yi > cy) != (yj > cy) and counts when cx < x(cy). */
static int inside_fixed(const int64_t* X, const int64_t* Y, int n, int64_t cx, int64_t cy)
{
    int c = 0;
    for (int i = 0, j = n - 1; i < n; j = i++) {
        int64_t yi = Y[i], yj = Y[j];
        if ((yi > cy) != (yj > cy)) {
            int64_t dy = yj - yi;
            int64_t lhs = (cx - X[i]) * dy;
            int64_t rhs = (X[j] - X[i]) * (cy - yi);
            if (dy > 0 ? lhs < rhs : lhs > rhs) c ^= 1;
        }
    }
    return c;
}

/* Edge-flag fill of the centers of a grid of cells of `cs` subunits (256 =
 * pixels, 64 = quarter pixels) over cells [c0, c1] x [r0, r1]. Each straddling
 * edge toggles the first cell whose center is not left of the crossing; a
 * running XOR along the row gives the crossing parity of every center. */
static void fill_region(const int64_t* X, const int64_t* Y, int n, int64_t cs, int64_t c0,
                        int64_t c1, int64_t r0, int64_t r1, uint8_t* out)
{
    int64_t bw = c1 - c0 + 1, bh = r1 - r0 + 1;
    int64_t h = cs / 2;
    memset(out, 0, (size_t)(bw * bh));
    for (int i = 0, j = n - 1; i < n; j = i++) {
        int64_t ax = X[i], ay = Y[i], bx = X[j], by = Y[j];
        if (ay == by) continue;
        if (ay > by) {
            int64_t t = ax; ax = bx; bx = t;
            t = ay; ay = by; by = t;
        }
        int64_t dy = by - ay, dx = bx - ax;
        int64_t jlo = ceil_div(ay - h, cs), jhi = ceil_div(by - h, cs) - 1;
        if (jlo < r0) jlo = r0;
        if (jhi > r1) jhi = r1;
        for (int64_t row = jlo; row <= jhi; row++) {
            int64_t cy = row * cs + h;
            int64_t N = ax * dy + dx * (cy - ay);
            int64_t k = ceil_div(N - h * dy, cs * dy);
            int64_t local = k - c0;
            if (local < 0) local = 0;
            if (local >= bw) continue;
            out[(row - r0) * bw + local] ^= 1;
        }
    }
    for (int64_t r = 0; r < bh; r++) {
        uint8_t acc = 0;
        for (int64_t c = 0; c < bw; c++) {
            acc ^= out[r * bw + c];
            out[r * bw + c] = acc;
        }
    }
}

int orc_coverage(const orc_plan* p, const float* xy, int32_t nv, int32_t method, uint8_t* out)
{
    /* method 0/1: pixel centers; 2/3: quarter-pixel centers (4W x 4H) */
    int sub = (method >= 2) ? 4 : 1;
    int64_t cs = 256 / sub;
    int64_t W = (int64_t)p->W * sub, H = (int64_t)p->H * sub;
    int64_t* X = (int64_t*)malloc(sizeof(int64_t) * nv);
    int64_t* Y = (int64_t*)malloc(sizeof(int64_t) * nv);
    orc_snap(p, xy, nv, X, Y);
    if ((method & 1) == 0) {
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < H; r++)
            for (int64_t c = 0; c < W; c++)
                out[r * W + c] = (uint8_t)inside_fixed(X, Y, nv, c * cs + cs / 2, r * cs + cs / 2);
    } else {
        fill_region(X, Y, nv, cs, 0, W - 1, 0, H - 1, out);
    }
    free(X); free(Y);
    return 0;
}

/* Pixel bbox (inclusive) of the cells whose centers can be inside. */
static void poly_cells(const int64_t* X, const int64_t* Y, int n, int64_t cs, int64_t W, int64_t H,
                       int64_t pad, int64_t* c0, int64_t* c1, int64_t* r0, int64_t* r1)
{
    int64_t mnx = X[0], mxx = X[0], mny = Y[0], mxy = Y[0];
    for (int k = 1; k < n; k++) {
        if (X[k] < mnx) mnx = X[k];
        if (X[k] > mxx) mxx = X[k];
        if (Y[k] < mny) mny = Y[k];
        if (Y[k] > mxy) mxy = Y[k];
    }
    *c0 = floor_div(mnx, cs) - 1 - pad; *c1 = floor_div(mxx, cs) + 1 + pad;
    *r0 = floor_div(mny, cs) - 1 - pad; *r1 = floor_div(mxy, cs) + 1 + pad;
    if (*c0 < 0) *c0 = 0;
    if (*r0 < 0) *r0 = 0;
    if (*c1 > W - 1) *c1 = W - 1;
    if (*r1 > H - 1) *r1 = H - 1;
}

/* ------------------------------------------------------------------------- */
/* Raster join: splat (points.vert + additive blend) then per-polygon sums   */
/* of the canvas over covered pixels (polygon.frag:12-28).                   */
/* ------------------------------------------------------------------------- */

int orc_raster_join(const orc_plan* p, const float* pxy, const int32_t* off, int32_t np,
                    const orc_points* P, int64_t* counts, int64_t* sum_hi, uint64_t* sum_lo,
                    uint32_t* canvas_out)
{
    int64_t npx = (int64_t)p->W * p->H;
    int want_sum = (P->agg_type != 0);
    uint32_t* cnt = canvas_out ? canvas_out : (uint32_t*)malloc(sizeof(uint32_t) * npx);
    int64_t* chi = want_sum ? (int64_t*)calloc(npx, sizeof(int64_t)) : NULL;
    uint64_t* clo = want_sum ? (uint64_t*)calloc(npx, sizeof(uint64_t)) : NULL;
    memset(cnt, 0, sizeof(uint32_t) * npx);
    for (int64_t i = 0; i < P->n; i++) {
        if (P->ncons && !pass_filters(P, i)) continue;
        int64_t px = orc_pixel(p, P->xy[2 * i], P->xy[2 * i + 1]);
        if (px < 0) continue;
        cnt[px]++;
        if (want_sum) {
            int64_t hi; uint64_t lo; double d;
            agg_value(P, i, &hi, &lo, &d);
            chi[px] += hi;
            clo[px] += lo;
        }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < np; q++) {
        int n = off[q + 1] - off[q];
        int64_t* X = (int64_t*)malloc(sizeof(int64_t) * n);
        int64_t* Y = (int64_t*)malloc(sizeof(int64_t) * n);
        orc_snap(p, pxy + 2 * (int64_t)off[q], n, X, Y);
        int64_t c0, c1, r0, r1;
        poly_cells(X, Y, n, 256, p->W, p->H, 0, &c0, &c1, &r0, &r1);
        int64_t bw = c1 - c0 + 1, bh = r1 - r0 + 1;
        int64_t cs = 0, hs = 0; uint64_t ls = 0;
        if (bw > 0 && bh > 0) {
            uint8_t* in = (uint8_t*)malloc((size_t)(bw * bh));
            fill_region(X, Y, n, 256, c0, c1, r0, r1, in);
            for (int64_t r = 0; r < bh; r++)
                for (int64_t c = 0; c < bw; c++)
                    if (in[r * bw + c]) {
                        int64_t pix = (r + r0) * p->W + (c + c0);
                        cs += cnt[pix];
                        if (want_sum) { hs += chi[pix]; ls += clo[pix]; }
                    }
            free(in);
        }
        if (counts) counts[q] = cs;
        if (sum_hi) sum_hi[q] = hs;
        if (sum_lo) sum_lo[q] = ls;
        free(X); free(Y);
    }
    if (!canvas_out) free(cnt);
    free(chi); free(clo);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Boundary pixels: every pixel whose closed square meets a snapped edge,    */
/* dilated by r pixels (Chebyshev).                                          */
/* ------------------------------------------------------------------------- */

typedef void (*mark_fn)(void* ctx, int64_t i, int64_t j);

static void touched(int64_t ax, int64_t ay, int64_t bx, int64_t by, int64_t W, int64_t H,
                    mark_fn mark, void* ctx)
{
    if (ay > by) {
        int64_t t = ax; ax = bx; bx = t;
        t = ay; ay = by; by = t;
    }
    int64_t jlo = ceil_div(ay, 256) - 1, jhi = floor_div(by, 256);
    if (jlo < 0) jlo = 0;
    if (jhi > H - 1) jhi = H - 1;
    for (int64_t j = jlo; j <= jhi; j++) {
        int64_t xlo, xhi;
        if (ay == by) {
            xlo = ax < bx ? ax : bx;
            xhi = ax < bx ? bx : ax;
        } else {
            int64_t ylo = ay > j * 256 ? ay : j * 256;
            int64_t yhi = by < (j + 1) * 256 ? by : (j + 1) * 256;
            int64_t den = by - ay;
            int64_t n0 = ax * den + (bx - ax) * (ylo - ay);
            int64_t n1 = ax * den + (bx - ax) * (yhi - ay);
            int64_t f0 = floor_div(n0, den), f1 = floor_div(n1, den);
            int64_t g0 = ceil_div(n0, den), g1 = ceil_div(n1, den);
            xlo = f0 < f1 ? f0 : f1;
            xhi = g0 > g1 ? g0 : g1;
        }
        int64_t ilo = ceil_div(xlo, 256) - 1, ihi = floor_div(xhi, 256);
        if (ilo < 0) ilo = 0;
        if (ihi > W - 1) ihi = W - 1;
        for (int64_t i = ilo; i <= ihi; i++) mark(ctx, i, j);
    }
}

typedef struct { uint8_t* m; int64_t W, H, r; int64_t c0, r0, bw, bh; } mark_ctx;

static void mark_dilate(void* vctx, int64_t i, int64_t j)
{
    mark_ctx* c = (mark_ctx*)vctx;
    for (int64_t dj = -c->r; dj <= c->r; dj++) {
        int64_t jj = j + dj - c->r0;
        if (jj < 0 || jj >= c->bh) continue;
        for (int64_t di = -c->r; di <= c->r; di++) {
            int64_t ii = i + di - c->c0;
            if (ii < 0 || ii >= c->bw) continue;
            c->m[jj * c->bw + ii] = 1;
        }
    }
}

int orc_boundary(const orc_plan* p, const float* xy, int32_t nv, int32_t r, uint8_t* out)
{
    int64_t* X = (int64_t*)malloc(sizeof(int64_t) * nv);
    int64_t* Y = (int64_t*)malloc(sizeof(int64_t) * nv);
    orc_snap(p, xy, nv, X, Y);
    mark_ctx c = { out, p->W, p->H, r, 0, 0, p->W, p->H };
    for (int i = 0, j = nv - 1; i < nv; j = i++)
        touched(X[j], Y[j], X[i], Y[i], p->W, p->H, mark_dilate, &c);
    free(X); free(Y);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Area of a ring inside each pixel (the area fraction of line-bound.frag's   */
/* computeFraction, :48-187, taken over the whole polygon): for every row a  */
/* snapped edge crosses, the area of the row's piece between the edge and    */
/* each column's right side, G(X) = int max(X - x(y), 0) dy floored to 1/256 */
/* square subunit, is deposited as column differences; the row prefix sum of */
/* the deposits of all edges, signed by edge direction, is the pixel's area  */
/* (2^24 = the whole pixel).                                                 */
/* ------------------------------------------------------------------------- */

#define ORC_AREA_ONE 16777216LL

static int64_t area_g(double X, double lo, double hi, double mid, double hd, double k)
{
    double g;
    if (X <= lo) return 0;
    if (X >= hi) g = hd * (X - mid);
    else g = (X - lo) * (X - lo) * k;
    return (int64_t)floor(g * 256.0);
}

/* Deposits of edge (ax,ay)-(bx,by) into acc (bh rows of bw, first pixel     */
/* (c0, r0)); columns left of c0 land on c0, columns past the region drop.   */
static void area_edge(int64_t ax, int64_t ay, int64_t bx, int64_t by, int64_t c0, int64_t r0,
                      int64_t bw, int64_t bh, uint32_t* acc)
{
    if (ay == by) return;
    uint32_t sg = 0xFFFFFFFFu;  /* upward edge: -1, so a counter-clockwise ring sums positive */
    if (by < ay) {
        int64_t t = ax; ax = bx; bx = t;
        t = ay; ay = by; by = t;
        sg = 1u;
    }
    int64_t den = by - ay;
    int64_t j0 = floor_div(ay, 256), j1 = ceil_div(by, 256) - 1;
    if (j0 < r0) j0 = r0;
    if (j1 > r0 + bh - 1) j1 = r0 + bh - 1;
    for (int64_t j = j0; j <= j1; j++) {
        int64_t ya = ay > 256 * j ? ay : 256 * j;
        int64_t yb = by < 256 * j + 256 ? by : 256 * j + 256;
        int64_t na = ax * den + (bx - ax) * (ya - ay);
        int64_t nb = ax * den + (bx - ax) * (yb - ay);
        double ua = (double)na / (double)den, ub = (double)nb / (double)den;
        double lo = ua < ub ? ua : ub, hi = ua < ub ? ub : ua;
        double mid = 0.5 * (ua + ub), hd = (double)(yb - ya);
        double k = hi > lo ? hd / (2.0 * (hi - lo)) : 0.0;
        int64_t cl = floor_div(na < nb ? na : nb, 256 * den);
        int64_t ch = floor_div(na < nb ? nb : na, 256 * den);
        uint32_t* row = acc + (j - r0) * bw;
        int64_t gl = 0, fl = 0;
        for (int64_t c = cl; c <= ch + 1; c++) {
            int64_t f;
            if (c <= ch) {
                int64_t g = area_g((double)(256 * (c + 1)), lo, hi, mid, hd, k);
                f = g - gl;
                gl = g;
            } else {
                f = 65536 * (yb - ya);
            }
            uint32_t d = (uint32_t)(f - fl) * sg;
            fl = f;
            int64_t i = c - c0 < 0 ? 0 : c - c0;
            if (i < bw) row[i] += d;
        }
    }
}

/* Area shares of one ring over the rectangle [c0, c0+bw) x [r0, r0+bh):    */
/* acc receives the signed per-pixel areas (int32 in 2^-24 of a pixel,      */
/* positive for a counter-clockwise ring).                                   */
static void area_region(const int64_t* X, const int64_t* Y, int n, int64_t c0, int64_t r0,
                        int64_t bw, int64_t bh, uint32_t* acc)
{
    memset(acc, 0, sizeof(uint32_t) * (size_t)(bw * bh));
    for (int i = 0, j = n - 1; i < n; j = i++)
        area_edge(X[j], Y[j], X[i], Y[i], c0, r0, bw, bh, acc);
    for (int64_t r = 0; r < bh; r++) {
        uint32_t run = 0;
        for (int64_t c = 0; c < bw; c++) {
            run += acc[r * bw + c];
            acc[r * bw + c] = run;
        }
    }
}

/* min(a, 1 - a) of a pixel's signed area sum, |area| capped at one pixel. */
static uint32_t area_min_share(uint32_t acc)
{
    int64_t v = (int32_t)acc;
    if (v < 0) v = -v;
    if (v > ORC_AREA_ONE) v = ORC_AREA_ONE;
    return (uint32_t)(v < ORC_AREA_ONE - v ? v : ORC_AREA_ONE - v);
}

int orc_area(const orc_plan* p, const float* xy, int32_t nv, int32_t* out)
{
    int64_t* X = (int64_t*)malloc(sizeof(int64_t) * nv);
    int64_t* Y = (int64_t*)malloc(sizeof(int64_t) * nv);
    orc_snap(p, xy, nv, X, Y);
    area_region(X, Y, nv, 0, 0, p->W, p->H, (uint32_t*)out);
    free(X); free(Y);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Errorbounds (RasterJoinBounds.cpp:306-381, line-bound.frag:219-262) with  */
/* boundary pixels = dilate1(touched) per polygon, inside = pixel-center     */
/* rule, and the estimate ceil(min(a, 1 - a) * count) per boundary pixel     */
/* from the exact area share a of the polygon in the pixel (area_region).    */
/* ------------------------------------------------------------------------- */

int orc_errorbounds(const orc_plan* p, const float* pxy, const int32_t* off, int32_t np,
                    const uint32_t* cnt, int64_t* bounds)
{
    memset(bounds, 0, sizeof(int64_t) * 4 * (size_t)np);
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < np; q++) {
        int n = off[q + 1] - off[q];
        int64_t* X = (int64_t*)malloc(sizeof(int64_t) * n);
        int64_t* Y = (int64_t*)malloc(sizeof(int64_t) * n);
        orc_snap(p, pxy + 2 * (int64_t)off[q], n, X, Y);
        int64_t c0, c1, r0, r1;
        poly_cells(X, Y, n, 256, p->W, p->H, 1, &c0, &c1, &r0, &r1);
        int64_t bw = c1 - c0 + 1, bh = r1 - r0 + 1;
        int64_t fn = 0, fp = 0, fn1 = 0, fp1 = 0;
        if (bw > 0 && bh > 0) {
            uint8_t* bm = (uint8_t*)calloc((size_t)(bw * bh), 1);
            uint8_t* in = (uint8_t*)malloc((size_t)(bw * bh));
            uint32_t* area = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(bw * bh));
            mark_ctx c = { bm, p->W, p->H, 1, c0, r0, bw, bh };
            for (int i = 0, j = n - 1; i < n; j = i++)
                touched(X[j], Y[j], X[i], Y[i], p->W, p->H, mark_dilate, &c);
            fill_region(X, Y, n, 256, c0, c1, r0, r1, in);
            area_region(X, Y, n, c0, r0, bw, bh, area);
            for (int64_t r = 0; r < bh; r++)
                for (int64_t cc = 0; cc < bw; cc++) {
                    if (!bm[r * bw + cc]) continue;
                    uint32_t val = cnt[(r + r0) * p->W + (cc + c0)];
                    if (val == 0) continue;
                    uint64_t m = area_min_share(area[r * bw + cc]);
                    int64_t add = (int64_t)((m * val + (uint64_t)(ORC_AREA_ONE - 1)) >> 24);
                    if (in[r * bw + cc]) { fp += add; fp1 += val; }
                    else { fn += add; fn1 += val; }
                }
            free(bm); free(in); free(area);
        }
        bounds[q] = fn;
        bounds[q + np] = fp;
        bounds[q + 2 * (size_t)np] = fn1;
        bounds[q + 3 * (size_t)np] = fp1;
        free(X); free(Y);
    }
    return 0;
}
