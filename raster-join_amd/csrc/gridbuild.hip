// gridbuild.hip -- the CPUJoin grid index of a polygon collection, built on
// the device: the candidate polygons of every grid cell
// (GridIndexF::buildGrid, CPUJoin/GridIndexF.cpp:60-155, with the cell test
// polygonRectIntersectionCPU, :461-496), the cells whose candidate list
// misses a polygon that meets them, and the edge bands of the exact test.
// The reference builds its GPU index the same way in two dispatches, sizes
// then lists (GLFunction::createPolyIndex, GLFunction.cpp:65-156;
// index-size.glsl, create-index.glsl), with the GL shader's exclusive cell
// loop; here the cell range and the test follow CPUJoin, whose lists decide
// the index and hybrid joins' results.
//
// Compiled with -ffp-contract=off: the cell bounds and the cell test must
// round as CPUJoin's float code does, and the candidate lists come out in
// CPUJoin's order (cell-major, polygons ascending).
//
//   1. k_grid_cells: one workgroup per rectangle of up to 64 x 32 of a
//      polygon's bounding-box cells. The cell test is an OR of terms that
//      each follow from one vertex or one edge: the cells holding a vertex,
//      the cells an edge (i, i+1) meets, and the cell corners inside the
//      polygon, whose crossing parity comes from per-row lists of the
//      straddling edges' crossings; the deficient test likewise. So the
//      work grows with the vertices plus the cells, not their product. An
//      item whose float grid is too fine for that decomposition (or whose
//      row lists overflow) tests each cell against every edge directly.
//      Candidates append (cell, polygon), deficient cells the cell.
//   2. radix sort of the (cell << 32 | polygon) keys; k_grid_offsets writes
//      the candidate list and every cell's first entry.
//   3. edge bands: k_edge_rows gives every edge its band range (and each
//      polygon its first and last band), then one thread per edge counts
//      itself into each of its bands and, after a scan, writes itself into
//      them (atomic cursors: a band's edges in no particular order).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "engine.hpp"

namespace rj {
namespace {

constexpr int kGridLdsVerts = 2048;  // vertices a workgroup stages in LDS
constexpr int kItemW = 64;           // cell columns of one work item
constexpr int kItemH = 32;           // cell rows of one work item
constexpr int kPtRowCap = 120;       // crossings kept per grid-point row of an item
constexpr int kCenRowCap = 128;      // straddling edges kept per cell-center row of an item

struct GridItem {
    int32_t poly;
    int32_t x0, y0;  // first cell column and row
    int32_t wh;      // cell columns | cell rows << 16
};

// Crossing number, GridIndexF::isInsidePoly (GridIndexF.cpp:383-402).
__device__ __forceinline__ bool pip(const float* v, int n, float tx, float ty)
{
    bool c = false;
    for (int i = 0, j = n - 1; i < n; j = i++) {
        const float vix = v[2 * i], viy = v[2 * i + 1], vjx = v[2 * j], vjy = v[2 * j + 1];
        if (((viy > ty) != (vjy > ty)) && (tx < (vjx - vix) * (ty - viy) / (vjy - viy) + vix)) c = !c;
    }
    return c;
}

// Cohen-Sutherland region code, GridIndexF::ComputeOutCode (:499-511).
__device__ __forceinline__ int region_code(float x, float y, float lbx, float lby, float rtx, float rty)
{
    int code = 0;
    if (x < lbx) code |= 1;
    else if (x > rtx) code |= 2;
    if (y < lby) code |= 4;
    else if (y > rty) code |= 8;
    return code;
}

// Ray from o along d against segment p0-p1, GridIndexF::rayToLineSegmentIntersection
// (:513-549): parallel when the two slopes compare equal, hit when r >= 0 and
// 0 <= s <= 1; the hit point is o + d*r.
__device__ __forceinline__ bool ray_hits(float ox, float oy, float dx, float dy, float p0x, float p0y, float p1x,
                                         float p1y, float* hx, float* hy)
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

// GridIndexF::lineRectIntersection (:404-424); the second test compares the
// hit's y with the rectangle's right x, as the reference does (:412).
__device__ __forceinline__ bool segment_meets_cell(float ax, float ay, float bx, float by, float lbx, float lby,
                                                   float rtx, float rty)
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
__device__ bool cell_candidate(const float* v, int n, float lbx, float lby, float rtx, float rty)
{
    for (int i = 0; i < n; i++) {
        const float x = v[2 * i], y = v[2 * i + 1];
        if (x >= lbx && x <= rtx && y >= lby && y <= rty) return true;
    }
    if (pip(v, n, lbx, lby) || pip(v, n, rtx, rty) || pip(v, n, lbx, rty) || pip(v, n, rtx, lby)) return true;
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

// Does the closed polygon meet the closed cell, grown by a relative margin of
// 1e-6 (double precision)? A cell that meets a polygon that is not among its
// candidates is deficient: the hybrid join sends its points to the exact test.
__device__ bool polygon_meets_rect(const float* v, int n, double lbx, double lby, double rtx, double rty)
{
    const double mx = 1e-6 * (fabs(rtx) + fabs(lbx) + (rtx - lbx));
    const double my = 1e-6 * (fabs(rty) + fabs(lby) + (rty - lby));
    lbx -= mx; rtx += mx; lby -= my; rty += my;
    for (int i = 0; i < n; i++) {
        const double x = v[2 * i], y = v[2 * i + 1];
        if (x >= lbx && x <= rtx && y >= lby && y <= rty) return true;
    }
    {  // the cell's center inside (the cell may lie in the polygon's interior)
        const double cx = 0.5 * (lbx + rtx), cy = 0.5 * (lby + rty);
        bool c = false;
        for (int i = 0, j = n - 1; i < n; j = i++) {
            const double xi = v[2 * i], yi = v[2 * i + 1], xj = v[2 * j], yj = v[2 * j + 1];
            if (((yi > cy) != (yj > cy)) && (cx < (xj - xi) * (cy - yi) / (yj - yi) + xi)) c = !c;
        }
        if (c) return true;
    }
    // any edge, the closing one included, crossing the cell (Liang-Barsky clip)
    for (int i = 0, j = n - 1; i < n; j = i++) {
        const double x0 = v[2 * j], y0 = v[2 * j + 1], dx = v[2 * i] - x0, dy = v[2 * i + 1] - y0;
        double t0 = 0, t1 = 1;
        const double pp[4] = {-dx, dx, -dy, dy};
        const double qq[4] = {x0 - lbx, rtx - x0, y0 - lby, rty - y0};
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!ok) break;
            if (pp[k] == 0) {
                if (qq[k] < 0) ok = false;
            } else {
                const double t = qq[k] / pp[k];
                if (pp[k] < 0) t0 = (t0 < t) ? t : t0;
                else t1 = (t < t1) ? t : t1;
                if (t0 > t1) ok = false;
            }
        }
        if (ok) return true;
    }
    return false;
}

struct GridArgs {
    const float* xy;         // collection vertices (CSR)
    const int32_t* off;
    int32_t direct;          // 1: every item tests each cell against every edge
    int32_t xs;
    float minx, miny, width, height;
    unsigned long long* keys;  // candidates: cell << 32 | polygon
    int32_t* defi;             // deficient cells
    unsigned long long* count; // [0] candidates, [1] deficient cells
    int64_t cap;               // room of keys and defi
};

// Append the lanes' flagged values: one atomic per wave.
template <class T>
__device__ __forceinline__ void wave_append(bool flag, T val, T* out, unsigned long long* counter, int64_t cap)
{
    const unsigned long long m = __ballot(flag);
    if (!m) return;
    const int lane = (int)__lane_id();
    const int lead = __ffsll((long long)m) - 1;
    unsigned long long base = 0;
    if (lane == lead) base = atomicAdd(counter, (unsigned long long)__popcll(m));
    base = __shfl(base, lead, 64);
    const unsigned long long below = lane ? (m & (~0ull >> (64 - lane))) : 0ull;
    const unsigned long long pos = base + (unsigned long long)__popcll(below);
    if (flag && (int64_t)pos < cap) out[pos] = val;
}

// Grid line c (column side) and r (row side) as GridIndexF::buildGrid
// computes cell bounds (:96-99).
__device__ __forceinline__ float grid_x(const GridArgs& a, int c) { return (float)c * a.width + a.minx; }
__device__ __forceinline__ float grid_y(const GridArgs& a, int r) { return (float)r * a.height + a.miny; }

// A cell side's bounds grown as polygon_meets_rect grows them.
__device__ __forceinline__ void grown(float l, float h, double& gl, double& gh)
{
    const double lo = l, hi = h;
    const double m = 1e-6 * (fabs(hi) + fabs(lo) + (hi - lo));
    gl = lo - m;
    gh = hi + m;
}

// The column (row) of a coordinate, from one division; within one of the
// true one on a regular item.
__device__ __forceinline__ int est_cell(float v, float mn, float step, int lo, int hi)
{
    const float f = floorf((v - mn) / step);
    if (!(f >= (float)lo)) return lo;  // NaN included
    if (f > (float)hi) return hi;
    return (int)f;
}

// The decomposed test needs the item's float grid lines at least 1e-5 of
// their magnitude apart (a coordinate's column estimate is then off by at
// most one, and a grown cell's margin stays below a tenth of a cell).
__device__ __forceinline__ bool item_regular(const GridArgs& a, int x0, int y0, int w, int h)
{
    const float mx = fmaxf(fabsf(grid_x(a, x0 - 4)), fabsf(grid_x(a, x0 + w + 4)));
    const float my = fmaxf(fabsf(grid_y(a, y0 - 4)), fabsf(grid_y(a, y0 + h + 4)));
    return a.width > 1e-5f * mx && a.height > 1e-5f * my;
}

__device__ __forceinline__ void set_bit(uint32_t* m, int r, int c) { atomicOr(&m[r * 2 + (c >> 5)], 1u << (c & 31)); }
__device__ __forceinline__ bool get_bit(const uint32_t* m, int r, int c) { return (m[r * 2 + (c >> 5)] >> (c & 31)) & 1u; }

__global__ __launch_bounds__(256) void k_grid_cells(GridArgs a, const GridItem* __restrict__ items)
{
    __shared__ float sv[2 * kGridLdsVerts];
    __shared__ uint32_t s_cand[kItemW * kItemH / 32];  // a vertex in the cell, or an edge (i, i+1) meets it
    __shared__ uint32_t s_meet[kItemW * kItemH / 32];  // a vertex in the grown cell, or an edge crosses it
    __shared__ uint32_t s_corner[(kItemH + 1) * 3];    // crossing parity of each grid point
    __shared__ float s_px[(kItemH + 1) * kPtRowCap];   // per grid-point row: its straddling edges' crossings
    __shared__ int32_t s_ce[kItemH * kCenRowCap];      // per cell-center row: its straddling edges
    __shared__ int32_t s_np[kItemH + 1], s_nc[kItemH];
    __shared__ int s_direct;
    const GridItem it = items[blockIdx.x];
    const int p = it.poly;
    const int v0 = a.off[p], n = a.off[p + 1] - v0;
    const float* v = a.xy + 2 * (int64_t)v0;
    const int x0 = it.x0, y0 = it.y0, w = it.wh & 0xFFFF, h = it.wh >> 16;
    if (n <= kGridLdsVerts) {
        for (int k = threadIdx.x; k < 2 * n; k += blockDim.x) sv[k] = v[k];
        v = sv;
    }
    for (int k = threadIdx.x; k < kItemW * kItemH / 32; k += blockDim.x) s_cand[k] = s_meet[k] = 0u;
    for (int k = threadIdx.x; k < (kItemH + 1) * 3; k += blockDim.x) s_corner[k] = 0u;
    for (int k = threadIdx.x; k <= kItemH; k += blockDim.x) {
        s_np[k] = 0;
        if (k < kItemH) s_nc[k] = 0;
    }
    if (threadIdx.x == 0) s_direct = a.direct || !item_regular(a, x0, y0, w, h);
    __syncthreads();
    const bool marks = !s_direct;
    if (marks) {
        const int cl = x0 - 4, ch = x0 + w + 3, rl = y0 - 4, rh = y0 + h + 3;
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            const int j = k == 0 ? n - 1 : k - 1;
            const float vix = v[2 * k], viy = v[2 * k + 1], vjx = v[2 * j], vjy = v[2 * j + 1];
            // vertex k in a cell (cell_candidate) or a grown cell (polygon_meets_rect)
            {
                const int ec = est_cell(vix, a.minx, a.width, cl, ch), er = est_cell(viy, a.miny, a.height, rl, rh);
                for (int dc = -2; dc <= 2; dc++) {
                    const int c = ec + dc;
                    if (c < x0 || c >= x0 + w) continue;
                    const float lx = grid_x(a, c), hx = grid_x(a, c + 1);
                    double glx, ghx;
                    grown(lx, hx, glx, ghx);
                    const bool inx = vix >= lx && vix <= hx, ingx = (double)vix >= glx && (double)vix <= ghx;
                    if (!inx && !ingx) continue;
                    for (int dr = -2; dr <= 2; dr++) {
                        const int r = er + dr;
                        if (r < y0 || r >= y0 + h) continue;
                        const float ly = grid_y(a, r), hy = grid_y(a, r + 1);
                        double gly, ghy;
                        grown(ly, hy, gly, ghy);
                        if (inx && viy >= ly && viy <= hy) set_bit(s_cand, r - y0, c - x0);
                        if (ingx && (double)viy >= gly && (double)viy <= ghy) set_bit(s_meet, r - y0, c - x0);
                    }
                }
            }
            // crossings of grid-point rows (the corner tests, GridIndexF::isInsidePoly)
            for (int r = 0; r <= h; r++) {
                const float ty = grid_y(a, y0 + r);
                if ((viy > ty) != (vjy > ty)) {
                    const float X = (vjx - vix) * (ty - viy) / (vjy - viy) + vix;
                    const int q = atomicAdd(&s_np[r], 1);
                    if (q < kPtRowCap) s_px[r * kPtRowCap + q] = X;
                    else s_direct = 1;
                }
            }
            // edges straddling cell-center rows (the grown cell's center test)
            for (int r = 0; r < h; r++) {
                double gly, ghy;
                grown(grid_y(a, y0 + r), grid_y(a, y0 + r + 1), gly, ghy);
                const double cy = 0.5 * (gly + ghy);
                if (((double)viy > cy) != ((double)vjy > cy)) {
                    const int q = atomicAdd(&s_nc[r], 1);
                    if (q < kCenRowCap) s_ce[r * kCenRowCap + q] = k;
                    else s_direct = 1;
                }
            }
            // edge j-k against the cells of its box: edge (k-1, k) of
            // cell_candidate (k >= 1; the closing edge is not tested there)
            // and, for every edge, the grown cells it crosses (Liang-Barsky)
            const int c0 = est_cell(fminf(vix, vjx), a.minx, a.width, cl, ch) - 2;
            const int c1 = est_cell(fmaxf(vix, vjx), a.minx, a.width, cl, ch) + 2;
            const int r0 = est_cell(fminf(viy, vjy), a.miny, a.height, rl, rh) - 2;
            const int r1 = est_cell(fmaxf(viy, vjy), a.miny, a.height, rl, rh) + 2;
            const double ex0 = vjx, ey0 = vjy, edx = (double)vix - ex0, edy = (double)viy - ey0;
            for (int c = max(c0, x0); c <= min(c1, x0 + w - 1); c++) {
                const float lx = grid_x(a, c), hx = grid_x(a, c + 1);
                double glx, ghx;
                grown(lx, hx, glx, ghx);
                for (int r = max(r0, y0); r <= min(r1, y0 + h - 1); r++) {
                    const float ly = grid_y(a, r), hy = grid_y(a, r + 1);
                    if (k >= 1) {
                        const int ca = region_code(vjx, vjy, lx, ly, hx, hy);
                        const int cb = region_code(vix, viy, lx, ly, hx, hy);
                        if (!(ca & cb) && segment_meets_cell(vjx, vjy, vix, viy, lx, ly, hx, hy))
                            set_bit(s_cand, r - y0, c - x0);
                    }
                    double gly, ghy;
                    grown(ly, hy, gly, ghy);
                    double t0 = 0, t1 = 1;
                    const double pp[4] = {-edx, edx, -edy, edy};
                    const double qq[4] = {ex0 - glx, ghx - ex0, ey0 - gly, ghy - ey0};
                    bool ok = true;
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        if (!ok) break;
                        if (pp[m] == 0) {
                            if (qq[m] < 0) ok = false;
                        } else {
                            const double t = qq[m] / pp[m];
                            if (pp[m] < 0) t0 = (t0 < t) ? t : t0;
                            else t1 = (t < t1) ? t : t1;
                            if (t0 > t1) ok = false;
                        }
                    }
                    if (ok) set_bit(s_meet, r - y0, c - x0);
                }
            }
        }
    }
    __syncthreads();
    const bool dec = !s_direct;
    if (dec) {
        // crossing parity of every grid point of the item
        for (int k = threadIdx.x; k < (w + 1) * (h + 1); k += blockDim.x) {
            const int r = k / (w + 1), c = k % (w + 1);
            const float tx = grid_x(a, x0 + c);
            const int m = s_np[r];
            bool in = false;
            for (int q = 0; q < m; q++) in ^= tx < s_px[r * kPtRowCap + q];
            if (in) atomicOr(&s_corner[r * 3 + (c >> 5)], 1u << (c & 31));
        }
        __syncthreads();
    }
    // every lane takes part in the appends: the loop runs to a multiple of the block
    const int cells = w * h;
    const int rounds = (cells + (int)blockDim.x - 1) / (int)blockDim.x;
    for (int rr = 0; rr < rounds; rr++) {
        const int i = rr * (int)blockDim.x + (int)threadIdx.x;
        bool cand = false, def = false;
        int cell = 0;
        if (i < cells) {
            const int r = i / w, c = i % w;
            const int x = x0 + c, y = y0 + r;
            cell = y * a.xs + x;
            const float lbx = grid_x(a, x), lby = grid_y(a, y), rtx = grid_x(a, x + 1), rty = grid_y(a, y + 1);
            if (dec) {
                auto corner = [&](int rr2, int cc2) { return (s_corner[rr2 * 3 + (cc2 >> 5)] >> (cc2 & 31)) & 1u; };
                cand = get_bit(s_cand, r, c) || corner(r, c) || corner(r + 1, c + 1) || corner(r + 1, c) ||
                       corner(r, c + 1);
                if (!cand) {
                    def = get_bit(s_meet, r, c);
                    if (!def) {
                        // the grown cell's center inside (polygon_meets_rect)
                        double glx, ghx, gly, ghy;
                        grown(lbx, rtx, glx, ghx);
                        grown(lby, rty, gly, ghy);
                        const double cx = 0.5 * (glx + ghx), cy = 0.5 * (gly + ghy);
                        const int m = s_nc[r];
                        bool in = false;
                        for (int q = 0; q < m; q++) {
                            const int e = s_ce[r * kCenRowCap + q];
                            const int ej = e == 0 ? n - 1 : e - 1;
                            const double xi = v[2 * e], yi = v[2 * e + 1], xj = v[2 * ej], yj = v[2 * ej + 1];
                            in ^= cx < (xj - xi) * (cy - yi) / (yj - yi) + xi;
                        }
                        def = in;
                    }
                }
            } else {
                cand = cell_candidate(v, n, lbx, lby, rtx, rty);
                if (!cand) def = polygon_meets_rect(v, n, lbx, lby, rtx, rty);
            }
        }
        wave_append<unsigned long long>(cand, ((unsigned long long)(uint32_t)cell << 32) | (uint32_t)p, a.keys,
                                        &a.count[0], a.cap);
        wave_append<int32_t>(def, cell, a.defi, &a.count[1], a.cap);
    }
}

// From the sorted keys: the candidate list and, for every cell c, the index
// of its first candidate (keys of cells >= c start there). Thread k writes the
// cells between the previous key's cell and its own; thread n the cells after
// the last key.
__global__ void k_grid_offsets(const unsigned long long* __restrict__ keys, int64_t n, int64_t ncells,
                               int32_t* __restrict__ cell_off, int32_t* __restrict__ cand)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n) return;
    const int64_t prev = k == 0 ? -1 : (int64_t)(keys[k - 1] >> 32);
    const int64_t cur = k < n ? (int64_t)(keys[k] >> 32) : ncells;
    if (k < n) cand[k] = (int32_t)(uint32_t)keys[k];
    for (int64_t c = prev + 1; c <= cur; c++) cell_off[c] = (int32_t)k;
}

// The polygon of vertex k: the last p with off[p] <= k.
__device__ __forceinline__ int vertex_poly(const int32_t* __restrict__ off, int np, int64_t k)
{
    int a = 0, b = np - 1;
    while (a < b) {
        const int m = (a + b + 1) >> 1;
        if (off[m] <= k) a = m;
        else b = m - 1;
    }
    return a;
}

// Band range of every edge (vertex k and its predecessor), as the host's
// band_row: the edge toggles the crossing test of a point at ty iff
// min(y) <= ty < max(y), whose bands run from band_row(min) to band_row(the
// largest float below max). Each polygon gets its first and last band: the
// lanes of a wave's first and last polygon combine before their atomics
// (consecutive vertices mostly share a polygon).
__global__ void k_edge_rows(const float* __restrict__ xy, const int32_t* __restrict__ prev,
                            const int32_t* __restrict__ off, int np, int64_t nv, float miny, float band_h, int rows,
                            int2* __restrict__ er, int32_t* __restrict__ lo, int32_t* __restrict__ hi,
                            unsigned long long* __restrict__ total)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long span = 0;
    int poly = -1, r0 = INT_MAX, r1 = INT_MIN;
    bool has = false;
    if (k < nv) {
        const float yi = xy[2 * k + 1], yj = xy[2 * (int64_t)prev[k] + 1];
        const float y0 = fminf(yi, yj), y1 = fmaxf(yi, yj);
        int b0 = 0, b1 = -1;
        if (y0 < y1) {
            b0 = band_row(y0, miny, band_h, rows);
            b1 = band_row(nextafterf(y1, -INFINITY), miny, band_h, rows);
        }
        er[k] = make_int2(b0, b1);
        poly = vertex_poly(off, np, k);
        if (b0 <= b1) {
            has = true;
            r0 = b0;
            r1 = b1;
            span = (unsigned long long)(b1 - b0 + 1);
        }
    }
    const int pa = __shfl(poly, 0, 64), pb = __shfl(poly, 63, 64);
    int la = (has && poly == pa) ? r0 : INT_MAX, ha = (has && poly == pa) ? r1 : INT_MIN;
    int lb = (has && poly == pb && pb != pa) ? r0 : INT_MAX, hb = (has && poly == pb && pb != pa) ? r1 : INT_MIN;
    for (int d = 32; d > 0; d >>= 1) {
        la = min(la, __shfl_xor(la, d, 64));
        ha = max(ha, __shfl_xor(ha, d, 64));
        lb = min(lb, __shfl_xor(lb, d, 64));
        hb = max(hb, __shfl_xor(hb, d, 64));
        span += __shfl_xor(span, d, 64);
    }
    const int lane = (int)__lane_id();
    if (lane == 0 && pa >= 0 && la <= ha) {
        atomicMin(&lo[pa], la);
        atomicMax(&hi[pa], ha);
    }
    if (lane == 63 && pb >= 0 && lb <= hb) {
        atomicMin(&lo[pb], lb);
        atomicMax(&hi[pb], hb);
    }
    if (has && poly != pa && poly != pb) {
        atomicMin(&lo[poly], r0);
        atomicMax(&hi[poly], r1);
    }
    if (lane == 0 && span) atomicAdd(total, span);
}

// One thread per edge (vertex k and its predecessor): +1 to the edge count of
// every band of its polygon its rows cover (FILL = false; counts at
// slot[band + 1]), or the edge as (vi, vj) into the next free entry of each
// such band (FILL = true; slot[band] is the band's cursor, from zero). The
// entries of a band are taken in no particular order: the crossing parity of
// the exact test does not depend on it.
template <bool FILL>
__global__ void k_band_edges(const float* __restrict__ xy, const int32_t* __restrict__ off, int np, int64_t nv,
                             const int32_t* __restrict__ prev, const int2* __restrict__ er,
                             const int32_t* __restrict__ blo, const int32_t* __restrict__ pband,
                             const int32_t* __restrict__ band_off, int32_t* __restrict__ slot,
                             float4* __restrict__ edges)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nv) return;
    const int2 r = er[k];
    if (r.x > r.y) return;
    const int p = vertex_poly(off, np, k);
    const int64_t base = (int64_t)pband[p] - blo[p];
    float4 e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (FILL) {
        const int j = prev[k];
        e = make_float4(xy[2 * k], xy[2 * k + 1], xy[2 * (int64_t)j], xy[2 * (int64_t)j + 1]);
    }
    for (int b = r.x; b <= r.y; b++) {
        if (FILL) {
            const int q = atomicAdd(&slot[base + b], 1);
            edges[band_off[base + b] + q] = e;
        } else {
            atomicAdd(&slot[base + b + 1], 1);
        }
    }
}

}  // namespace
}  // namespace rj

using namespace rj;

// CPUJoin grid of collection P at resolution res, built on the device into G
// (its device buffers, scalars and deficient cells).
int rj_ctx::grid_device(rj_polyset* P, int res, Grid* G)
{
    const PolySet& ps = P->ps;
    const int np = ps.np;
    GridIndex& g = G->g;
    g.xs = g.ys = res;
    g.minx = ps.bboxf[0];
    g.miny = ps.bboxf[1];
    g.width = (ps.bboxf[2] - ps.bboxf[0]) / (float)res;  // GridIndexF.cpp:66-67
    g.height = (ps.bboxf[3] - ps.bboxf[1]) / (float)res;
    g.deficient.clear();
    const int64_t ncells = (int64_t)res * res;
    int rc;

    // each polygon's box of cells (GridIndexF.cpp:82-91) and the work items
    std::vector<GridItem> items;
    int64_t total_cells = 0;
    for (int p = 0; p < np; p++) {
        const float* b = &ps.pbox[4 * (size_t)p];
        int stx = (int)((b[0] - g.minx) / g.width);
        int enx = (int)((b[2] - g.minx) / g.width) + 1;
        int sty = (int)((b[1] - g.miny) / g.height);
        int eny = (int)((b[3] - g.miny) / g.height) + 1;
        if (enx >= res) enx = res - 1;
        if (eny >= res) eny = res - 1;
        if (stx < 0) stx = 0;
        if (sty < 0) sty = 0;
        const int64_t cells = (enx >= stx && eny >= sty) ? (int64_t)(enx - stx + 1) * (eny - sty + 1) : 0;
        if (cells)
            for (int x = stx; x <= enx; x += kItemW)
                for (int y = sty; y <= eny; y += kItemH) {
                    const int w = std::min(kItemW, enx - x + 1), h = std::min(kItemH, eny - y + 1);
                    items.push_back({p, x, y, w | (h << 16)});
                }
        total_cells += cells;
    }
    if (total_cells >= ((int64_t)1 << 31) || ncells >= ((int64_t)1 << 31))
        return fail(RJ_ERR_LIMIT, "grid too fine for this collection");

    DevBuf d_items, d_keys, d_keys2, d_defi, d_count, d_tmp;
    const int64_t cap = std::max<int64_t>(1, total_cells);
    if ((rc = upload(d_items, items.data(), std::max<size_t>(1, items.size()) * sizeof(GridItem), "grid items")) ||
        (rc = alloc(d_keys, (size_t)cap * 8, "grid keys")) || (rc = alloc(d_defi, (size_t)cap * 4, "grid cells")) ||
        (rc = alloc(d_count, 16, "grid counters")))
        return rc;
    unsigned long long* cnt = d_count.as<unsigned long long>();
    if ((rc = launch("grid_clear", [&] { launch_fill_u32(stream, reinterpret_cast<uint32_t*>(cnt), 4, 0); })))
        return rc;
    GridArgs a{P->d_xy.as<float>(), P->d_off.as<int32_t>(), opt_grid_direct ? 1 : 0, res, g.minx, g.miny, g.width,
               g.height, d_keys.as<unsigned long long>(), d_defi.as<int32_t>(), cnt, cap};
    if (!items.empty() &&
        (rc = launch("grid_cells", [&] {
             hipLaunchKernelGGL(k_grid_cells, dim3((unsigned)items.size()), dim3(256), 0, stream, a, d_items.as<GridItem>());
         })))
        return rc;
    unsigned long long hc[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(hc, cnt, 16, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "grid cells");
    const int64_t nk = (int64_t)hc[0], nd = (int64_t)hc[1];
    if (nk > cap || nd > cap) return fail(RJ_ERR_STATE, "grid build overflow");
    G->ncand = nk;
    if (nd) {
        g.deficient.resize(nd);
        e = hipMemcpyAsync(g.deficient.data(), d_defi.p, nd * 4, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "grid deficient cells");
        std::sort(g.deficient.begin(), g.deficient.end());
        g.deficient.erase(std::unique(g.deficient.begin(), g.deficient.end()), g.deficient.end());
    }
    d_defi.release();  // temporaries go as soon as they are read: a tight budget holds the peak

    // candidates sorted by (cell, polygon): CPUJoin's lists
    if (nk == 0) {
        if ((rc = alloc(G->d_cell_off, (size_t)(ncells + 1) * 4, "grid cells")) ||
            (rc = alloc(G->d_cand, 4, "grid candidates")))
            return rc;
        int32_t* co = G->d_cell_off.as<int32_t>();
        if ((rc = launch("grid_offsets", [&] { launch_fill_u32(stream, (uint32_t*)co, ncells + 1, 0); }))) return rc;
    } else {
        int cell_bits = 1;
        while (((int64_t)1 << cell_bits) < ncells) cell_bits++;
        size_t tmp = 0;
        unsigned long long* k0 = d_keys.as<unsigned long long>();
        if ((rc = alloc(d_keys2, (size_t)nk * 8, "grid keys"))) return rc;
        unsigned long long* k1 = d_keys2.as<unsigned long long>();
        (void)hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, k0, k1, (int)nk, 0, 32 + cell_bits, stream);
        if ((rc = alloc(d_tmp, tmp + 16, "grid sort"))) return rc;
        e = hipcub::DeviceRadixSort::SortKeys(d_tmp.p, tmp, k0, k1, (int)nk, 0, 32 + cell_bits, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "grid sort");
        d_tmp.release();
        d_keys.release();
        if ((rc = alloc(G->d_cell_off, (size_t)(ncells + 1) * 4, "grid cells")) ||
            (rc = alloc(G->d_cand, (size_t)nk * 4, "grid candidates")))
            return rc;
        int32_t* co = G->d_cell_off.as<int32_t>();
        int32_t* cd = G->d_cand.as<int32_t>();
        if ((rc = launch("grid_offsets", [&] {
                 hipLaunchKernelGGL(k_grid_offsets, dim3((unsigned)((nk + 1 + 255) / 256)), dim3(256), 0, stream, k1,
                                    nk, ncells, co, cd);
             })))
            return rc;
    }
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "grid offsets");
    d_keys.release();
    d_keys2.release();

    // edge bands: the grid rows, split k ways for about 4 edges per band
    const int64_t nv = ps.off[np];
    DevBuf d_er, d_lohi, d_scan, d_bcnt;
    if ((rc = alloc(d_er, (size_t)std::max<int64_t>(1, nv) * 8, "edge rows")) ||
        (rc = alloc(d_lohi, (size_t)np * 8 + 16, "band ranges")))
        return rc;
    int2* er = d_er.as<int2>();
    int32_t* dlo = d_lohi.as<int32_t>();
    int32_t* dhi = dlo + np;
    std::vector<int32_t> lo(np), hi(np);
    auto edge_rows = [&](float band_h, int rows, unsigned long long* total) -> int {
        int r2;
        if ((r2 = launch("band_init", [&] {
                 launch_fill_u32(stream, (uint32_t*)dlo, np, (uint32_t)INT_MAX);
                 launch_fill_u32(stream, (uint32_t*)dhi, np, 0xFFFFFFFFu);
                 launch_fill_u32(stream, (uint32_t*)cnt, 4, 0);
             })))
            return r2;
        const int32_t* prev = P->d_prev.as<int32_t>();
        if ((r2 = launch("edge_rows", [&] {
                 hipLaunchKernelGGL(k_edge_rows, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, stream,
                                    P->d_xy.as<float>(), prev, P->d_off.as<int32_t>(), np, nv, g.miny, band_h, rows,
                                    er, dlo, dhi, cnt);
             })))
            return r2;
        hipError_t e2 = hipMemcpyAsync(lo.data(), dlo, (size_t)np * 4, hipMemcpyDeviceToHost, stream);
        if (e2 == hipSuccess) e2 = hipMemcpyAsync(hi.data(), dhi, (size_t)np * 4, hipMemcpyDeviceToHost, stream);
        if (e2 == hipSuccess) e2 = hipMemcpyAsync(total, cnt, 8, hipMemcpyDeviceToHost, stream);
        if (e2 == hipSuccess) e2 = hipStreamSynchronize(stream);
        if (e2 != hipSuccess) return hip_fail(e2, "edge rows");
        // polygons without a band-crossing edge: no bands (lo 0, hi -1)
        for (int p = 0; p < np; p++)
            if (hi[p] < lo[p]) {
                lo[p] = 0;
                hi[p] = -1;
            }
        return RJ_OK;
    };
    unsigned long long entries = 0;
    if ((rc = edge_rows(g.height, res, &entries))) return rc;
    int64_t nb = 0;
    for (int p = 0; p < np; p++) nb += hi[p] - lo[p] + 1;
    const double per = (double)entries / (double)std::max<int64_t>(1, nb);
    int ksplit = 1;
    while (ksplit < 64 && per / ksplit > 4.0) ksplit *= 2;
    g.band_h = g.height;
    g.band_rows = res;
    if (ksplit > 1) {
        g.band_h = g.height / (float)ksplit;
        g.band_rows = res * ksplit;
        if ((rc = edge_rows(g.band_h, g.band_rows, &entries))) return rc;
    }
    // band_lo | band_hi | pband (np + 1) | band_off (nbands + 1)
    std::vector<int32_t> pband(np + 1, 0);
    for (int p = 0; p < np; p++) pband[p + 1] = pband[p] + (hi[p] - lo[p] + 1);
    const int64_t nbands = pband[np];
    G->np = np;
    if ((rc = alloc(G->d_bands, ((size_t)3 * np + 1 + nbands + 1) * 4, "grid row bands"))) return rc;
    int32_t* bands = G->d_bands.as<int32_t>();
    e = hipMemcpyAsync(bands, lo.data(), (size_t)np * 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(bands + np, hi.data(), (size_t)np * 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(bands + 2 * np, pband.data(), (size_t)(np + 1) * 4, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "grid row bands");
    int32_t* band_off = bands + 3 * np + 1;
    const int32_t* blo = bands;
    const int32_t* pb = bands + 2 * np;
    const unsigned nvk = (unsigned)((std::max<int64_t>(1, nv) + 255) / 256);
    if ((rc = alloc(d_bcnt, (size_t)(std::max<int64_t>(1, nbands) + 1) * 4, "band counts"))) return rc;
    int32_t* bcnt = d_bcnt.as<int32_t>();  // bcnt[b + 1]: edges of band b; then band b's cursor at bcnt[b]
    if ((rc = launch("band_count", [&] {
             launch_fill_u32(stream, (uint32_t*)band_off, 1, 0);
             launch_fill_u32(stream, (uint32_t*)bcnt, std::max<int64_t>(1, nbands) + 1, 0);
             if (nbands)
                 hipLaunchKernelGGL(k_band_edges<false>, dim3(nvk), dim3(256), 0, stream, P->d_xy.as<float>(),
                                    P->d_off.as<int32_t>(), np, nv, P->d_prev.as<int32_t>(), er, blo, pb, band_off,
                                    bcnt, nullptr);
         })))
        return rc;
    int32_t nedges = 0;
    if (nbands) {
        size_t tmp = 0;
        (void)hipcub::DeviceScan::InclusiveSum(nullptr, tmp, bcnt + 1, band_off + 1, (int)nbands, stream);
        if ((rc = alloc(d_scan, tmp + 16, "band scan"))) return rc;
        e = hipcub::DeviceScan::InclusiveSum(d_scan.p, tmp, bcnt + 1, band_off + 1, (int)nbands, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&nedges, band_off + nbands, 4, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, "band scan");
    }
    if ((rc = alloc(G->d_edges, (size_t)std::max<int32_t>(1, nedges) * 16, "grid band edges"))) return rc;
    float4* ed = G->d_edges.as<float4>();
    if (nbands && (rc = launch("band_fill", [&] {
                       launch_fill_u32(stream, (uint32_t*)bcnt, nbands, 0);
                       hipLaunchKernelGGL(k_band_edges<true>, dim3(nvk), dim3(256), 0, stream, P->d_xy.as<float>(),
                                          P->d_off.as<int32_t>(), np, nv, P->d_prev.as<int32_t>(), er, blo, pb,
                                          band_off, bcnt, ed);
                   })))
        return rc;
    // the temporaries above are freed when this returns: wait for the kernels
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "grid build");
    return RJ_OK;
}
