// kernels.hip -- gfx950 kernels of the spatial aggregation join.
//
//  k_band      scan conversion of one band of rows of one polygon with the
//              edge-flag rule on the fixed-point canvas grid; per mode it sums
//              the canvas under the polygon (polygon.frag), writes the polygon
//              label map, or accumulates the error bounds (line-bound.frag).
//  k_touch     boundary pixels of every edge (hybrid border mask, the
//              conservative line raster of HybridJoin::drawOutline).
//  k_points    one pass over the points: CPUJoin grid + exact float PIP
//              (index.glsl), label-map gather (raster), label gather with the
//              exact test on border pixels (hybrid.glsl), or the splat into
//              the count/sum canvas (RasterJoin.cpp:91-148, points.vert/.frag).
//
// Compiled with -ffp-contract=off: the float PIP, the grid cell and the pixel
// formula must round exactly like the host restatement of CPUJoin.
#include <algorithm>

#include "point_dev.hpp"
#include "scan_ops.hpp"

namespace rj {
namespace {

#ifndef RJ_POINT_ROWS
#define RJ_POINT_ROWS 8
#endif
constexpr int kRows = RJ_POINT_ROWS;  // consecutive rows per thread and iteration of a point pass
static_assert(kRows == 4 || kRows == 8 || kRows == 16, "rows per thread: 4, 8 or 16");


// ---------------------------------------------------------------------------
// fill, snap
// ---------------------------------------------------------------------------

// Vertices to the fixed-point canvas grid: rint(((double)v - min) * scale),
// round half to even, the same IEEE double operations as the host.
__global__ void k_snap(const float* __restrict__ xy, int64_t nv, double minx, double miny, double fxs,
                       double fys, int32_t* __restrict__ X, int32_t* __restrict__ Y)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nv) return;
    X[k] = (int32_t)rint(((double)xy[2 * k] - minx) * fxs);
    Y[k] = (int32_t)rint(((double)xy[2 * k + 1] - miny) * fys);
}

// The edge list of a collection: the predecessor of every vertex in its ring.
__global__ void k_prev(const int32_t* __restrict__ off, int np, int64_t nv, int32_t* __restrict__ prev)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nv) return;
    int a = 0, b = np - 1;  // the polygon of vertex k: the last p with off[p] <= k
    while (a < b) {
        const int m = (a + b + 1) >> 1;
        if (off[m] <= k) a = m;
        else b = m - 1;
    }
    prev[k] = (k == off[a]) ? off[a + 1] - 1 : (int32_t)(k - 1);
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// Up to two ranges of 64-bit words copied by one kernel; either side may be
// pinned host memory (its device address), so small uploads and readbacks
// run on the compute queue in stream order instead of as copy-engine
// transfers that the next kernel has to wait for.
__global__ void k_copy_words(CopyWords r0, CopyWords r1)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < r0.n + r1.n; i += stride) {
        if (i < r0.n) r0.dst[i] = r0.src[i];
        else r1.dst[i - r0.n] = r1.src[i - r0.n];
    }
}

// ---------------------------------------------------------------------------
// scan conversion of a band of one polygon
// ---------------------------------------------------------------------------

// Edge-flag toggles of the centers of cells of size cs (subunits) for rows
// [R0, R0+NR) and columns from C0 (BW columns, WPR words per row): each
// crossing toggles the first column at or right of it.
__device__ __forceinline__ void edge_toggles(int64_t ax, int64_t ay, int64_t bx, int64_t by,
                                             int64_t cs, int R0, int NR, int C0, int BW, int WPR,
                                             uint32_t* buf)
{
    edge_crossings(ax, ay, bx, by, cs, R0, NR, [&](int64_t row, int64_t k) {
        int64_t local = k - C0;
        if (local < 0) local = 0;
        if (local >= BW) return;
        atomicXor(&buf[(row - R0) * WPR + (local >> 5)], 1u << (local & 31));
    });
}

// Pixels whose closed square meets segment ab, rows [TR0, TR0+TNR) of the
// mask, columns [C0, C0+BW) (WPR words per row); canvas W x H.
__device__ __forceinline__ void edge_touch(int64_t ax, int64_t ay, int64_t bx, int64_t by, int TR0,
                                           int TNR, int C0, int BW, int WPR, int W, int H,
                                           uint32_t* buf)
{
    edge_touch_rows(ax, ay, bx, by, TR0, TNR, H, [&](int64_t j, int64_t xlo, int64_t xhi) {
        int64_t ilo = ceil_div(xlo, kSub) - 1, ihi = floor_div(xhi, kSub);
        if (ilo < 0) ilo = 0;
        if (ihi > W - 1) ihi = W - 1;
        if (ilo < C0) ilo = C0;
        if (ihi > C0 + BW - 1) ihi = C0 + BW - 1;
        for (int64_t i = ilo; i <= ihi; i++) {
            int64_t l = i - C0;
            atomicOr(&buf[(j - TR0) * WPR + (l >> 5)], 1u << (l & 31));
        }
    });
}

__device__ __forceinline__ void prefix_xor_row(uint32_t* w, int nw)
{
    uint32_t carry = 0;
    for (int k = 0; k < nw; k++) {
        uint32_t x = w[k];
        x ^= x << 1;
        x ^= x << 2;
        x ^= x << 4;
        x ^= x << 8;
        x ^= x << 16;
        x ^= carry;
        w[k] = x;
        carry = (x >> 31) ? 0xFFFFFFFFu : 0u;
    }
}

// Inclusive prefix sums (wrapping) along nr rows of bw words, row stride
// `stride`, in LDS: a power-of-two group of threads per row, each thread a
// contiguous chunk; tot holds blockDim.x chunk totals. Every thread calls it.
__device__ __forceinline__ void prefix_sum_rows(uint32_t* a, int nr, int bw, int stride, uint32_t* tot)
{
    const int tpr = 1 << (31 - __clz((int)blockDim.x / nr));
    const int row = threadIdx.x / tpr, part = threadIdx.x % tpr;
    const int chunk = (bw + tpr - 1) / tpr;
    const int b = min(part * chunk, bw), e = min(b + chunk, bw);
    uint32_t* ar = a + row * stride;
    uint32_t s = 0;
    if (row < nr)
        for (int c = b; c < e; c++) s += ar[c];
    tot[threadIdx.x] = s;
    __syncthreads();
    if (row < nr) {
        uint32_t run = 0;
        for (int q = row * tpr; q < (int)threadIdx.x; q++) run += tot[q];
        for (int c = b; c < e; c++) {
            run += ar[c];
            ar[c] = run;
        }
    }
}

// A row of bits grown by one column each way (bit k set when k-1, k or k+1
// was), in place; bits past the row's last word stay clear.
__device__ __forceinline__ void hdilate_row(uint32_t* w, int nw)
{
    uint32_t prev = 0u, cur = w[0];
    for (int k = 0; k < nw; k++) {
        const uint32_t next = k + 1 < nw ? w[k + 1] : 0u;
        w[k] = cur | (cur << 1) | (prev >> 31) | (cur >> 1) | (next << 31);
        prev = cur;
        cur = next;
    }
}

__device__ __forceinline__ int bit_at(const uint32_t* row, int col)
{
    return (row[col >> 5] >> (col & 31)) & 1;
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_band(const Band* __restrict__ bands, DevPolys P, DevCanvas cv,
                                                 const uint32_t* __restrict__ cnt,
                                                 const unsigned long long* __restrict__ chi,
                                                 const unsigned long long* __restrict__ clo, DevResult res,
                                                 unsigned long long* __restrict__ bounds, void* label,
                                                 int label16, unsigned long long* aux, int opt)
{
    extern __shared__ uint32_t lds[];
    const Band bd = bands[blockIdx.x];
    const int p = bd.poly;
    const int c0 = bd.c0, c1 = bd.c1;
    const int r0 = bd.r0, nr = bd.nrows, wpr = bd.wpr;
    const int bw = c1 - c0 + 1;
    uint32_t* tog = lds;
    uint32_t* tch = tog + nr * wpr;
    const int astride = 32 * wpr;
    uint32_t* area = tch + (nr + 2) * wpr;  // BAND_BOUNDS: nr rows of per-pixel area sums
    uint32_t* scan = area + nr * astride;    // BAND_BOUNDS: kBandScan chunk totals
    const int nw = (MODE == BAND_BOUNDS) ? (nr * wpr + (nr + 2) * wpr + nr * astride) : nr * wpr;
    for (int k = threadIdx.x; k < nw; k += blockDim.x) lds[k] = 0;
    __syncthreads();

    const int e0 = P.off[p], e1 = P.off[p + 1], n = e1 - e0;
    const int64_t band_ylo = (int64_t)(r0 - 1) * kSub, band_yhi = (int64_t)(r0 + nr + 1) * kSub;
    if constexpr (MODE == BAND_BOUNDS) {
        // (edge, row) items: the edges that meet the band are compacted in
        // LDS and every one of their rows r0-1 .. r0+nr goes to a lane of its
        // own (a band holds few edges; one lane per edge ran their rows one
        // after the other). The list and its count use the scan totals'
        // words, free until the prefix sums.
        constexpr int kChunk = kBandScan - 1;
        int* s_in = reinterpret_cast<int*>(scan);
        int* s_nin = s_in + kChunk;
        const int nrows = nr + 2;
        for (int base = 0; base < n; base += kChunk) {
            if (threadIdx.x == 0) *s_nin = 0;
            __syncthreads();
            const int e = base + (int)threadIdx.x;
            if ((int)threadIdx.x < kChunk && e < n) {
                const int i = e0 + e;
                const int j = (e == 0) ? e1 - 1 : i - 1;
                const int64_t ay = P.Y[i], by = P.Y[j];
                if (!(max(ay, by) < band_ylo || min(ay, by) > band_yhi)) s_in[atomicAdd(s_nin, 1)] = e;
            }
            __syncthreads();
            const int nitems = *s_nin * nrows;
            for (int it = threadIdx.x; it < nitems; it += blockDim.x) {
                const int ei = s_in[it / nrows], row = r0 - 1 + it % nrows;
                const int i = e0 + ei;
                const int j = (ei == 0) ? e1 - 1 : i - 1;
                const int64_t ax = P.X[i], ay = P.Y[i], bx = P.X[j], by = P.Y[j];
                if (row >= r0 && row < r0 + nr) {
                    edge_toggles(ax, ay, bx, by, kSub, row, 1, c0, bw, wpr, tog + (row - r0) * wpr);
                    edge_area(ax, ay, bx, by, row, 1, [&](int64_t rw, int64_t c, uint32_t d) {
                        int64_t l = c - c0;
                        if (l < 0) l = 0;  // everything left of the band enters at its first column
                        if (l >= bw) return;
                        atomicAdd(&area[(rw - r0) * astride + l], d);
                    });
                }
                edge_touch(ax, ay, bx, by, row, 1, c0, bw, wpr, cv.W, cv.H, tch + (row - (r0 - 1)) * wpr);
            }
            __syncthreads();  // s_in takes the next edges
        }
    } else {
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int i = e0 + e;
            const int j = (e == 0) ? e1 - 1 : i - 1;
            const int64_t ax = P.X[i], ay = P.Y[i], bx = P.X[j], by = P.Y[j];
            if (max(ay, by) < band_ylo || min(ay, by) > band_yhi) continue;
            edge_toggles(ax, ay, bx, by, kSub, r0, nr, c0, bw, wpr, tog);
        }
    }
    __syncthreads();
    for (int r = threadIdx.x; r < nr; r += blockDim.x) prefix_xor_row(tog + r * wpr, wpr);
    if (MODE == BAND_BOUNDS) {
        // touched rows grown by one column; the rows above and below join at use
        for (int r = threadIdx.x; r < nr + 2; r += blockDim.x) hdilate_row(tch + r * wpr, wpr);
        prefix_sum_rows(area, nr, bw, astride, scan);
    }
    __syncthreads();

    unsigned long long s_cnt = 0, s_hi = 0, s_lo = 0;
    unsigned long long s_fn = 0, s_fp = 0, s_fn1 = 0, s_fp1 = 0;
    if (MODE == BAND_LABEL && label16) {
        // 16-bit label map: aligned groups of 8 pixels, one 16-byte store
        // where the polygon covers all 8 (rows of the map are 64-byte aligned)
        const int g0 = c0 >> 3, ng = (c1 >> 3) - g0 + 1;
        const uint32_t pp = (uint32_t)p | ((uint32_t)p << 16);
        for (int idx = threadIdx.x; idx < nr * ng; idx += blockDim.x) {
            const int row = idx / ng, g = g0 + idx % ng;
            const uint32_t m = bits8(tog + row * wpr, 8 * g - c0, bw, wpr);
            if (!m) continue;
            uint16_t* dst = static_cast<uint16_t*>(label) + (int64_t)(r0 + row) * cv.LW + 8 * g;
            if (m == 0xFFu) {
                *reinterpret_cast<uint4*>(dst) = make_uint4(pp, pp, pp, pp);
            } else {
#pragma unroll
                for (int k = 0; k < 8; k++)
                    if ((m >> k) & 1u) dst[k] = (uint16_t)p;
            }
            s_cnt += __popc(m);
        }
        block_add4(s_cnt, 0, 0, 0, aux, nullptr, nullptr, nullptr);
        return;
    }
    if (MODE == BAND_BOUNDS && (opt & kBandNoCounts)) {
        // only boundary pixels count: a thread per 32-pixel word of a row,
        // over the word's boundary bits alone
        for (int idx = threadIdx.x; idx < nr * wpr; idx += blockDim.x) {
            const int row = idx / wpr, w = idx - row * wpr;
            uint32_t d = tch[row * wpr + w] | tch[(row + 1) * wpr + w] | tch[(row + 2) * wpr + w];
            if (w == wpr - 1 && (bw & 31)) d &= (1u << (bw & 31)) - 1u;  // columns past the band
            const uint32_t in_w = tog[row * wpr + w];
            const int64_t prow = (int64_t)(r0 + row) * cv.W + c0;
            while (d) {
                const int b = __ffs(d) - 1;
                d &= d - 1;
                const int col = 32 * w + b;
                const uint32_t val = cnt[prow + col];
                if (val == 0) continue;
                const uint32_t a = area_share(area[row * astride + col]);
                const uint32_t m = min(a, (uint32_t)kAreaOne - a);
                const unsigned long long add =
                    ((unsigned long long)m * val + (unsigned long long)(kAreaOne - 1)) >> 24;
                if ((in_w >> b) & 1u) {
                    s_fp += add;
                    s_fp1 += val;
                } else {
                    s_fn += add;
                    s_fn1 += val;
                }
            }
        }
        block_add4(s_fn, s_fp, s_fn1, s_fp1, bounds + p, bounds + P.np + p, bounds + 2 * P.np + p,
                   bounds + 3 * P.np + p);
        return;
    }
    // Threads cover 2^lg columns of (256 >> lg) rows at a time: no division.
    const int lg = bw >= kBlock ? 8 : (32 - __clz(bw - 1));
    const int colstep = 1 << lg;
    for (int row = threadIdx.x >> lg; row < nr; row += (kBlock >> lg)) {
        const uint32_t* trow = tog + row * wpr;
        const int64_t prow = (int64_t)(r0 + row) * cv.W + c0;
        const int64_t lrow = (int64_t)(r0 + row) * cv.LW + c0;
        for (int col = threadIdx.x & (colstep - 1); col < bw; col += colstep) {
            const int inside = bit_at(trow, col);
            const int64_t pix = prow + col;
            if (MODE == BAND_SUM) {
                if (inside) {
                    s_cnt += cnt[pix];
                    if (opt & kBandSums) {
                        s_hi += chi[pix];
                        s_lo += clo[pix];
                    }
                }
            } else if (MODE == BAND_LABEL) {
                if (inside) {
                    if (label16) static_cast<uint16_t*>(label)[lrow + col] = (uint16_t)p;
                    else static_cast<uint32_t*>(label)[lrow + col] = (uint32_t)p;
                    s_cnt++;
                }
            } else if (MODE == BAND_LABEL_ATOMIC) {
                if (inside) {
                    uint32_t* l32 = static_cast<uint32_t*>(label);
                    const uint32_t old = atomicCAS(&l32[lrow + col], kLabelNone, (uint32_t)p);
                    if (old != kLabelNone) atomicExch(&l32[lrow + col], kLabelBorder);
                }
            } else {
                // boundary pixel: a touched pixel within one row and column
                const int w = col >> 5;
                const uint32_t d = tch[row * wpr + w] | tch[(row + 1) * wpr + w] | tch[(row + 2) * wpr + w];
                const bool bd = (d >> (col & 31)) & 1u;
                // with the counts from a label pass only boundary pixels are read
                const uint32_t val = (bd || (inside && !(opt & kBandNoCounts))) ? cnt[pix] : 0u;
                if (inside) s_cnt += val;
                if (bd && val != 0) {
                    // min(share, 1 - share) of the pixel's area times its count, rounded up
                    const uint32_t a = area_share(area[row * astride + col]);
                    const uint32_t m = min(a, (uint32_t)kAreaOne - a);
                    const unsigned long long add =
                        ((unsigned long long)m * val + (unsigned long long)(kAreaOne - 1)) >> 24;
                    if (inside) {
                        s_fp += add;
                        s_fp1 += val;
                    } else {
                        s_fn += add;
                        s_fn1 += val;
                    }
                }
            }
        }
    }
    if (MODE == BAND_SUM) {
        block_add4(s_cnt, s_hi, s_lo, 0, res.cnt + p, (unsigned long long*)res.hi + p,
                   res.lo + p, nullptr);
    } else if (MODE == BAND_LABEL) {
        // total coverage; equals the labeled pixel count iff no pixel is in two polygons
        block_add4(s_cnt, 0, 0, 0, aux, nullptr, nullptr, nullptr);
    } else if (MODE == BAND_BOUNDS) {
        if (!(opt & kBandNoCounts)) {
            block_add4(s_cnt, 0, 0, 0, res.cnt + p, nullptr, nullptr, nullptr);
            __syncthreads();
        }
        block_add4(s_fn, s_fp, s_fn1, s_fp1, bounds + p, bounds + P.np + p, bounds + 2 * P.np + p,
                   bounds + 3 * P.np + p);
    }
}

// ---------------------------------------------------------------------------
// border mask for the hybrid join
// ---------------------------------------------------------------------------

// One thread per polygon edge (vertex i and its predecessor prev[i]).
__global__ void k_touch(const int32_t* __restrict__ X, const int32_t* __restrict__ Y,
                        const int32_t* __restrict__ prev, int64_t nverts, int W, int H, int wpr,
                        uint32_t* mask)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nverts) return;
    int64_t j = prev[i];
    edge_touch(X[i], Y[i], X[j], Y[j], 0, H, 0, W, wpr, W, H, mask);
}

// Inclusive pixel rectangles (i0, i1, j0, j1) set in the mask.
__global__ void k_mark_rects(const int32_t* __restrict__ rects, int nrects, int W, int H, int wpr,
                             uint32_t* mask)
{
    int r = blockIdx.x;
    if (r >= nrects) return;
    int i0 = max(0, rects[4 * r]), i1 = min(W - 1, rects[4 * r + 1]);
    int j0 = max(0, rects[4 * r + 2]), j1 = min(H - 1, rects[4 * r + 3]);
    int w = i1 - i0 + 1;
    if (w <= 0 || j1 < j0) return;
    for (int t = threadIdx.x; t < w * (j1 - j0 + 1); t += blockDim.x) {
        int j = j0 + t / w, i = i0 + t % w;
        atomicOr(&mask[(int64_t)j * wpr + (i >> 5)], 1u << (i & 31));
    }
}

// Dilate the mask by r pixels (Chebyshev) and mark the result in the labels.
__global__ void k_dilate_to_label(const uint32_t* __restrict__ mask, int W, int H, int LW, int wpr, int r,
                                  void* label, int label16)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)wpr * H) return;
    const int row = (int)(t / wpr), w = (int)(t % wpr);
    uint32_t m = 0, ml = 0, mr = 0;
    for (int dr = -r; dr <= r; dr++) {
        int rr = row + dr;
        if (rr < 0 || rr >= H) continue;
        const uint32_t* mrow = mask + (int64_t)rr * wpr;
        m |= mrow[w];
        if (w > 0) ml |= mrow[w - 1];
        if (w + 1 < wpr) mr |= mrow[w + 1];
    }
    uint32_t d = m;
    for (int s = 1; s <= r; s++) {
        d |= (m << s) | (ml >> (32 - s));
        d |= (m >> s) | (mr << (32 - s));
    }
    if (!d) return;
    const int64_t rowpix = (int64_t)row * LW;
    for (int b = 0; b < 32; b++) {
        const int col = w * 32 + b;
        if (col >= W) break;
        if ((d >> b) & 1u) {
            if (label16) static_cast<uint16_t*>(label)[rowpix + col] = (uint16_t)kLabelBorder;
            else static_cast<uint32_t*>(label)[rowpix + col] = kLabelBorder;
        }
    }
}

// Tile map of the label map: per tile of 2^TS x 2^TS pixels its single label,
// or kLabelMixed, plus the count of labeled pixels (the overlap test). Threads
// read 16-byte vectors (8 or 4 labels) down a strip of 2^TS rows and the
// threads of one tile combine with shuffles (rows of the label map are padded
// to a multiple of 32 pixels, so a tile's threads are all active or all
// idle). The map is small enough to stay in L2 during the point pass.
template <int TS>
__global__ void k_tile_labels(const void* __restrict__ label, int label16, int LW, int H, int TW, int TH,
                              uint32_t* __restrict__ tiles, unsigned long long* __restrict__ labeled)
{
    constexpr int T = 1 << TS;
    constexpr int KT = TS == 3 ? 4 : 1;  // tile rows per workgroup (8-px tiles: 1, 2 and 8 measured slower)
    const int per = label16 ? 8 : 4;  // labels per vector
    const int tpt = T / per;          // threads per tile
    const int nvec = LW / per;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = v < nvec;
    unsigned long long nlab = 0;
    for (int kt = 0; kt < KT; kt++) {
        const int ty = blockIdx.y * KT + kt;
        if (ty >= TH) break;
        uint32_t first = 0;
        int uni = 1;
        if (active) {
            const int y0 = ty << TS;
            const int rows = min(T, H - y0);
            const size_t row_bytes = (size_t)LW * (label16 ? 2 : 4);
            const char* base = static_cast<const char*>(label) + (size_t)y0 * row_bytes;
            const uint4 w0 = reinterpret_cast<const uint4*>(base)[v];
            const uint32_t a0 = w0.x & 0xFFFFu;
            first = label16 ? (a0 >= 0xFFF0u ? (0xFFFF0000u | a0) : a0) : w0.x;
#pragma unroll 8
            for (int k = 0; k < T; k++) {
                if (k < rows) {
                    const uint4 w = reinterpret_cast<const uint4*>(base + (size_t)k * row_bytes)[v];
                    const uint32_t x[4] = {w.x, w.y, w.z, w.w};
                    if (label16) {
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const uint32_t a = x[j] & 0xFFFFu, b = x[j] >> 16;
                            const uint32_t va = a >= 0xFFF0u ? (0xFFFF0000u | a) : a;
                            const uint32_t vb = b >= 0xFFF0u ? (0xFFFF0000u | b) : b;
                            uni &= (va == first) & (vb == first);
                            nlab += (va != kLabelNone) + (vb != kLabelNone);
                        }
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            uni &= (x[j] == first);
                            nlab += (x[j] != kLabelNone);
                        }
                    }
                }
            }
        }
        for (int d = 1; d < tpt; d <<= 1) {
            const uint32_t of = __shfl_xor(first, d, 64);
            const int ou = __shfl_xor(uni, d, 64);
            uni = uni && ou && (of == first);
        }
        if (active && (v % tpt) == 0) tiles[(size_t)ty * TW + v / tpt] = uni ? first : kLabelMixed;
    }
    if (labeled) block_add4(nlab, 0, 0, 0, labeled, nullptr, nullptr, nullptr);
}

// ---------------------------------------------------------------------------
// point pass: index join, label-map raster join, hybrid join, canvas splat
// ---------------------------------------------------------------------------

// The hybrid pass with the border queue over many rows takes 4 rows per
// thread (its per-row state at kRows costs occupancy); short passes keep
// kRows (fewer iterations of dependent loads per workgroup).
constexpr int64_t kQueuePassRows4 = 1 << 24;

// Rows per thread of the two-stage passes (rows_compact).
#ifndef RJ_COMPACT_ROWS
#define RJ_COMPACT_ROWS 8
#endif
constexpr int kCompactRows = RJ_COMPACT_ROWS;

// One pass over rows [begin, end): each workgroup takes a contiguous block of
// rows (point_rows) and sums per polygon in LDS (acc_init / acc_flush).
// CMP: the two-stage rows of a query with constraints (rows_compact).
template <int MODE, bool AVG, int R, bool CMP = false>
__global__ __launch_bounds__(kBlock) void k_points(PointArgs a)
{
    extern __shared__ unsigned long long sacc[];
    Acc A;
    if (MODE == PT_SPLAT) {
        A.cnt = A.hi = A.lo = nullptr;
        A.tag = nullptr;
        A.g = a.res;
        A.lds_np = A.hash = 0;
        splat_init(A, sacc, AVG);
    } else {
        acc_init(A, sacc, a.P.np, a.lds_np, AVG, a.res);
    }
    const int64_t b0 = a.start + (int64_t)blockIdx.x * a.per_block;
    const int64_t b1 = min(a.end, b0 + a.per_block);
    Run run{-1, 0, 0, 0};
    point_rows<MODE, AVG, R, true, CMP>(a, b0, max(a.begin, b0), b1, A, run);
    if (MODE == PT_HYBRID_QUEUE) border_finish(a, run);
    if (MODE == PT_LABEL_BOUNDS && run.bc) atomicAdd(&a.cnt[run.bp], run.bc);
    if (run.ck >= 0) run_flush<MODE, AVG>(a, A, run.ck, run.cc, run.ch, run.cl);
    if (MODE == PT_SPLAT) splat_flush(A, a.cnt, a.chi, a.clo, AVG);
    else acc_flush(A, AVG);
}

// The queued border points of a PT_HYBRID_QUEUE pass: the exact test of
// each (one per lane, no idle lanes), summed per workgroup in LDS.
template <bool AVG>
__global__ __launch_bounds__(kBlock) void k_exact(PointArgs a)
{
    extern __shared__ unsigned long long sacc[];
    const uint32_t n = min(*a.qcount, a.qcap);
    // a workgroup past the queued points leaves before its LDS sums (the
    // whole workgroup: the test is uniform)
    if ((uint64_t)blockIdx.x * blockDim.x >= n) return;
    Acc A;
    acc_init(A, sacc, a.P.np, a.lds_np, AVG, a.res);
    const DevQuery& q = a.q;
    const int agg_t = AVG ? agg_kind(a.cols, q) : kAggUnsigned;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 e = a.queue[i];
        const uint32_t v = __float_as_uint(e.z);
        const int key = exact_first<AVG, true, 4>(a.g, a.P, A, e.x, e.y, agg_t, v);
        if (key >= 0) {
            unsigned long long h = 0, l = 0;
            if (AVG) agg_bits(agg_t, v, h, l);
            acc_add(A, key, 1ull, h, l, AVG);
        }
    }
    acc_flush(A, AVG);
}

inline int blocks_for(int64_t n, int64_t per, int cap)
{
    int64_t b = (n + per - 1) / per;
    if (b < 1) b = 1;
    return (int)(b > cap ? cap : b);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

void launch_snap(hipStream_t s, const float* xy, int64_t nv, double minx, double miny, double fxs, double fys,
                 int32_t* X, int32_t* Y)
{
    if (nv <= 0) return;
    hipLaunchKernelGGL(k_snap, dim3((unsigned)((nv + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, xy, nv, minx,
                       miny, fxs, fys, X, Y);
}

void launch_prev(hipStream_t s, const int32_t* off, int np, int64_t nv, int32_t* prev)
{
    if (nv <= 0) return;
    hipLaunchKernelGGL(k_prev, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, off, np, nv, prev);
}

void launch_copy_words(hipStream_t s, CopyWords r0, CopyWords r1)
{
    const size_t n = r0.n + r1.n;
    if (n == 0) return;
    int blocks = (int)std::min<size_t>((n + kBlock - 1) / kBlock, 64);
    hipLaunchKernelGGL(k_copy_words, dim3(blocks), dim3(kBlock), 0, s, r0, r1);
}

void launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v)
{
    if (n == 0) return;
    int blocks = (int)std::min<size_t>((n + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(kBlock), 0, s, p, n, v);
}

// Contiguous row chunks per block (a multiple of the block size) so that a
// wave reads consecutive rows and spatially ordered tables stay local.
static void chunking(int64_t n, int cap, int* blocks, int64_t* per_block)
{
    int b = blocks_for(n, kBlock, cap);
    int64_t per = (n + b - 1) / b;
    per = (per + kBlock - 1) / kBlock * kBlock;
    if (per < kBlock) per = kBlock;
    *blocks = (int)std::max<int64_t>(1, (n + per - 1) / per);
    *per_block = per;
}

void launch_band(hipStream_t s, int mode, const Band* bands, int nbands, const DevPolys& P,
                 const DevCanvas& c, const uint32_t* cnt, const long long* chi,
                 const unsigned long long* clo, DevResult res, long long* bounds, void* label,
                 int label16, unsigned long long* aux, size_t lds_bytes, int opt)
{
    if (nbands <= 0) return;
    const unsigned long long* hi = (const unsigned long long*)chi;
    unsigned long long* bd = (unsigned long long*)bounds;
    if (mode == BAND_SUM)
        hipLaunchKernelGGL(k_band<BAND_SUM>, dim3(nbands), dim3(kBlock), lds_bytes, s, bands, P, c, cnt,
                           hi, clo, res, bd, label, label16, aux, opt);
    else if (mode == BAND_LABEL)
        hipLaunchKernelGGL(k_band<BAND_LABEL>, dim3(nbands), dim3(kBlock), lds_bytes, s, bands, P, c,
                           cnt, hi, clo, res, bd, label, label16, aux, opt);
    else if (mode == BAND_LABEL_ATOMIC)
        hipLaunchKernelGGL(k_band<BAND_LABEL_ATOMIC>, dim3(nbands), dim3(kBlock), lds_bytes, s, bands, P, c,
                           cnt, hi, clo, res, bd, label, label16, aux, opt);
    else
        hipLaunchKernelGGL(k_band<BAND_BOUNDS>, dim3(nbands), dim3(kBlock), lds_bytes, s, bands, P, c,
                           cnt, hi, clo, res, bd, label, label16, aux, opt);
}

void launch_touch(hipStream_t s, const int32_t* X, const int32_t* Y, const int32_t* prev,
                  int64_t nverts, const DevCanvas& c, uint32_t* mask_bits)
{
    if (nverts <= 0) return;
    int wpr = (c.W + 31) / 32;
    int blocks = (int)((nverts + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_touch, dim3(blocks), dim3(kBlock), 0, s, X, Y, prev, nverts, c.W, c.H, wpr,
                       mask_bits);
}

void launch_mark_rects(hipStream_t s, const int32_t* rects, int nrects, const DevCanvas& c,
                       uint32_t* mask_bits)
{
    if (nrects <= 0) return;
    int wpr = (c.W + 31) / 32;
    hipLaunchKernelGGL(k_mark_rects, dim3(nrects), dim3(kBlock), 0, s, rects, nrects, c.W, c.H, wpr,
                       mask_bits);
}

void launch_dilate_to_label(hipStream_t s, const uint32_t* mask_bits, const DevCanvas& c, int r,
                            void* label, int label16)
{
    int wpr = (c.W + 31) / 32;
    int64_t n = (int64_t)wpr * c.H;
    int blocks = (int)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_dilate_to_label, dim3(blocks), dim3(kBlock), 0, s, mask_bits, c.W, c.H, c.LW, wpr, r,
                       label, label16);
}

void launch_tile_labels(hipStream_t s, const void* label, int label16, const DevCanvas& c, int tshift,
                        uint32_t* tiles, unsigned long long* labeled)
{
    const int TW = c.LW >> tshift, TH = (c.H + (1 << tshift) - 1) >> tshift;
    const int nvec = c.LW / (label16 ? 8 : 4);
    const int kt = tshift == 3 ? 4 : 1;  // tile rows per workgroup, as the kernel
    const dim3 grid((unsigned)((nvec + kBlock - 1) / kBlock), (unsigned)((TH + kt - 1) / kt));
    if (tshift == 3)
        hipLaunchKernelGGL(k_tile_labels<3>, grid, dim3(kBlock), 0, s, label, label16, c.LW, c.H, TW, TH, tiles,
                           labeled);
    else if (tshift == 5)
        hipLaunchKernelGGL(k_tile_labels<5>, grid, dim3(kBlock), 0, s, label, label16, c.LW, c.H, TW, TH, tiles,
                           labeled);
    else
        hipLaunchKernelGGL(k_tile_labels<4>, grid, dim3(kBlock), 0, s, label, label16, c.LW, c.H, TW, TH, tiles,
                           labeled);
}

// Hybrid label maps in the halo mode: every pixel that is not itself border
// and has a touched pixel among its 8 neighbours gets its label flagged (or
// kLabelNoneHalo), so that the point pass runs the halo test (which reads the
// touched bitmap) only for points of flagged pixels. One thread per 32-pixel
// word of a row.
__global__ void k_halo_flags(const uint32_t* __restrict__ mask, int W, int H, int LW, int wpr, void* label,
                             int label16)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)wpr * H) return;
    const int row = (int)(t / wpr), w = (int)(t % wpr);
    uint32_t m = 0, ml = 0, mr = 0;
    for (int dr = -1; dr <= 1; dr++) {
        const int rr = row + dr;
        if (rr < 0 || rr >= H) continue;
        const uint32_t* mrow = mask + (int64_t)rr * wpr;
        m |= mrow[w];
        if (w > 0) ml |= mrow[w - 1];
        if (w + 1 < wpr) mr |= mrow[w + 1];
    }
    const uint32_t d = m | (m << 1) | (ml >> 31) | (m >> 1) | (mr << 31);
    if (!d) return;
    const int64_t rowpix = (int64_t)row * LW;
    for (int b = 0; b < 32; b++) {
        const int col = w * 32 + b;
        if (col >= W) break;
        if (!((d >> b) & 1u)) continue;
        if (label16) {
            uint16_t* l = static_cast<uint16_t*>(label) + rowpix + col;
            const uint16_t v = *l;
            if (v == (uint16_t)kLabelBorder) continue;
            *l = v == (uint16_t)kLabelNone ? (uint16_t)kLabelNoneHalo : (uint16_t)(v | kHaloFlag16);
        } else {
            uint32_t* l = static_cast<uint32_t*>(label) + rowpix + col;
            const uint32_t v = *l;
            if (v == kLabelBorder) continue;
            *l = v == kLabelNone ? kLabelNoneHalo : (v | kHaloFlag32);
        }
    }
}

void launch_halo_flags(hipStream_t s, const uint32_t* mask, const DevCanvas& c, void* label, int label16)
{
    const int wpr = (c.W + 31) / 32;
    const int64_t n = (int64_t)wpr * c.H;
    hipLaunchKernelGGL(k_halo_flags, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, mask, c.W, c.H,
                       c.LW, wpr, label, label16);
}

void launch_points(hipStream_t s, int mode, PointArgs a)
{
    if (a.end <= a.begin) return;
    // rows that pass every constraint run as a query without constraints;
    // the rows before and after them keep the filters (the sums, the canvas
    // and the border queue accumulate over the launches)
    if (!a.q.plain && a.in_b < a.in_e) {
        const int64_t ib = std::max(a.begin, a.in_b), ie = std::min(a.end, a.in_e);
        a.in_b = a.in_e = 0;
        if (ib < ie) {
            PointArgs part = a;
            part.end = ib;
            launch_points(s, mode, part);
            part.begin = ib;
            part.end = ie;
            part.q.plain = 1;
            launch_points(s, mode, part);
            part = a;
            part.begin = ie;
            launch_points(s, mode, part);
            return;
        }
    }
    // Row groups start at a multiple of kRows so that vector loads are aligned.
    a.start = a.begin - a.begin % kRows;
    const int64_t n = a.end - a.start;
    const int rows = (mode == PT_HYBRID_QUEUE && n >= kQueuePassRows4) ? 4 : kRows;
    const int64_t iter_rows = (int64_t)kBlock * rows;
    int64_t cap = a.cap;
    // a splat over a spatially ordered table takes fewer rows per workgroup:
    // they then fit its LDS pixel slots (4096 rows: 1.52 -> 1.23 ms for an
    // errorbounds splat of 100M rows at eps = 10 m, 5.2 -> 4.0 ms with sums)
    const int64_t wg_rows = (mode == PT_SPLAT && a.wg_rows > 0) ? kSplatRowsPerWorkgroup : a.wg_rows;
    if (cap <= 0)
        cap = wg_rows > 0 ? std::min<int64_t>(kPassWorkgroupsMax, std::max<int64_t>(kPassWorkgroupsMin, n / wg_rows))
                          : kPassWorkgroupsMin;
    int64_t per = (n + cap - 1) / cap;
    per = (per + iter_rows - 1) / iter_rows * iter_rows;
    a.per_block = per;
    const int blocks = (int)((n + per - 1) / per);
    const bool avg = a.q.agg_type != RJ_AGG_COUNT;
    const size_t lds = (mode == PT_SPLAT) ? (avg ? (size_t)(kSplatSlots / 2) * 24 : (size_t)kSplatSlots * 8)
                                          : acc_lds_bytes(a.P.np, a.lds_np, avg);
#define RJ_LAUNCH_PTS(M, A) \
    hipLaunchKernelGGL((k_points<M, A, kRows>), dim3(blocks), dim3(kBlock), lds, s, a)
    // a label or hybrid-queue pass with constraints: two-stage rows
    if (!a.q.plain && (mode == PT_LABEL || mode == PT_HYBRID_QUEUE)) {
        const int64_t per_c = (n + cap - 1) / cap;
        a.per_block = (per_c + kBlock * kCompactRows - 1) / (kBlock * kCompactRows) * (kBlock * kCompactRows);
        const int blocks_c = (int)((n + a.per_block - 1) / a.per_block);
#define RJ_LAUNCH_CMP(M, A) \
    hipLaunchKernelGGL((k_points<M, A, kCompactRows, true>), dim3(blocks_c), dim3(kBlock), lds, s, a)
        if (mode == PT_LABEL) {
            if (avg) RJ_LAUNCH_CMP(PT_LABEL, true); else RJ_LAUNCH_CMP(PT_LABEL, false);
        } else {
            if (avg) RJ_LAUNCH_CMP(PT_HYBRID_QUEUE, true); else RJ_LAUNCH_CMP(PT_HYBRID_QUEUE, false);
        }
#undef RJ_LAUNCH_CMP
        return;
    }
    switch (mode) {
    case PT_INDEX: if (avg) RJ_LAUNCH_PTS(PT_INDEX, true); else RJ_LAUNCH_PTS(PT_INDEX, false); break;
    case PT_LABEL: if (avg) RJ_LAUNCH_PTS(PT_LABEL, true); else RJ_LAUNCH_PTS(PT_LABEL, false); break;
    case PT_LABEL_BOUNDS: RJ_LAUNCH_PTS(PT_LABEL_BOUNDS, false); break;  // counts only
    case PT_LABEL_HYBRID:
        if (avg) RJ_LAUNCH_PTS(PT_LABEL_HYBRID, true); else RJ_LAUNCH_PTS(PT_LABEL_HYBRID, false);
        break;
    case PT_HYBRID_QUEUE:
        if (rows == 4) {
            if (avg) hipLaunchKernelGGL((k_points<PT_HYBRID_QUEUE, true, 4>), dim3(blocks), dim3(kBlock), lds, s, a);
            else hipLaunchKernelGGL((k_points<PT_HYBRID_QUEUE, false, 4>), dim3(blocks), dim3(kBlock), lds, s, a);
        } else {
            if (avg) RJ_LAUNCH_PTS(PT_HYBRID_QUEUE, true); else RJ_LAUNCH_PTS(PT_HYBRID_QUEUE, false);
        }
        break;
    default: if (avg) RJ_LAUNCH_PTS(PT_SPLAT, true); else RJ_LAUNCH_PTS(PT_SPLAT, false); break;
    }
#undef RJ_LAUNCH_PTS
}

void launch_exact(hipStream_t s, const PointArgs& a)
{
    const bool avg = a.q.agg_type != RJ_AGG_COUNT;
    const size_t lds = acc_lds_bytes(a.P.np, a.lds_np, avg);
    // grid-stride over the queue (the count is read on the device), about one
    // workgroup per 512 entries of room, 64..2048
    const int blocks = (int)std::min<uint32_t>(2048, std::max<uint32_t>(64, a.qcap / 512));
    if (avg) hipLaunchKernelGGL(k_exact<true>, dim3(blocks), dim3(kBlock), lds, s, a);
    else hipLaunchKernelGGL(k_exact<false>, dim3(blocks), dim3(kBlock), lds, s, a);
}

}  // namespace rj
