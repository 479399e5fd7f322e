// point_dev.hpp -- device helpers shared by the point-pass kernels
// (kernels.hip: one query per pass; batch.hip: fused passes over many
// queries): typed filters, exact fixed-point aggregates, vector row loads, the
// pixel formula, the CPUJoin grid cell and point-in-polygon test, label-map
// reads and the per-polygon accumulators.
//
// Included only by HIP sources compiled with -ffp-contract=off: the float PIP,
// the grid cell and the pixel formula must round exactly like the host
// restatement of CPUJoin.
#pragma once
#include <type_traits>

#include "rj_dev.hpp"

namespace rj {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;

__device__ __forceinline__ int64_t floor_div(int64_t a, int64_t b)
{
    int64_t q = a / b;
    if ((a % b != 0) && (a < 0)) q--;
    return q;
}

__device__ __forceinline__ int64_t ceil_div(int64_t a, int64_t b)
{
    int64_t q = a / b;
    if ((a % b != 0) && (a > 0)) q++;
    return q;
}

// The constraints of one slot on R rows: ok[r] stays true when the row's
// value lies in the slot's key range (DevQuery). Typed comparisons: Float
// columns compare as floats, Int as signed and Uint (and u8) as unsigned
// integers; the reference's float-bit equality on LTE/GTE
// (points.vert:99,107) is not kept. A slot without constraints (its whole
// key range) is skipped.
template <int R>
__device__ __forceinline__ void slot_filter(int type, uint32_t lo, uint32_t hi, const uint32_t (&v)[R], bool (&ok)[R])
{
    if (lo == 0u && hi == 0xFFFFFFFFu) return;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t k = col_key(type, v[r]);
        ok[r] = ok[r] & (k >= lo) & (k <= hi);
    }
}

// How a pass decodes its aggregate column (wave-uniform): unsigned (Uint,
// u8), signed (Int), float, or float known finite and in [-2^31, 2^31) on
// every row of the table (no range tests).
constexpr int kAggUnsigned = 0, kAggSigned = 1, kAggFloat = 2, kAggFloatInRange = 3;

__device__ __forceinline__ int agg_kind(const DevCols& C, const DevQuery& q)
{
    const int col = q.slot_col[q.agg_slot < 0 ? 0 : q.agg_slot];
    const int t = C.type[col];
    if (t == RJ_ATTR_FLOAT) return ((C.in_range >> col) & 1u) ? kAggFloatInRange : kAggFloat;
    return t == RJ_ATTR_INT ? kAggSigned : kAggUnsigned;
}

// Exact fixed-point aggregate value of a 32-bit pattern: hi = floor(v),
// lo = frac(v) * 2^32 (integer columns: hi = value, lo = 0). For v in
// [-2^-25, 0) the float difference v - floor(v) rounds to 1.0: such a v
// counts as 0 (hi = floor(v) + 1, lo = 0). Otherwise the difference is at
// most 1 - 2^-24, so lo <= 2^32 - 256 fits 32 bits. Non-finite values and
// |v| >= 2^62 count nothing; floor(v) outside 32 bits takes the 64-bit
// conversion. For a column in range, floor(v) fits 32 bits and the
// decode is the same without branches.
// The integer value is set first and the float branches overwrite it: hipcc
// (ROCm 7.2, -O3) compiled the form if (Float) .. else if (Int) .. else
// hi = value so that u8 columns left hi unset.
__device__ __forceinline__ void agg_bits(int kind, uint32_t a, unsigned long long& hi, unsigned long long& lo)
{
    lo = 0;
    hi = kind == kAggSigned ? (unsigned long long)(long long)(int32_t)a : (unsigned long long)a;
    if (kind == kAggFloatInRange) {
        const float f = __uint_as_float(a);
        const float fl = floorf(f);
        const float fr = f - fl;
        const bool one = fr >= 1.0f;
        hi = (unsigned long long)(long long)((int32_t)fl + (one ? 1 : 0));
        lo = one ? 0ull : (unsigned long long)(uint32_t)(fr * 4294967296.0f);
    } else if (kind == kAggFloat) {
        const float f = __uint_as_float(a);
        const float fl = floorf(f);
        if (fabsf(fl) < 2147483648.0f) {
            hi = (unsigned long long)(long long)(int32_t)fl;  // 32-bit conversion suffices
        } else if (fabsf(f) < 4.611686018427387904e18f) {
            hi = (unsigned long long)(long long)fl;
        } else {
            hi = 0;
            return;
        }
        const float fr = f - fl;
        if (fr >= 1.0f) {
            hi += 1ull;
            return;
        }
        lo = (uint32_t)(fr * 4294967296.0f);
    }
}

// Rows [i0, i0+R) of one column as 32-bit patterns (u8 widened). Vector loads
// when the whole group is inside the table, guarded loads otherwise.
template <int R>
__device__ __forceinline__ void load_col(const void* col, int t, int64_t i0, bool full, int64_t end,
                                         uint32_t (&v)[R])
{
    static_assert(R == 4 || R == 8 || R == 16, "rows per thread: 4, 8 or 16");
    if (t == RJ_ATTR_UINT8) {
        const uint8_t* p = static_cast<const uint8_t*>(col);
        if (full) {
            uint32_t w[R / 4];
            if (R == 4) {
                w[0] = *reinterpret_cast<const uint32_t*>(p + i0);
            } else if (R == 8) {
                const uint2 u = *reinterpret_cast<const uint2*>(p + i0);
                w[0] = u.x; w[R / 4 > 1 ? 1 : 0] = u.y;
            } else {
                const uint4 u = *reinterpret_cast<const uint4*>(p + i0);
                w[0] = u.x; w[R / 4 > 1 ? 1 : 0] = u.y;
                w[R / 4 > 2 ? 2 : 0] = u.z;
                w[R / 4 > 3 ? 3 : 0] = u.w;
            }
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = (w[r >> 2] >> (8 * (r & 3))) & 0xFFu;
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = (i0 + r < end) ? (uint32_t)p[i0 + r] : 0u;
        }
    } else {
        const uint32_t* p = static_cast<const uint32_t*>(col);
        if (full) {
#pragma unroll
            for (int k = 0; k < R / 4; k++) {
                const uint4 a = *reinterpret_cast<const uint4*>(p + i0 + 4 * k);
                v[4 * k] = a.x; v[4 * k + 1] = a.y; v[4 * k + 2] = a.z; v[4 * k + 3] = a.w;
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = (i0 + r < end) ? p[i0 + r] : 0u;
        }
    }
}

template <int R>
__device__ __forceinline__ void load_rows(const DevCols& C, int col, int64_t i0, bool full, int64_t end,
                                          uint32_t (&v)[R])
{
    load_col<R>(C.col[col], C.type[col], i0, full, end, v);
}

// x,y of rows [i0, i0+R): vector loads when the group is inside the table,
// guarded loads otherwise; nothing when i0 >= b1.
template <int R>
__device__ __forceinline__ void load_xy(const float* xy, int64_t i0, int64_t b1, int64_t end, float (&px)[R],
                                        float (&py)[R])
{
    if (i0 >= b1) return;
    if (i0 + R <= end) {
        const float4* p4 = reinterpret_cast<const float4*>(xy + 2 * i0);
#pragma unroll
        for (int k = 0; k < R / 2; k++) {
            const float4 f = p4[k];
            px[2 * k] = f.x; py[2 * k] = f.y;
            px[2 * k + 1] = f.z; py[2 * k + 1] = f.w;
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const bool in = i0 + r < end;
            px[r] = in ? xy[2 * (i0 + r)] : 0.0f;
            py[r] = in ? xy[2 * (i0 + r) + 1] : 0.0f;
        }
    }
}

// The pixel formula shared by every kernel (and the host oracle): canvas
// membership by float comparison with the bbox ([min, max) per axis), then
// int((v - min) * s) clamped to the last pixel. Returns -1 outside.
__device__ __forceinline__ int pixel_xy(const DevCanvas& c, float x, float y, int& ix, int& iy)
{
    if (!(x >= c.minx && x < c.maxx && y >= c.miny && y < c.maxy)) {
        ix = iy = 0;
        return -1;
    }
    const float fx = (x - c.minx) * c.sx;
    const float fy = (y - c.miny) * c.sy;
    ix = min((int)fx, c.W - 1);
    iy = min((int)fy, c.H - 1);
    return iy * c.W + ix;
}

// The pixel formula for a point known to be inside the canvas: the same
// operations as pixel_xy without its membership test.
__device__ __forceinline__ int pixel_in(const DevCanvas& c, float x, float y, int& ix, int& iy)
{
    const float fx = (x - c.minx) * c.sx;
    const float fy = (y - c.miny) * c.sy;
    ix = min((int)fx, c.W - 1);
    iy = min((int)fy, c.H - 1);
    return iy * c.W + ix;
}

// Every location of the table lies inside the canvas (its bounds do, by the
// comparisons of pixel_xy, and none is NaN or infinite): a pass then takes
// pixel_in for every row (wave-uniform).
__device__ __forceinline__ bool rows_inside(const DevCols& t, const DevCanvas& c)
{
    return t.box_ok && t.box[0] >= c.minx && t.box[2] < c.maxx && t.box[1] >= c.miny && t.box[3] < c.maxy;
}

// Pixels of R rows (rows not ok: -1).
template <int R>
__device__ __forceinline__ void row_pixels(const DevCanvas& c, bool inside, const float (&px)[R],
                                           const float (&py)[R], const bool (&ok)[R], int (&pix)[R], int (&ixs)[R],
                                           int (&iys)[R])
{
    if (inside) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            int ix, iy;
            const int p = pixel_in(c, px[r], py[r], ix, iy);
            pix[r] = ok[r] ? p : -1;
            ixs[r] = ok[r] ? ix : 0;
            iys[r] = ok[r] ? iy : 0;
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
            pix[r] = pixel_xy(c, px[r], py[r], ixs[r], iys[r]);
            if (!ok[r]) pix[r] = -1;
        }
    }
}

// CPUJoin getXIndex/getYIndex (GridIndexF.cpp:39-47) with the in-grid test of
// getRegion (:163); (int) truncation done only on in-range values. row
// receives the cell's grid row.
__device__ __forceinline__ int grid_cell(const DevGrid& g, float x, float y, int& row)
{
    float fx = (x - g.minx) / g.width;
    float fy = (y - g.miny) / g.height;
    row = 0;
    if (!(fx > -1.0f && fx < (float)g.xs && fy > -1.0f && fy < (float)g.ys)) return -1;
    row = (int)fy;
    return row * g.xs + (int)fx;
}

// Crossing-number PIP (GridIndexF.cpp:383-402, the same operation order per
// edge) over the edges of one row band of a polygon: the edges a point of that
// row can cross, so the parity is that of the loop over all edges.
// U > 1: edges are loaded U at a time (loads in flight together; the
// parity does not depend on the order of the edges).
template <int U = 1>
__device__ __forceinline__ bool pip_band(const DevGrid& g, int p, int row, float tx, float ty)
{
    const int lo = g.band_lo[p];
    if (row < lo || row > g.band_hi[p]) return false;
    const int b = g.pband[p] + (row - lo);
    const int k1 = g.band_off[b + 1];
    bool c = false;
    int k = g.band_off[b];
    if (U > 1) {
        for (; k + U <= k1; k += U) {
            float4 e[U];
#pragma unroll
            for (int j = 0; j < U; j++) e[j] = g.edges[k + j];
#pragma unroll
            for (int j = 0; j < U; j++)
                if (((e[j].y > ty) != (e[j].w > ty)) && (tx < (e[j].z - e[j].x) * (ty - e[j].y) / (e[j].w - e[j].y) + e[j].x))
                    c = !c;
        }
    }
    for (; k < k1; k++) {
        const float4 e = g.edges[k];  // vi = (x, y), vj = (z, w)
        if (((e.y > ty) != (e.w > ty)) && (tx < (e.z - e.x) * (ty - e.y) / (e.w - e.y) + e.x)) c = !c;
    }
    return c;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Block-wide sum of up to 4 values, then one atomic each from thread 0.
__device__ __forceinline__ void block_add4(unsigned long long a0, unsigned long long a1,
                                           unsigned long long a2, unsigned long long a3,
                                           unsigned long long* d0, unsigned long long* d1,
                                           unsigned long long* d2, unsigned long long* d3)
{
    __shared__ unsigned long long red[4][kBlock / kWave];
    a0 = wave_sum(a0); a1 = wave_sum(a1); a2 = wave_sum(a2); a3 = wave_sum(a3);
    const int w = threadIdx.x / kWave;
    if (__lane_id() == 0) {
        red[0][w] = a0; red[1][w] = a1; red[2][w] = a2; red[3][w] = a3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s[4] = {0, 0, 0, 0};
        for (int k = 0; k < (int)(blockDim.x / kWave); k++)
            for (int m = 0; m < 4; m++) s[m] += red[m][k];
        if (d0 && s[0]) atomicAdd(d0, s[0]);
        if (d1 && s[1]) atomicAdd(d1, s[1]);
        if (d2 && s[2]) atomicAdd(d2, s[2]);
        if (d3 && s[3]) atomicAdd(d3, s[3]);
    }
}

__device__ __forceinline__ uint32_t label_at(const void* label, int label16, int64_t pix)
{
    if (label16) {
        const uint32_t v16 = static_cast<const uint16_t*>(label)[pix];
        return v16 >= 0xFFF0u ? (0xFFFF0000u | v16) : v16;  // top codes map onto none/border
    }
    return static_cast<const uint32_t*>(label)[pix];
}

// Per-polygon accumulators of a workgroup in LDS, added to the result arrays
// at the end (acc_flush). Direct: polygon k < lds_np at slot k. Hashed (for
// collections too large to index directly): kHashSlots tagged slots; a
// workgroup's rows are spatially close and meet few polygons, so nearly every
// key finds its slot, and a key whose slot holds another polygon goes to the
// result arrays directly.
constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;

struct Acc {
    unsigned long long* cnt;  // LDS
    unsigned long long* hi;
    unsigned long long* lo;
    uint32_t* tag;            // hashed mode: the polygon of each slot
    DevResult g;
    int lds_np;               // direct mode: polygons in LDS (0: none)
    int hash;                 // 1: hashed mode
    // canvas splat (PT_SPLAT): per-pixel sums of the workgroup in LDS slots
    uint32_t* stag;           // pixel of each slot
    uint32_t* scnt;
    unsigned long long* shi;
    unsigned long long* slo;
    int sslots;
};

// Canvas splat in LDS: kSplatSlots tagged slots keyed by pixel (half as many
// with sums), linear probing over 8 slots; a pixel that finds none is added
// to the canvas directly. A workgroup's rows are spatially close (a spatially
// ordered table), so its points repeat pixels and each pixel reaches the
// canvas once per workgroup instead of once per run of a lane.
constexpr int kSplatSlots = 2048;

__device__ __forceinline__ void splat_init(Acc& A, unsigned long long* sacc, bool avg)
{
    const int slots = avg ? kSplatSlots / 2 : kSplatSlots;
    A.sslots = slots;
    A.stag = reinterpret_cast<uint32_t*>(sacc);
    A.scnt = A.stag + slots;
    A.shi = reinterpret_cast<unsigned long long*>(A.scnt + slots);
    A.slo = A.shi + slots;
    for (int k = threadIdx.x; k < slots; k += blockDim.x) {
        A.stag[k] = kHashEmpty;
        A.scnt[k] = 0;
        if (avg) {
            A.shi[k] = 0;
            A.slo[k] = 0;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void splat_add(const Acc& A, uint32_t* cnt, unsigned long long* chi,
                                          unsigned long long* clo, int key, uint32_t c, unsigned long long hi,
                                          unsigned long long lo, bool avg)
{
    const uint32_t h = (uint32_t)key * 2654435761u;
    for (int p = 0; p < 8; p++) {
        const int sl = (int)((h + (uint32_t)p) & (uint32_t)(A.sslots - 1));
        uint32_t t = A.stag[sl];
        if (t == kHashEmpty) {
            const uint32_t old = atomicCAS(&A.stag[sl], kHashEmpty, (uint32_t)key);
            t = old == kHashEmpty ? (uint32_t)key : old;
        }
        if (t == (uint32_t)key) {
            atomicAdd(&A.scnt[sl], c);
            if (avg) {
                atomicAdd(&A.shi[sl], hi);
                atomicAdd(&A.slo[sl], lo);
            }
            return;
        }
    }
    atomicAdd(&cnt[key], c);
    if (avg) {
        atomicAdd(&chi[key], hi);
        atomicAdd(&clo[key], lo);
    }
}

// The workgroup's LDS pixels to the canvas (after a barrier inside).
__device__ __forceinline__ void splat_flush(const Acc& A, uint32_t* cnt, unsigned long long* chi,
                                            unsigned long long* clo, bool avg)
{
    __syncthreads();
    for (int k = threadIdx.x; k < A.sslots; k += blockDim.x) {
        const uint32_t key = A.stag[k];
        if (key == kHashEmpty) continue;
        atomicAdd(&cnt[key], A.scnt[k]);
        if (avg) {
            if (A.shi[k]) atomicAdd(&chi[key], A.shi[k]);
            if (A.slo[k]) atomicAdd(&clo[key], A.slo[k]);
        }
    }
}

__device__ __forceinline__ int hash_slot(int key)
{
    return (int)(((uint32_t)key * 2654435761u) >> (32 - 9));
}
static_assert(kHashSlots == 512, "hash_slot yields 9 bits");

// Point A at the LDS buffer `sacc` (direct: np <= lds_cap; else hashed) and
// clear it. Every thread of the workgroup calls it; a barrier follows inside.
__device__ __forceinline__ void acc_init(Acc& A, unsigned long long* sacc, int np, int lds_cap, bool avg,
                                         DevResult res)
{
    A.g = res;
    A.hash = np > lds_cap ? 1 : 0;
    A.lds_np = A.hash ? 0 : np;
    const int slots = A.hash ? kHashSlots : np;
    A.cnt = sacc;
    A.hi = sacc + slots;
    A.lo = sacc + 2 * slots;
    A.tag = reinterpret_cast<uint32_t*>(sacc + (avg ? 3 : 1) * slots);
    const int n = (avg ? 3 : 1) * slots;
    for (int k = threadIdx.x; k < n; k += blockDim.x) sacc[k] = 0;
    if (A.hash)
        for (int k = threadIdx.x; k < kHashSlots; k += blockDim.x) A.tag[k] = kHashEmpty;
    __syncthreads();
}

template <bool LDS = true>
__device__ __forceinline__ void acc_add(const Acc& A, int key, unsigned long long c, unsigned long long hi,
                                        unsigned long long lo, bool avg)
{
    if (LDS) {
        int slot = -1;
        if (key < A.lds_np) {
            slot = key;
        } else if (A.hash) {
            const int h = hash_slot(key);
            uint32_t t = A.tag[h];
            if (t == kHashEmpty) {
                const uint32_t old = atomicCAS(&A.tag[h], kHashEmpty, (uint32_t)key);
                t = old == kHashEmpty ? (uint32_t)key : old;
            }
            if (t == (uint32_t)key) slot = h;
        }
        if (slot >= 0) {
            atomicAdd(&A.cnt[slot], c);
            if (avg) {
                atomicAdd(&A.hi[slot], hi);
                atomicAdd(&A.lo[slot], lo);
            }
            return;
        }
    }
    atomicAdd(&A.g.cnt[key], c);
    if (avg) {
        atomicAdd((unsigned long long*)&A.g.hi[key], hi);
        atomicAdd(&A.g.lo[key], lo);
    }
}

// The workgroup's LDS sums to the result arrays (after a barrier inside).
__device__ __forceinline__ void acc_flush(const Acc& A, bool avg)
{
    __syncthreads();
    const int slots = A.hash ? kHashSlots : A.lds_np;
    for (int k = threadIdx.x; k < slots; k += blockDim.x) {
        const unsigned long long c = A.cnt[k];
        if (!c) continue;
        const int key = A.hash ? (int)A.tag[k] : k;
        atomicAdd(&A.g.cnt[key], c);
        if (avg) {
            atomicAdd((unsigned long long*)&A.g.hi[key], A.hi[k]);
            atomicAdd(&A.g.lo[key], A.lo[k]);
        }
    }
}

// CPUJoin getRegion over the cell's candidates. Returns the first polygon that
// contains the point; further containing polygons are added directly.
template <bool AVG, bool LDS = true, int U = 1>
__device__ __forceinline__ int exact_first(const DevGrid& g, const DevPolys& P, const Acc& A, float x,
                                           float y, int agg_t, uint32_t vraw)
{
    int grow;
    const int cell = grid_cell(g, x, y, grow);
    if (cell < 0) return -1;
    const int row = band_row(y, g.miny, g.band_h, g.band_rows);
    int first = -1;
    const int k1 = g.cell_off[cell + 1];
    for (int k = g.cell_off[cell]; k < k1; k++) {
        const int p = g.cand[k];
        if (pip_band<U>(g, p, row, x, y)) {
            if (first < 0) {
                first = p;
            } else {
                unsigned long long h = 0, l = 0;
                if (AVG) agg_bits(agg_t, vraw, h, l);
                acc_add<LDS>(A, p, 1ull, h, l, AVG);
            }
        }
    }
    return first;
}

// A lane's running (key, count, sum) of a point pass, and the wave's staged
// border rows (queue_border): ring position qh and count qn, the same in
// every lane that ran the wave's last row group.
struct Run {
    int ck;
    uint32_t cc;
    unsigned long long ch, cl;
    int qh, qn;
    int bp;        // PT_LABEL_BOUNDS: the lane's run of one boundary pixel
    uint32_t bc;   // and its points (0: none)
};

template <int MODE, bool AVG, bool LDS = true>
__device__ __forceinline__ void run_flush(const PointArgs& a, const Acc& A, int ck, uint32_t cc,
                                          unsigned long long ch, unsigned long long cl)
{
    if (MODE == PT_SPLAT) {
        splat_add(A, a.cnt, a.chi, a.clo, ck, cc, ch, cl, AVG);
    } else {
        acc_add<LDS>(A, ck, cc, ch, cl, AVG);
    }
}

// Hybrid join: does the point (in untouched pixel ix, iy) lie within halo_d
// pixels of a side or corner it shares with a touched pixel? Edges lie only
// in touched pixels, so a point farther than that from every touched
// neighbour is farther than the float and snapping error from every edge and
// takes its pixel center's label. The fraction uses the pixel formula's own
// float operations.
__device__ __forceinline__ bool halo_near(const PointArgs& a, float x, float y, int ix, int iy)
{
    const float fx = (x - a.cv.minx) * a.cv.sx - (float)ix;
    const float fy = (y - a.cv.miny) * a.cv.sy - (float)iy;
    const float d = a.halo_d;
    const int dx = fx < d ? -1 : (fx > 1.0f - d ? 1 : 0);
    const int dy = fy < d ? -1 : (fy > 1.0f - d ? 1 : 0);
    if (!dx && !dy) return false;
    auto touched = [&](int cx, int cy) -> bool {
        if (cx < 0 || cy < 0 || cx >= a.cv.W || cy >= a.cv.H) return false;
        return (a.halo[(int64_t)cy * a.halo_wpr + (cx >> 5)] >> (cx & 31)) & 1u;
    };
    if (dx && touched(ix + dx, iy)) return true;
    if (dy && touched(ix, iy + dy)) return true;
    return dx && dy && touched(ix + dx, iy + dy);
}

// Exact tests of a wave's border rows (hybrid join), compacted: the rows of
// all lanes that need the test are queued in LDS and the lanes take them one
// each, so a wave runs about one round of the test instead of one per row of
// its lanes. Only the wave itself touches its queue (no workgroup barrier).
constexpr int kExactQueue = 128;  // entries per wave

template <bool AVG, bool LDS, int R>
__device__ __forceinline__ void exact_compact(const PointArgs& a, const Acc& A, int agg_t, const float (&px)[R],
                                              const float (&py)[R], const uint32_t (&vraw)[R], bool (&need)[R],
                                              int (&key)[R])
{
    __shared__ float4 s_q[kBlock / kWave][kExactQueue];
    __shared__ int s_res[kBlock / kWave][kExactQueue];
    const int w = threadIdx.x / kWave;
    const int lane = (int)__lane_id();
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // lanes past the end of the rows may have left the loop: the active ones
    // share the queue by rank
    const unsigned long long act = __ballot(1);
    const int rank = __popcll(act & below);
    const int nact = __popcll(act);
    for (;;) {
        int cnt = 0;
        int pos[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const unsigned long long m = __ballot(need[r]);
            const int p = cnt + __popcll(m & below);
            pos[r] = -1;
            if (need[r] && p < kExactQueue) {
                s_q[w][p] = make_float4(px[r], py[r], __uint_as_float(vraw[r]), 0.0f);
                pos[r] = p;
                need[r] = false;
            }
            cnt += __popcll(m);
        }
        cnt = min(cnt, kExactQueue);
        if (cnt == 0) break;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int i = rank; i < cnt; i += nact) {
            const float4 e = s_q[w][i];
            s_res[w][i] = exact_first<AVG, LDS>(a.g, a.P, A, e.x, e.y, agg_t, __float_as_uint(e.z));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int r = 0; r < R; r++)
            if (pos[r] >= 0) key[r] = s_res[w][pos[r]];
        __builtin_amdgcn_wave_barrier();  // the queue is rewritten by the next round
    }
}

// LDS writes of a wave's lanes visible to its other lanes (and its reads
// done before later writes).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The border queue of a pass is filled a wave's worth at a time: a wave
// stages its border rows in an LDS ring and reserves queue room with one
// atomic per 64 rows (a returning atomic on the pass's one counter for every
// row group of every wave serialized the pass: hybrid counts over 900M and
// 580M rows of a day-ordered table, 260 and 900 polygons, 12.8 -> 2.8 ms
// and 13.3 -> 2.5 ms). Rows past the capacity are dropped; the count still
// grows, so the host sees the overflow.
constexpr int kBorderStage = 128;  // staged rows per wave (a power of two, >= 128)

__device__ __forceinline__ float4* border_stage()
{
    __shared__ float4 s_bq[kBlock / kWave][kBorderStage];
    return s_bq[threadIdx.x / kWave];
}

// The first n (<= 64) staged rows into the queue (the wave's active lanes).
__device__ __forceinline__ void border_flush(const PointArgs& a, Run& run, int n)
{
    float4* stage = border_stage();
    wave_sync();
    const int lane = (int)__lane_id();
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const unsigned long long act = __ballot(1);
    const int lead = __ffsll((long long)act) - 1;
    const int rank = __popcll(act & below), nact = __popcll(act);
    uint32_t base = 0;
    if (lane == lead) base = atomicAdd(a.qcount, (uint32_t)n);
    base = __shfl(base, lead, kWave);
    for (int i = rank; i < n; i += nact) {
        const uint32_t k = base + (uint32_t)i;
        if (k < a.qcap) a.queue[k] = stage[(run.qh + i) & (kBorderStage - 1)];
    }
    run.qh = (run.qh + n) & (kBorderStage - 1);
    run.qn -= n;
    wave_sync();
}

// Stage a wave's border rows (need[r]) in ballot order.
template <int R>
__device__ __forceinline__ void queue_border(const PointArgs& a, const float (&px)[R], const float (&py)[R],
                                             const uint32_t (&vraw)[R], const bool (&need)[R], Run& run)
{
    float4* stage = border_stage();
    const int lane = (int)__lane_id();
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < R; r++) {
        const unsigned long long m = __ballot(need[r]);
        if (m == 0) continue;
        if (need[r])
            stage[(run.qh + run.qn + __popcll(m & below)) & (kBorderStage - 1)] =
                make_float4(px[r], py[r], __uint_as_float(vraw[r]), 0.0f);
        run.qn += __popcll(m);
        if (run.qn >= kWave) border_flush(a, run, kWave);
    }
}

// The staged rows left at the end of a pass (every lane of the wave; lane
// 0, whose rows come first, ran every row group the wave ran).
__device__ __forceinline__ void border_finish(const PointArgs& a, Run& run)
{
    run.qh = __shfl(run.qh, 0, kWave);
    run.qn = __shfl(run.qn, 0, kWave);
    if (run.qn > 0) border_flush(a, run, run.qn);
}

// Labels of R rows from their pixels (pix < 0: no label), as point_rows
// computes them (kept apart from point_rows, whose code generation is tuned
// on its own): the tile map answers for whole tiles, mixed tiles read the
// pixel, and a row in the same tile (pixel) as the row before it takes that
// row's label without a load.
// HYB: hybrid maps in the halo mode, whose flagged pixels (next to a touched
// pixel) send points within the float margin of the touched pixel to the
// exact test (kLabelBorder). `between` runs after the tile loads are issued
// and before their values are used (a prefetch placed there does not delay
// them: a wave's vector loads complete in order).
template <bool HYB, int R, class F>
__device__ __forceinline__ void row_labels(const PointArgs& a, const float (&px)[R], const float (&py)[R],
                                           const int (&pix)[R], const int (&ixs)[R], const int (&iys)[R],
                                           uint32_t (&lab)[R], F&& between)
{
    int tix[R];
#pragma unroll
    for (int r = 0; r < R; r++) tix[r] = pix[r] < 0 ? -1 : (iys[r] >> a.tshift) * a.tiles_w + (ixs[r] >> a.tshift);
    uint32_t tl[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        tl[r] = kLabelNone;
        if (tix[r] >= 0 && (r == 0 || tix[r] != tix[r - 1])) tl[r] = a.tiles[tix[r]];
    }
    between();
#pragma unroll
    for (int r = 0; r < R; r++) lab[r] = (r == 0 || tix[r] != tix[r - 1]) ? tl[r] : lab[r > 0 ? r - 1 : 0];
    uint32_t pl[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        pl[r] = 0;
        if (lab[r] == kLabelMixed && (r == 0 || pix[r] != pix[r - 1]))
            pl[r] = label_at(a.label, a.label16, (int64_t)iys[r] * a.cv.LW + ixs[r]);
    }
#pragma unroll
    for (int r = 0; r < R; r++)
        if (lab[r] == kLabelMixed) lab[r] = (r == 0 || pix[r] != pix[r - 1]) ? pl[r] : lab[r > 0 ? r - 1 : 0];
    if (HYB && a.halo) {
        // only pixels flagged next to a touched pixel take the halo test
        const uint32_t fb = a.label16 ? kHaloFlag16 : kHaloFlag32;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const bool f = lab[r] == kLabelNoneHalo || (lab[r] < kLabelMaxId && (lab[r] & fb));
            lab[r] = lab[r] == kLabelNoneHalo ? kLabelNone : (f ? lab[r] & ~fb : lab[r]);
            if (f && halo_near(a, px[r], py[r], ixs[r], iys[r])) lab[r] = kLabelBorder;
        }
    }
}

// The slots of one query decoded once per pass into wave-uniform values:
// the column of each slot that is read (constrained, or the aggregate's),
// and for a constrained slot its key range as lo and span = hi - lo, with
// col_key computed as k = v ^ (xr | (sign(v) & fl)) (Float: xr = 2^31,
// fl = ~0; Int: xr = 2^31; Uint and u8: neither). A row passes a slot when
// k - lo <= span (unsigned); no row passes a slot in none_mask (lo > hi).
struct SlotPlan {
    const void* col[kMaxSlots];
    int32_t type[kMaxSlots];
    uint32_t xr[kMaxSlots], fl[kMaxSlots], lo[kMaxSlots], span[kMaxSlots];
    uint32_t read_mask;  // bit s: slot s is read
    uint32_t filt_mask;  // bit s: slot s has constraints
    uint32_t none_mask;  // bit s: no value passes slot s
    int agg;             // slot of the aggregate column (-1: count)
};

__device__ __forceinline__ SlotPlan make_plan(const DevQuery& q, const DevCols& C, bool avg)
{
    SlotPlan P;
    P.read_mask = P.filt_mask = P.none_mask = 0;
    P.agg = avg ? q.agg_slot : -1;
#pragma unroll
    for (int s = 0; s < kMaxSlots; s++) {
        const bool in = s < q.nslots;
        const int col = in ? q.slot_col[s] : 0;
        const int t = C.type[col];
        const uint32_t lo = q.slot_lo[s], hi = q.slot_hi[s];
        const bool filt = in && !(lo == 0u && hi == 0xFFFFFFFFu);
        P.col[s] = C.col[col];
        P.type[s] = t;
        P.xr[s] = (t == RJ_ATTR_FLOAT || t == RJ_ATTR_INT) ? 0x80000000u : 0u;
        P.fl[s] = t == RJ_ATTR_FLOAT ? 0xFFFFFFFFu : 0u;
        P.lo[s] = lo;
        P.span[s] = hi - lo;
        if (filt) P.filt_mask |= 1u << s;
        if (filt && lo > hi) P.none_mask |= 1u << s;
        if (filt || (in && s == P.agg)) P.read_mask |= 1u << s;
    }
    return P;
}

// The constraints of slot s on R rows with values v (the test of
// slot_filter): ok[r] stays true when the row passes.
template <int R>
__device__ __forceinline__ void plan_filter(const SlotPlan& P, int s, const uint32_t (&v)[R], bool (&ok)[R])
{
    if ((P.none_mask >> s) & 1u) {
#pragma unroll
        for (int r = 0; r < R; r++) ok[r] = false;
        return;
    }
    if (P.xr[s] | P.fl[s]) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t k = v[r] ^ (P.xr[s] | ((uint32_t)((int32_t)v[r] >> 31) & P.fl[s]));
            ok[r] = ok[r] & (k - P.lo[s] <= P.span[s]);
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) ok[r] = ok[r] & (v[r] - P.lo[s] <= P.span[s]);
    }
}

// One row into its polygon's running (key, count, sum) of the lane: a new
// key flushes the run first.
template <int MODE, bool AVG, bool LDS>
__device__ __forceinline__ void run_add(const PointArgs& a, const Acc& A, Run& run, int k, int agg_t, uint32_t v)
{
    if (k != run.ck) {
        if (run.ck >= 0) run_flush<MODE, AVG, LDS>(a, A, run.ck, run.cc, run.ch, run.cl);
        run.ck = k;
        run.cc = 0;
        run.ch = run.cl = 0;
    }
    run.cc++;
    if (AVG) {
        unsigned long long h, l;
        agg_bits(agg_t, v, h, l);
        run.ch += h;
        run.cl += l;
    }
}

// Rows [lo, hi) of a label or hybrid pass whose query has constraints, in
// two stages. Each lane reads the columns of its R rows and tests the
// constraints; the rows that pass go to the wave's queue in LDS (x, y and
// the aggregate's value). Whenever the queue holds two waves' worth, every
// lane takes two of them (rows lane and lane + 64) through the pixel, the
// label maps (and, in a hybrid pass, the halo test and the border queue)
// into its polygon's run, the two rows' gathers in flight together. The
// label work then scales with the rows that pass, not the rows read (500M
// rows, 10% passing, hybrid: 2.50 -> 1.62 ms; 33% passing, label: 1.54 ->
// 1.25 ms). Only the wave touches its queue (no workgroup barrier); rows
// enter it two per lane at a time, so it never holds more than 127 + 128.
// (Three rows a lane per drain, with a queue of 512: hybrid 1.67 ms, label
// 1.40 ms; one: hybrid 2.12 ms, label 1.35 ms.)
constexpr int kDrainRows = 2;
constexpr int kRowQueue = 256;  // entries per wave (a power of two above 64 * kDrainRows + 127)

template <int MODE, bool AVG, int R, bool LDS>
__device__ __forceinline__ void rows_compact(const PointArgs& a, int64_t start, int64_t lo, int64_t hi,
                                             const Acc& A, Run& run)
{
    static_assert(MODE == PT_LABEL || MODE == PT_HYBRID_QUEUE, "label and hybrid-queue passes");
    static_assert(R % 2 == 0, "rows enter the queue two at a time");
    __shared__ float s_x[kBlock / kWave][kRowQueue];
    __shared__ float s_y[kBlock / kWave][kRowQueue];
    __shared__ uint32_t s_v[AVG ? kBlock / kWave : 1][AVG ? kRowQueue : 1];
    constexpr int kMask = kRowQueue - 1;
    const int w = threadIdx.x / kWave;
    const int lane = (int)__lane_id();
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const DevQuery& q = a.q;
    const int agg_t = AVG ? agg_kind(a.cols, q) : kAggUnsigned;
    const SlotPlan SP = make_plan(q, a.cols, AVG);
    const int64_t stride = (int64_t)blockDim.x * R;
    const bool all_in = rows_inside(a.cols, a.cv);
    int head = 0, cnt = 0;  // the wave's queue: entries head .. head + cnt - 1 (mod kRowQueue)
    // the first n (<= K * 64) rows of the queue, K per lane (rows lane,
    // lane + 64, ...): their label-map gathers are in flight together
    auto drain = [&](auto kk, int n) {
        constexpr int K = decltype(kk)::value;
        wave_sync();
        float px[K], py[K];
        uint32_t v[K];
        int pix[K], ixs[K], iys[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const bool in = lane + k * kWave < n;
            const int e = (head + lane + k * kWave) & kMask;
            px[k] = in ? s_x[w][e] : 0.0f;
            py[k] = in ? s_y[w][e] : 0.0f;
            v[k] = (AVG && in) ? s_v[w][e] : 0u;
            ixs[k] = iys[k] = 0;
            if (all_in) {
                int ix, iy;
                const int p = pixel_in(a.cv, px[k], py[k], ix, iy);
                pix[k] = in ? p : -1;
                ixs[k] = in ? ix : 0;
                iys[k] = in ? iy : 0;
            } else {
                pix[k] = in ? pixel_xy(a.cv, px[k], py[k], ixs[k], iys[k]) : -1;
            }
        }
        head = (head + n) & kMask;
        cnt -= n;
        uint32_t lab[K];
        row_labels<MODE == PT_HYBRID_QUEUE, K>(a, px, py, pix, ixs, iys, lab, [] {});
        if (MODE == PT_HYBRID_QUEUE) {
            bool need[K];
#pragma unroll
            for (int k = 0; k < K; k++) need[k] = lab[k] == kLabelBorder;
            queue_border<K>(a, px, py, v, need, run);
        }
#pragma unroll
        for (int k = 0; k < K; k++)
            if (lab[k] < kLabelMaxId) run_add<MODE, AVG, LDS>(a, A, run, (int)lab[k], agg_t, v[k]);
    };
    using one = std::integral_constant<int, 1>;
    using many = std::integral_constant<int, kDrainRows>;
    float nx[R], ny[R];
    load_xy<R>(a.xy, start + (int64_t)threadIdx.x * R, hi, a.end, nx, ny);
    for (int64_t base = start; base < hi; base += stride) {
        const int64_t i0 = base + (int64_t)threadIdx.x * R;
        float px[R], py[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            px[r] = nx[r];
            py[r] = ny[r];
        }
        load_xy<R>(a.xy, i0 + stride, hi, a.end, nx, ny);
        // interior: the workgroup's whole row group lies in [lo, hi) (and so
        // in the table): no per-row bounds
        const bool interior = base >= lo && base + stride <= hi;
        bool ok[R];
        if (interior) {
#pragma unroll
            for (int r = 0; r < R; r++) ok[r] = true;
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) ok[r] = (i0 + r >= lo) && (i0 + r < hi);
        }
        uint32_t vraw[R];
#pragma unroll
        for (int r = 0; r < R; r++) vraw[r] = 0;
        if (interior || i0 < hi) {
            const bool full = interior || (i0 + R <= a.end);
#pragma unroll
            for (int s = 0; s < kMaxSlots; s++) {
                if (!((SP.read_mask >> s) & 1u)) continue;
                uint32_t v[R];
                if (interior) load_col<R>(SP.col[s], SP.type[s], i0, true, a.end, v);
                else load_col<R>(SP.col[s], SP.type[s], i0, full, a.end, v);
                if ((SP.filt_mask >> s) & 1u) plan_filter<R>(SP, s, v, ok);
                if (AVG && s == SP.agg) {
#pragma unroll
                    for (int r = 0; r < R; r++) vraw[r] = v[r];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; r += 2) {
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const unsigned long long m = __ballot(ok[r + u]);
                if (ok[r + u]) {
                    const int e = (head + cnt + __popcll(m & below)) & kMask;
                    s_x[w][e] = px[r + u];
                    s_y[w][e] = py[r + u];
                    if (AVG) s_v[w][e] = vraw[r + u];
                }
                cnt += __popcll(m);
            }
            if (cnt >= kDrainRows * kWave) drain(many{}, kDrainRows * kWave);
        }
    }
    if (cnt > kWave) drain(many{}, cnt);
    else if (cnt > 0) drain(one{}, cnt);
}

// Rows [lo, hi) of one point pass, read in groups of R consecutive rows per
// thread from `start` (a multiple of R, <= lo). Every row gets one key
// (polygon id, or pixel for PT_SPLAT) and the lane keeps a running
// (key, count, sum) that it flushes only when the key changes: rows of a
// spatially ordered table mostly repeat the key, so flushes are rare. x,y of
// the next row group are loaded one iteration ahead: the tile gather needs
// only x,y and then overlaps the column loads of its group. a.end is the last
// row the loads may touch. LDS = false: every sum goes to the result arrays
// (A.lds_np is ignored).
// CMP: the rows go through rows_compact (label and hybrid-queue passes;
// any query, though only one with constraints gains). plain_rows: rows
// [lo, hi) pass every constraint of the query (they are read as for a query
// without constraints).
template <int MODE, bool AVG, int R, bool LDS = true, bool CMP = false>
__device__ __forceinline__ void point_rows(const PointArgs& a, int64_t start, int64_t lo, int64_t hi, const Acc& A,
                                           Run& run, bool plain_rows = false)
{
    if constexpr (CMP) {
        rows_compact<MODE, AVG, R, LDS>(a, start, lo, hi, A, run);
        return;
    }
    const DevQuery& q = a.q;
    const int agg_t = AVG ? agg_kind(a.cols, q) : kAggUnsigned;
    const int64_t stride = (int64_t)blockDim.x * R;
    const bool all_in = MODE != PT_INDEX && rows_inside(a.cols, a.cv);
    float nx[R], ny[R];
    load_xy<R>(a.xy, start + (int64_t)threadIdx.x * R, hi, a.end, nx, ny);
    for (int64_t base = start; base < hi; base += stride) {
        const int64_t i0 = base + (int64_t)threadIdx.x * R;
        float px[R], py[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            px[r] = nx[r];
            py[r] = ny[r];
        }
        constexpr bool late =
            MODE == PT_LABEL || MODE == PT_LABEL_HYBRID || MODE == PT_HYBRID_QUEUE || MODE == PT_LABEL_BOUNDS;
        if (!late) load_xy<R>(a.xy, i0 + stride, hi, a.end, nx, ny);
        // interior: the workgroup's whole row group lies in [lo, hi) (and so
        // in the table): no per-row bounds
        const bool interior = base >= lo && base + stride <= hi;
        if (!interior && i0 >= hi) continue;
        const bool full = interior || (i0 + R <= a.end);
        bool ok[R];
        if (interior) {
#pragma unroll
            for (int r = 0; r < R; r++) ok[r] = true;
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) ok[r] = (i0 + r >= lo) && (i0 + r < hi);
        }
        // raw 32-bit aggregate values; decoded to fixed point only when added
        uint32_t vraw[R];
#pragma unroll
        for (int r = 0; r < R; r++) vraw[r] = 0;
        if (q.plain || plain_rows) {
            // no constraints: at most the aggregate column
            if (AVG) load_rows<R>(a.cols, q.slot_col[q.agg_slot], i0, full, a.end, vraw);
        } else
        // one slot at a time: fewer registers than every slot's loads in flight
#pragma unroll 1
        for (int s = 0; s < kMaxSlots; s++) {
            if (s >= q.nslots) break;
            uint32_t v[R];
            const int col = q.slot_col[s];
            load_rows<R>(a.cols, col, i0, full, a.end, v);
            slot_filter<R>(a.cols.type[col], q.slot_lo[s], q.slot_hi[s], v, ok);
            if (AVG && s == q.agg_slot) {
#pragma unroll
                for (int r = 0; r < R; r++) vraw[r] = v[r];
            }
        }
        int key[R];
        if (MODE == PT_INDEX) {
#pragma unroll
            for (int r = 0; r < R; r++)
                key[r] = ok[r] ? exact_first<AVG, LDS, 4>(a.g, a.P, A, px[r], py[r], agg_t, vraw[r]) : -1;
        } else {
            int pix[R], ixs[R], iys[R];
            row_pixels<R>(a.cv, all_in, px, py, ok, pix, ixs, iys);
            if (MODE == PT_SPLAT) {
#pragma unroll
                for (int r = 0; r < R; r++) key[r] = pix[r];
            } else {
                uint32_t lab[R];
                // the tile map answers for whole tiles; mixed tiles read the
                // pixel. A row in the same tile (pixel) as the row before it
                // takes that row's label without a load.
                int tix[R];
#pragma unroll
                for (int r = 0; r < R; r++)
                    tix[r] = pix[r] < 0 ? -1 : (iys[r] >> a.tshift) * a.tiles_w + (ixs[r] >> a.tshift);
                uint32_t tl[R];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    tl[r] = kLabelNone;
                    if (tix[r] >= 0 && (r == 0 || tix[r] != tix[r - 1])) tl[r] = a.tiles[tix[r]];
                }
                // the next group's x,y after the tile loads: waiting for the
                // tiles then does not wait for them
                if (late) load_xy<R>(a.xy, i0 + stride, hi, a.end, nx, ny);
#pragma unroll
                for (int r = 0; r < R; r++) lab[r] = (r == 0 || tix[r] != tix[r - 1]) ? tl[r] : lab[r > 0 ? r - 1 : 0];
                uint32_t pl[R];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    pl[r] = 0;
                    if (lab[r] == kLabelMixed && (r == 0 || pix[r] != pix[r - 1]))
                        pl[r] = label_at(a.label, a.label16, (int64_t)iys[r] * a.cv.LW + ixs[r]);
                }
#pragma unroll
                for (int r = 0; r < R; r++)
                    if (lab[r] == kLabelMixed) lab[r] = (r == 0 || pix[r] != pix[r - 1]) ? pl[r] : lab[r > 0 ? r - 1 : 0];
                if ((MODE == PT_LABEL_HYBRID || MODE == PT_HYBRID_QUEUE) && a.halo) {
                    // only pixels flagged next to a touched pixel take the halo test
                    const uint32_t fb = a.label16 ? kHaloFlag16 : kHaloFlag32;
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const bool f = lab[r] == kLabelNoneHalo || (lab[r] < kLabelMaxId && (lab[r] & fb));
                        lab[r] = lab[r] == kLabelNoneHalo ? kLabelNone : (f ? lab[r] & ~fb : lab[r]);
                        if (f && halo_near(a, px[r], py[r], ixs[r], iys[r])) lab[r] = kLabelBorder;
                    }
                }
                if (MODE == PT_LABEL_BOUNDS) {
                    // flagged (boundary) pixels: their points are also counted
                    // per pixel, in a run per lane
                    const uint32_t fb = a.label16 ? kHaloFlag16 : kHaloFlag32;
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const bool bd = lab[r] == kLabelNoneHalo || (lab[r] < kLabelMaxId && (lab[r] & fb));
                        lab[r] = lab[r] == kLabelNoneHalo ? kLabelNone : (bd ? lab[r] & ~fb : lab[r]);
                        if (bd) {
                            if (pix[r] != run.bp || run.bc == 0) {
                                if (run.bc) atomicAdd(&a.cnt[run.bp], run.bc);
                                run.bp = pix[r];
                                run.bc = 0;
                            }
                            run.bc++;
                        }
                    }
                }
#pragma unroll
                for (int r = 0; r < R; r++) key[r] = lab[r] < kLabelMaxId ? (int)lab[r] : -1;
                if (MODE == PT_LABEL_HYBRID || MODE == PT_HYBRID_QUEUE) {
                    bool need[R];
#pragma unroll
                    for (int r = 0; r < R; r++) need[r] = lab[r] == kLabelBorder;
                    if (MODE == PT_LABEL_HYBRID) exact_compact<AVG, LDS, R>(a, A, agg_t, px, py, vraw, need, key);
                    else queue_border<R>(a, px, py, vraw, need, run);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int k = key[r];
            if (k < 0) continue;
            if (k != run.ck) {
                if (run.ck >= 0) run_flush<MODE, AVG, LDS>(a, A, run.ck, run.cc, run.ch, run.cl);
                run.ck = k;
                run.cc = 0;
                run.ch = run.cl = 0;
            }
            run.cc++;
            if (AVG) {
                unsigned long long h, l;
                agg_bits(agg_t, vraw[r], h, l);
                run.ch += h;
                run.cl += l;
            }
        }
    }
}

}  // namespace
}  // namespace rj
