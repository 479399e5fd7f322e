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

template <class T>
__device__ __forceinline__ T cmp_op(int op, T a, T b)
{
    return op == RJ_EQ ? a == b : op == RJ_LT ? a < b : op == RJ_LTE ? a <= b : op == RJ_GT ? a > b : a >= b;
}

// Typed comparison of 32-bit column patterns: Float columns compare as
// floats, Int as signed and Uint (and u8) as unsigned integers; the
// reference's float-bit equality on LTE/GTE (points.vert:99,107) is not kept.
__device__ __forceinline__ bool cmp_bits(int type, int op, uint32_t a, uint32_t v)
{
    if (type == RJ_ATTR_FLOAT) return cmp_op<float>(op, __uint_as_float(a), __uint_as_float(v));
    if (type == RJ_ATTR_INT) return cmp_op<int32_t>(op, (int32_t)a, (int32_t)v);
    return cmp_op<uint32_t>(op, a, v);
}

// Exact fixed-point aggregate value of a 32-bit pattern: hi = floor(v),
// lo = frac(v) * 2^32 (integer columns: hi = value, lo = 0). The fraction
// times 2^32 is below 2^32 - 255, so lo always fits 32 bits.
__device__ __forceinline__ void agg_bits(int type, uint32_t a, unsigned long long& hi, unsigned long long& lo)
{
    lo = 0;
    if (type == RJ_ATTR_FLOAT) {
        const float f = __uint_as_float(a);
        const float fl = floorf(f);
        if (fabsf(fl) < 2147483648.0f) {
            hi = (unsigned long long)(long long)(int32_t)fl;  // 32-bit conversion suffices
        } else if (fabsf(f) < 4.611686018427387904e18f) {
            hi = (unsigned long long)(long long)fl;
        } else {
            hi = 0;  // non-finite or |v| >= 2^62 contributes nothing
            return;
        }
        lo = (uint32_t)((f - fl) * 4294967296.0f);
    } else if (type == RJ_ATTR_INT) {
        hi = (unsigned long long)(long long)(int32_t)a;
    } else {
        hi = a;
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

// CPUJoin getXIndex/getYIndex (GridIndexF.cpp:39-47) with the in-grid test of
// getRegion (:163); (int) truncation done only on in-range values.
__device__ __forceinline__ int grid_cell(const DevGrid& g, float x, float y)
{
    float fx = (x - g.minx) / g.width;
    float fy = (y - g.miny) / g.height;
    if (!(fx > -1.0f && fx < (float)g.xs && fy > -1.0f && fy < (float)g.ys)) return -1;
    return (int)fy * g.xs + (int)fx;
}

// Crossing-number PIP, GridIndexF.cpp:383-402 (same operation order).
__device__ __forceinline__ bool pip_f32(const float2* __restrict__ v, int n, float tx, float ty)
{
    bool c = false;
    float2 vj = v[n - 1];
    for (int i = 0; i < n; i++) {
        float2 vi = v[i];
        if (((vi.y > ty) != (vj.y > ty)) && (tx < (vj.x - vi.x) * (ty - vi.y) / (vj.y - vi.y) + vi.x))
            c = !c;
        vj = vi;
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

// Per-polygon accumulators: the first lds_np polygons in LDS, the rest (or
// all, lds_np = 0) straight into the result arrays.
struct Acc {
    unsigned long long* cnt;  // LDS (lds_np entries)
    unsigned long long* hi;
    unsigned long long* lo;
    DevResult g;
    int lds_np;
};

__device__ __forceinline__ void acc_add(const Acc& A, int key, unsigned long long c, unsigned long long hi,
                                        unsigned long long lo, bool avg)
{
    if (key < A.lds_np) {
        atomicAdd(&A.cnt[key], c);
        if (avg) {
            atomicAdd(&A.hi[key], hi);
            atomicAdd(&A.lo[key], lo);
        }
    } else {
        atomicAdd(&A.g.cnt[key], c);
        if (avg) {
            atomicAdd((unsigned long long*)&A.g.hi[key], hi);
            atomicAdd(&A.g.lo[key], lo);
        }
    }
}

// CPUJoin getRegion over the cell's candidates. Returns the first polygon that
// contains the point; further containing polygons are added directly.
template <bool AVG>
__device__ __forceinline__ int exact_first(const DevGrid& g, const DevPolys& P, const Acc& A, float x,
                                           float y, int agg_t, uint32_t vraw)
{
    const int cell = grid_cell(g, x, y);
    if (cell < 0) return -1;
    int first = -1;
    const int k1 = g.cell_off[cell + 1];
    for (int k = g.cell_off[cell]; k < k1; k++) {
        const int p = g.cand[k];
        const int v0 = P.off[p];
        if (pip_f32(reinterpret_cast<const float2*>(P.xy) + v0, P.off[p + 1] - v0, x, y)) {
            if (first < 0) {
                first = p;
            } else {
                unsigned long long h = 0, l = 0;
                if (AVG) agg_bits(agg_t, vraw, h, l);
                acc_add(A, p, 1ull, h, l, AVG);
            }
        }
    }
    return first;
}

}  // namespace
}  // namespace rj
