// scan_ops.hpp -- integer steps of the scan conversion shared by the band
// kernels (kernels.hip) and a host check: the crossing column of an edge on
// every row of a band, and an 8-pixel group of a row's coverage bits. Plain
// C++ with exact integer arithmetic, so that g++ compiles the same code.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define RJ_HD __host__ __device__ __forceinline__
#else
#define RJ_HD inline
#endif

namespace rj {

RJ_HD int64_t sc_floor_div(int64_t a, int64_t b)
{
    int64_t q = a / b;
    if ((a % b != 0) && (a < 0)) q--;
    return q;
}

RJ_HD int64_t sc_ceil_div(int64_t a, int64_t b)
{
    int64_t q = a / b;
    if ((a % b != 0) && (a > 0)) q++;
    return q;
}

// Crossings of edge a-b (fixed-point, cell size cs subunits) with the cell
// centers' rows of [R0, R0+NR): for each row whose center y = row*cs + cs/2
// lies in [min y, max y), emit(row, k) with k the first column whose center x
// is at or right of the crossing, k = ceil((ax*dy + dx*(row*cs + h - ay) -
// h*dy) / (cs*dy)). The numerator grows by dx*cs per row, so k follows from
// one division for the first row and a quotient/remainder step per row.
template <class F>
RJ_HD void edge_crossings(int64_t ax, int64_t ay, int64_t bx, int64_t by, int64_t cs, int R0, int NR, F&& emit)
{
    if (ay == by) return;
    if (ay > by) {
        int64_t t = ax; ax = bx; bx = t;
        t = ay; ay = by; by = t;
    }
    const int64_t h = cs / 2;
    int64_t jlo = sc_ceil_div(ay - h, cs), jhi = sc_ceil_div(by - h, cs) - 1;
    if (jlo < R0) jlo = R0;
    if (jhi > R0 + NR - 1) jhi = R0 + NR - 1;
    if (jlo > jhi) return;
    const int64_t dy = by - ay, dx = bx - ax;
    const int64_t den = cs * dy;
    const int64_t num0 = ax * dy + dx * (jlo * cs + h - ay) - h * dy;
    int64_t q = sc_floor_div(num0, den), r = num0 - q * den;  // num = q*den + r, 0 <= r < den
    const int64_t sq = sc_floor_div(dx * cs, den), sr = dx * cs - sq * den;
    for (int64_t row = jlo; row <= jhi; row++) {
        emit(row, q + (r != 0 ? 1 : 0));
        q += sq;
        r += sr;
        if (r >= den) {
            r -= den;
            q++;
        }
    }
}

// Bits col..col+7 of a row of wpr words (bit k of the result: column col+k),
// 0 for columns outside [0, bw); col in [-7, bw).
RJ_HD uint32_t bits8(const uint32_t* row, int col, int bw, int wpr)
{
    uint32_t m;
    if (col >= 0) {
        const int w = col >> 5, sh = col & 31;
        uint64_t v = row[w];
        if (sh > 24 && w + 1 < wpr) v |= (uint64_t)row[w + 1] << 32;
        m = (uint32_t)(v >> sh) & 0xFFu;
    } else {
        m = (row[0] << (-col)) & 0xFFu;
    }
    const int valid = bw - col;
    if (valid < 8) m &= (1u << valid) - 1u;
    return m;
}

}  // namespace rj
