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

// ---------------------------------------------------------------------------
// Exact area coverage of pixels by a ring (signed-area accumulation)
// ---------------------------------------------------------------------------

// One pixel's area in the units of the coverage sums: 1/256 of a square
// subunit, 256 subunits a pixel side.
constexpr int64_t kAreaOne = int64_t(1) << 24;

// Area of the piece's row band left of the vertical line x = X, floored to
// the units above: G(X) = integral over the piece's height h of
// max(X - x(t), 0) dt, x(t) running linearly from lo to hi (mid their mean,
// k = h / (2 (hi - lo))). IEEE double with no contraction, so host and device
// builds floor the same value.
RJ_HD int64_t area_left(double X, double lo, double hi, double mid, double hd, double k)
{
    if (X <= lo) return 0;
    double g;
    if (X >= hi) {
        g = hd * (X - mid);
    } else {
        const double t = X - lo;
        g = t * t * k;
    }
    return (int64_t)__builtin_floor(g * 256.0);
}

// Signed area contributions of edge a-b (fixed-point subunits) to the pixels
// of rows [R0, R0+NR): in every row the edge crosses, emit(row, col, d) adds d
// (wrapping 32-bit) at columns cmin..cmax+1 of the piece, so that a prefix sum
// along the row gives, per pixel, the area right of the edge within the
// row's piece (0 left of it, the full 256 x h subunits right of it), signed
// by the edge's direction. Summed over a ring's edges the prefix is the
// ring's winding-weighted area in each pixel, kAreaOne for a covered pixel.
template <class F>
RJ_HD void edge_area(int64_t ax, int64_t ay, int64_t bx, int64_t by, int R0, int NR, F&& emit)
{
    if (ay == by) return;
    uint32_t sgn = 1u;
    if (ay > by) {
        int64_t t = ax; ax = bx; bx = t;
        t = ay; ay = by; by = t;
        sgn = 0xFFFFFFFFu;  // -1
    }
    int64_t jlo = sc_floor_div(ay, 256), jhi = sc_ceil_div(by, 256) - 1;
    if (jlo < R0) jlo = R0;
    if (jhi > R0 + NR - 1) jhi = R0 + NR - 1;
    if (jlo > jhi) return;
    const int64_t den = by - ay, dx = bx - ax, D = 256 * den;
    const double dden = (double)den;
    // x at a row boundary y: num / den, as a double (one rounding) and as the
    // pixel column floor(num / (256 den)) with its remainder. Between interior
    // boundaries num grows by dx * 256: the column steps instead of dividing.
    auto at = [&](int64_t y, int64_t& n, double& u, int64_t& c, int64_t& r) {
        n = ax * den + dx * (y - ay);
        u = (double)n / dden;
        c = sc_floor_div(n, D);
        r = n - c * D;
    };
    int64_t yl = jlo * 256 > ay ? jlo * 256 : ay;
    int64_t n0, c0, r0;
    double u0;
    at(yl, n0, u0, c0, r0);
    const int64_t s = dx * 256;
    int64_t sq = 0, sr = 0;
    if (jlo < jhi) {
        sq = sc_floor_div(s, D);
        sr = s - sq * D;
    }
    for (int64_t row = jlo; row <= jhi; row++) {
        int64_t yh, n1, c1, r1;
        double u1;
        if (row < jhi && row > jlo) {
            yh = row * 256 + 256;
            n1 = n0 + s;
            u1 = (double)n1 / dden;
            c1 = c0 + sq;
            r1 = r0 + sr;
            if (r1 >= D) {
                r1 -= D;
                c1++;
            }
        } else {
            yh = row * 256 + 256 < by ? row * 256 + 256 : by;
            at(yh, n1, u1, c1, r1);
        }
        const int64_t h = yh - yl;
        // x at the piece's ends: num / den exactly, one rounding each
        const double lo = u0 < u1 ? u0 : u1, hi = u0 < u1 ? u1 : u0;
        const double mid = 0.5 * (u0 + u1), hd = (double)h;
        const double k = hi > lo ? hd / (2.0 * (hi - lo)) : 0.0;
        const int64_t cmin = n0 < n1 ? c0 : c1, cmax = n0 < n1 ? c1 : c0;
        int64_t gprev = 0, fprev = 0;  // G at the column's left side, F of the previous column
        for (int64_t c = cmin; c <= cmax; c++) {
            const int64_t g = area_left((double)((c + 1) * 256), lo, hi, mid, hd, k);
            const int64_t f = g - gprev;  // area right of the edge inside column c
            emit(row, c, (uint32_t)(f - fprev) * sgn);
            fprev = f;
            gprev = g;
        }
        emit(row, cmax + 1, (uint32_t)(65536 * h - fprev) * sgn);
        yl = yh;
        n0 = n1;
        u0 = u1;
        c0 = c1;
        r0 = r1;
    }
}

// The x-extent of edge a-b in every pixel row whose closed square row it
// meets (rows [R0, R0+NR) within [0, H)): emit(j, xlo, xhi) with xlo the
// floor of the smallest and xhi the ceiling of the largest x of the edge's
// piece in the row (subunits). Rows share their boundaries; between interior
// boundaries the quotient steps instead of dividing.
template <class F>
RJ_HD void edge_touch_rows(int64_t ax, int64_t ay, int64_t bx, int64_t by, int R0, int NR, int H, F&& emit)
{
    if (ay > by) {
        int64_t t = ax; ax = bx; bx = t;
        t = ay; ay = by; by = t;
    }
    int64_t jlo = sc_ceil_div(ay, 256) - 1, jhi = sc_floor_div(by, 256);
    if (jlo < 0) jlo = 0;
    if (jhi > H - 1) jhi = H - 1;
    if (jlo < R0) jlo = R0;
    if (jhi > R0 + NR - 1) jhi = R0 + NR - 1;
    if (jlo > jhi) return;
    if (ay == by) {
        const int64_t lo = ax < bx ? ax : bx, hi = ax < bx ? bx : ax;
        for (int64_t j = jlo; j <= jhi; j++) emit(j, lo, hi);
        return;
    }
    const int64_t den = by - ay, dx = bx - ax;
    auto at = [&](int64_t y, int64_t& q, int64_t& r) {
        const int64_t n = ax * den + dx * (y - ay);
        q = sc_floor_div(n, den);
        r = n - q * den;
    };
    int64_t ql, rl;  // x at the row's lower boundary: floor and remainder
    at(jlo * 256 > ay ? jlo * 256 : ay, ql, rl);
    const int64_t s = dx * 256;
    int64_t sq = 0, sr = 0;
    if (jlo < jhi) {
        sq = sc_floor_div(s, den);
        sr = s - sq * den;
    }
    for (int64_t j = jlo; j <= jhi; j++) {
        int64_t qu, ru;
        if (j < jhi && j > jlo) {
            qu = ql + sq;
            ru = rl + sr;
            if (ru >= den) {
                ru -= den;
                qu++;
            }
        } else {
            at((j + 1) * 256 < by ? (j + 1) * 256 : by, qu, ru);
        }
        const int64_t g0 = ql + (rl != 0 ? 1 : 0), g1 = qu + (ru != 0 ? 1 : 0);
        emit(j, ql < qu ? ql : qu, g0 > g1 ? g0 : g1);
        ql = qu;
        rl = ru;
    }
}

// Area fraction of a pixel from its accumulated coverage: |cov| capped at
// one pixel (exact for a simple ring; a self-overlapping ring's winding
// counts once), in units of kAreaOne.
RJ_HD uint32_t area_share(uint32_t acc)
{
    const int32_t v = (int32_t)acc;
    const int64_t a = v < 0 ? -(int64_t)v : (int64_t)v;
    return (uint32_t)(a < kAreaOne ? a : kAreaOne);
}

}  // namespace rj
