// Host check of the scan conversion's integer steps (csrc/scan_ops.hpp):
// edge_crossings' quotient/remainder stepping against the crossing column
// computed directly for every row in 128-bit arithmetic, edge_touch_rows and
// edge_area against the same values divided afresh for every row, and bits8
// against the bit-by-bit extraction, for random and extreme snapped edges and
// every column offset in [-7, bw). Prints "ok" and the counts checked, or the
// first mismatch (exit 1).
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "scan_ops.hpp"

#include <algorithm>
#include <array>
#include <cstdint>

static __int128 ceil_div128(__int128 a, __int128 b)
{
    __int128 q = a / b;
    if ((a % b != 0) && ((a > 0) == (b > 0))) q++;
    return q;
}

static __int128 floor_div128(__int128 a, __int128 b)
{
    __int128 q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}

static int check_edge(int64_t ax, int64_t ay, int64_t bx, int64_t by, int64_t cs, int R0, int NR, long* rows)
{
    std::vector<std::pair<int64_t, int64_t>> got;
    rj::edge_crossings(ax, ay, bx, by, cs, R0, NR, [&](int64_t row, int64_t k) { got.push_back({row, k}); });
    // direct: rows whose center y lies in [min y, max y), one division each
    std::vector<std::pair<int64_t, int64_t>> want;
    if (ay != by) {
        int64_t x0 = ax, y0 = ay, x1 = bx, y1 = by;
        if (y0 > y1) {
            std::swap(x0, x1);
            std::swap(y0, y1);
        }
        const int64_t h = cs / 2;
        for (int64_t row = R0; row < (int64_t)R0 + NR; row++) {
            const int64_t yc = row * cs + h;
            if (yc < y0 || yc >= y1) continue;
            // crossing x = x0 + (x1 - x0) * (yc - y0) / (y1 - y0); first column
            // whose center (col*cs + h) is at or right of it
            const __int128 dy = y1 - y0, dx = x1 - x0;
            const __int128 num = (__int128)x0 * dy + dx * (yc - y0) - (__int128)h * dy;
            want.push_back({row, (int64_t)ceil_div128(num, (__int128)cs * dy)});
        }
    }
    *rows += (long)want.size();
    if (got != want) {
        std::printf("mismatch: edge (%lld,%lld)-(%lld,%lld) cs %lld rows [%d,+%d): %zu vs %zu rows\n", (long long)ax,
                    (long long)ay, (long long)bx, (long long)by, (long long)cs, R0, NR, got.size(), want.size());
        for (size_t i = 0; i < got.size() && i < want.size(); i++)
            if (got[i] != want[i]) {
                std::printf("  row %lld: stepped %lld, direct %lld\n", (long long)got[i].first, (long long)got[i].second,
                            (long long)want[i].second);
                break;
            }
        return 1;
    }
    return 0;
}

// edge_touch_rows against the x-extent computed with one division per row
// boundary (floor and ceiling of num / den at both ends of the row's piece).
static int check_touch(int64_t ax, int64_t ay, int64_t bx, int64_t by, int R0, int NR, int H, long* rows)
{
    std::vector<std::array<int64_t, 3>> got, want;
    rj::edge_touch_rows(ax, ay, bx, by, R0, NR, H,
                        [&](int64_t j, int64_t lo, int64_t hi) { got.push_back({j, lo, hi}); });
    int64_t x0 = ax, y0 = ay, x1 = bx, y1 = by;
    if (y0 > y1) {
        std::swap(x0, x1);
        std::swap(y0, y1);
    }
    int64_t jlo = rj::sc_ceil_div(y0, 256) - 1, jhi = rj::sc_floor_div(y1, 256);
    jlo = std::max<int64_t>({jlo, 0, R0});
    jhi = std::min<int64_t>({jhi, H - 1, (int64_t)R0 + NR - 1});
    for (int64_t j = jlo; j <= jhi; j++) {
        int64_t lo, hi;
        if (y0 == y1) {
            lo = std::min(x0, x1);
            hi = std::max(x0, x1);
        } else {
            const __int128 den = y1 - y0, dx = x1 - x0;
            const int64_t ylo = std::max(y0, j * 256), yhi = std::min(y1, (j + 1) * 256);
            const __int128 n0 = (__int128)x0 * den + dx * (ylo - y0), n1 = (__int128)x0 * den + dx * (yhi - y0);
            const int64_t f0 = (int64_t)floor_div128(n0, den), f1 = (int64_t)floor_div128(n1, den);
            const int64_t g0 = (int64_t)ceil_div128(n0, den), g1 = (int64_t)ceil_div128(n1, den);
            lo = std::min(f0, f1);
            hi = std::max(g0, g1);
        }
        want.push_back({j, lo, hi});
    }
    *rows += (long)want.size();
    if (got != want) {
        std::printf("touch mismatch: edge (%lld,%lld)-(%lld,%lld) rows [%d,+%d) H %d: %zu vs %zu rows\n", (long long)ax,
                    (long long)ay, (long long)bx, (long long)by, R0, NR, H, got.size(), want.size());
        return 1;
    }
    return 0;
}

// edge_area against the same deposits computed with the divisions of every
// row done afresh (the double quotients at both ends and the columns by
// 128-bit floor division).
static int check_area(int64_t ax, int64_t ay, int64_t bx, int64_t by, int R0, int NR, long* rows)
{
    std::vector<std::array<int64_t, 3>> got, want;
    rj::edge_area(ax, ay, bx, by, R0, NR,
                  [&](int64_t row, int64_t c, uint32_t d) { got.push_back({row, c, (int64_t)d}); });
    if (ay != by) {
        uint32_t sgn = 1u;
        int64_t x0 = ax, y0 = ay, x1 = bx, y1 = by;
        if (y0 > y1) {
            std::swap(x0, x1);
            std::swap(y0, y1);
            sgn = 0xFFFFFFFFu;
        }
        int64_t jlo = std::max<int64_t>(rj::sc_floor_div(y0, 256), R0);
        int64_t jhi = std::min<int64_t>(rj::sc_ceil_div(y1, 256) - 1, (int64_t)R0 + NR - 1);
        const int64_t den = y1 - y0, dx = x1 - x0;
        const double dden = (double)den;
        for (int64_t row = jlo; row <= jhi; row++) {
            const int64_t ylo = std::max(row * 256, y0), yhi = std::min(row * 256 + 256, y1);
            const int64_t h = yhi - ylo;
            const int64_t n0 = x0 * den + dx * (ylo - y0), n1 = x0 * den + dx * (yhi - y0);
            const double u0 = (double)n0 / dden, u1 = (double)n1 / dden;
            const double lo = u0 < u1 ? u0 : u1, hi = u0 < u1 ? u1 : u0;
            const double mid = 0.5 * (u0 + u1), hd = (double)h;
            const double k = hi > lo ? hd / (2.0 * (hi - lo)) : 0.0;
            const int64_t cmin = (int64_t)floor_div128(std::min(n0, n1), (__int128)256 * den);
            const int64_t cmax = (int64_t)floor_div128(std::max(n0, n1), (__int128)256 * den);
            int64_t gprev = 0, fprev = 0;
            for (int64_t c = cmin; c <= cmax; c++) {
                const int64_t g = rj::area_left((double)((c + 1) * 256), lo, hi, mid, hd, k);
                const int64_t f = g - gprev;
                want.push_back({row, c, (int64_t)((uint32_t)(f - fprev) * sgn)});
                fprev = f;
                gprev = g;
            }
            want.push_back({row, cmax + 1, (int64_t)((uint32_t)(65536 * h - fprev) * sgn)});
        }
    }
    *rows += (long)want.size();
    if (got != want) {
        std::printf("area mismatch: edge (%lld,%lld)-(%lld,%lld) rows [%d,+%d): %zu vs %zu deposits\n", (long long)ax,
                    (long long)ay, (long long)bx, (long long)by, R0, NR, got.size(), want.size());
        return 1;
    }
    return 0;
}

int main()
{
    std::mt19937_64 rng(12345);
    long rows = 0;
    const int64_t sub = 256;
    for (int64_t cs : {sub, sub / 4}) {
        // random edges on canvases up to 8192 and up to 2^23 pixels a side
        for (int it = 0; it < 200000; it++) {
            const int64_t span = (it % 3 == 0) ? ((int64_t)1 << 31) - 1 : (it % 3 == 1 ? 8192 * sub : 64 * sub);
            std::uniform_int_distribution<int64_t> c(0, span - 1);
            int64_t ax = c(rng), ay = c(rng), bx, by;
            if (it % 5 == 0) {  // short edges near a vertex
                std::uniform_int_distribution<int64_t> d(-3 * cs, 3 * cs);
                bx = std::max<int64_t>(0, std::min(span - 1, ax + d(rng)));
                by = std::max<int64_t>(0, std::min(span - 1, ay + d(rng)));
            } else {
                bx = c(rng);
                by = c(rng);
            }
            if (it % 7 == 0) ay = (ay / cs) * cs + cs / 2;  // a vertex on a row center
            if (it % 11 == 0) bx = ax;                      // vertical
            const int64_t rmin = std::min(ay, by) / cs;
            std::uniform_int_distribution<int64_t> rr(std::max<int64_t>(0, rmin - 5), rmin + 5);
            const int R0 = (int)rr(rng);
            const int NR = (it % 64 == 0) ? 20000 : 1 + (int)(rng() % 64);
            if (check_edge(ax, ay, bx, by, cs, R0, NR, &rows)) return 1;
        }
        // extremes: the full canvas diagonal, horizontal, one subunit high
        const int64_t M = ((int64_t)1 << 31) - 1;
        const int64_t ex[][4] = {{0, 0, M, M}, {M, 0, 0, M}, {0, 5, M, 5}, {7, 0, 9, 1}, {0, M - 1, M, M},
                                 {M, M, 0, 0}, {3, 0, 3, M}, {123456, 128, 654321, 129}};
        for (auto& e : ex)
            for (int R0 : {0, 1000, 8388000})
                if (check_edge(e[0], e[1], e[2], e[3], cs, R0, 20000, &rows)) return 1;
    }
    // edge_touch_rows and edge_area against their per-row divisions: random
    // edges (negative coordinates too: polygons may leave the canvas), short
    // edges, vertices on row boundaries, horizontal and vertical edges, row
    // ranges clipped on either side
    long trows = 0, adeps = 0;
    for (int it = 0; it < 100000; it++) {
        const int64_t span = (it % 3 == 0) ? ((int64_t)1 << 24) : (it % 3 == 1 ? 8192 * sub : 64 * sub);
        std::uniform_int_distribution<int64_t> c(-span / 8, span - 1);
        int64_t ax = c(rng), ay = c(rng), bx, by;
        if (it % 5 == 0) {
            std::uniform_int_distribution<int64_t> d(-3 * sub, 3 * sub);
            bx = ax + d(rng);
            by = ay + d(rng);
        } else {
            bx = c(rng);
            by = c(rng);
        }
        if (it % 7 == 0) ay = rj::sc_floor_div(ay, 256) * 256;  // a vertex on a row boundary
        if (it % 13 == 0) by = ay;                               // horizontal
        if (it % 11 == 0) bx = ax;                               // vertical
        const int64_t rmin = rj::sc_floor_div(std::min(ay, by), 256);
        std::uniform_int_distribution<int64_t> rr(rmin - 5, rmin + 5);
        const int R0 = (int)rr(rng);
        const int NR = (it % 64 == 0) ? 20000 : 1 + (int)(rng() % 64);
        const int H = (it % 4 == 0) ? (int)std::max<int64_t>(1, rmin + (int64_t)(rng() % 16)) : 1 << 24;
        if (check_touch(ax, ay, bx, by, R0, NR, H, &trows)) return 1;
        if (span <= 8192 * sub && check_area(ax, ay, bx, by, R0, NR, &adeps)) return 1;
    }
    // bits8 against bit-by-bit extraction
    long groups = 0;
    for (int it = 0; it < 20000; it++) {
        const int wpr = 1 + (int)(rng() % 8);
        std::vector<uint32_t> row(wpr);
        for (auto& w : row) w = (uint32_t)rng();
        const int bw = 1 + (int)(rng() % (32 * wpr));
        for (int col = -7; col < bw; col++) {
            uint32_t want = 0;
            for (int k = 0; k < 8; k++) {
                const int c = col + k;
                if (c >= 0 && c < bw && ((row[c >> 5] >> (c & 31)) & 1u)) want |= 1u << k;
            }
            const uint32_t got = rj::bits8(row.data(), col, bw, wpr);
            groups++;
            if (got != want) {
                std::printf("bits8 mismatch: wpr %d bw %d col %d: %02x vs %02x\n", wpr, bw, col, got, want);
                return 1;
            }
        }
    }
    std::printf("ok %ld rows %ld groups %ld touch rows %ld area deposits\n", rows, groups, trows, adeps);
    return 0;
}
