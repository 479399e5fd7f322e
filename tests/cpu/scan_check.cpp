// Host check of the scan conversion's integer steps (csrc/scan_ops.hpp):
// edge_crossings' quotient/remainder stepping against the crossing column
// computed directly for every row in 128-bit arithmetic, and bits8 against
// the bit-by-bit extraction, for random and extreme snapped edges and every
// column offset in [-7, bw). Prints "ok" and the number of rows checked, or
// the first mismatch (exit 1).
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "scan_ops.hpp"

#include <algorithm>
#include <cstdint>

static __int128 ceil_div128(__int128 a, __int128 b)
{
    __int128 q = a / b;
    if ((a % b != 0) && ((a > 0) == (b > 0))) q++;
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
    std::printf("ok %ld rows %ld groups\n", rows, groups);
    return 0;
}
