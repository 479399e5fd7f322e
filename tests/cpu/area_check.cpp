// Host build of the error-bounds kernel's per-pixel area sums
// (csrc/scan_ops.hpp edge_area, deposited and prefix-summed the way k_band
// does for a band that spans the whole canvas). Reads "W H n" and n snapped
// vertices (x y, subunits) from stdin; writes the W*H signed area sums
// (int32, 2^-24 of a pixel) to stdout as raw bytes.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "scan_ops.hpp"

int main()
{
    long long W, H, n;
    if (std::scanf("%lld %lld %lld", &W, &H, &n) != 3 || W <= 0 || H <= 0 || n < 3) return 2;
    std::vector<long long> X(n), Y(n);
    for (long long i = 0; i < n; i++)
        if (std::scanf("%lld %lld", &X[i], &Y[i]) != 2) return 2;
    std::vector<uint32_t> acc((size_t)(W * H), 0u);
    for (long long i = 0, j = n - 1; i < n; j = i++)
        rj::edge_area(X[i], Y[i], X[j], Y[j], 0, (int)H, [&](int64_t row, int64_t c, uint32_t d) {
            int64_t l = c < 0 ? 0 : c;
            if (l >= W) return;
            acc[(size_t)(row * W + l)] += d;
        });
    for (long long r = 0; r < H; r++) {
        uint32_t run = 0;
        for (long long c = 0; c < W; c++) {
            run += acc[(size_t)(r * W + c)];
            acc[(size_t)(r * W + c)] = run;
        }
    }
    std::fwrite(acc.data(), sizeof(uint32_t), acc.size(), stdout);
    return 0;
}
