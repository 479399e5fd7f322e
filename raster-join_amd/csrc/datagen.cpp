// datagen.cpp -- seeded synthetic inputs of the benchmark configurations:
//  * points: 70% from a mixture of isotropic Gaussians, 30% uniform, in a
//    square box; time (u32 over one year), fare (lognormal f32) and
//    passengers (u8 in 1..6) columns;
//  * polygons: Voronoi cells of uniform sites clipped to the box. Every edge
//    shared by two cells is subdivided and its interior vertices are jittered
//    perpendicular to the edge from a hash of the edge, so both cells carry
//    the exact same vertex chain (no gaps, no overlaps).
// Every random number is a pure function of (seed, stream, index), so any
// slice of the point table can be generated independently and identically.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <thread>
#include <tuple>
#include <vector>

#include "rasterjoin_datagen.h"

namespace {

inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline double u01(uint64_t seed, uint64_t stream, uint64_t i)
{
    uint64_t k = mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ULL + 0x8CB92BA72F3D8DD7ULL));
    return (double)(mix64(k ^ mix64(i)) >> 11) * 0x1.0p-53;
}

inline void gauss2(double u1, double u2, double* z0, double* z1)
{
    double r = std::sqrt(-2.0 * std::log(1.0 - u1));
    double a = 2.0 * M_PI * u2;
    *z0 = r * std::cos(a);
    *z1 = r * std::sin(a);
}

template <class F>
void parallel_for(int64_t n, F f)
{
    unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    if (n < 65536) nt = 1;
    std::vector<std::thread> th;
    int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        int64_t b = (int64_t)t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=] { f(b, e); });
    }
    for (auto& t : th) t.join();
}

float clamp_box(double v, double L)
{
    float f = (float)v;
    float top = std::nextafter((float)L, 0.0f);
    if (f < 0.0f) f = 0.0f;
    if (f > top) f = top;
    return f;
}

}  // namespace

extern "C" int rj_gen_points(const rj_point_params* P, int64_t first, int64_t n, float* xy,
                             uint32_t* time, float* fare, uint8_t* passengers)
{
    if (!P || n < 0 || !(P->box > 0) || P->ncenters < 1) return -1;
    const int K = P->ncenters;
    std::vector<double> cx(K), cy(K), sg(K);
    for (int k = 0; k < K; k++) {
        cx[k] = u01(P->seed, 100, 2 * (uint64_t)k) * P->box;
        cy[k] = u01(P->seed, 100, 2 * (uint64_t)k + 1) * P->box;
        sg[k] = P->sigma_lo + u01(P->seed, 101, (uint64_t)k) * (P->sigma_hi - P->sigma_lo);
    }
    const double L = P->box;
    parallel_for(n, [&](int64_t b, int64_t e) {
        for (int64_t r = b; r < e; r++) {
            uint64_t i = (uint64_t)(first + r);
            double x = 0, y = 0;
            bool placed = false;
            if (u01(P->seed, 1, i) < P->frac_cluster) {
                int k = std::min(K - 1, (int)(u01(P->seed, 2, i) * K));
                for (int a = 0; a < 64 && !placed; a++) {
                    double z0, z1;
                    gauss2(u01(P->seed, 16 + 2 * (uint64_t)a, i),
                           u01(P->seed, 17 + 2 * (uint64_t)a, i), &z0, &z1);
                    x = cx[k] + sg[k] * z0;
                    y = cy[k] + sg[k] * z1;
                    placed = (x >= 0 && x < L && y >= 0 && y < L);
                }
            }
            if (!placed) {
                x = u01(P->seed, 4, i) * L;
                y = u01(P->seed, 5, i) * L;
            }
            if (xy) {
                xy[2 * r] = clamp_box(x, L);
                xy[2 * r + 1] = clamp_box(y, L);
            }
            if (time)
                time[r] = P->t0 + (uint32_t)(u01(P->seed, 6, i) * (double)P->tspan);
            if (fare) {
                double z0, z1;
                gauss2(u01(P->seed, 7, i), u01(P->seed, 8, i), &z0, &z1);
                fare[r] = (float)std::exp(P->fare_mu + P->fare_sigma * z0);
            }
            if (passengers) passengers[r] = (uint8_t)(1 + std::min(5, (int)(u01(P->seed, 9, i) * 6)));
        }
    });
    return 0;
}

// ---------------------------------------------------------------------------
// Voronoi polygon sets
// ---------------------------------------------------------------------------

namespace {

// Edge labels: sites are >= 0, box sides are negative.
enum { SIDE_BOTTOM = -1, SIDE_RIGHT = -2, SIDE_TOP = -3, SIDE_LEFT = -4 };

struct CVtx {
    double x, y;
    int out;  // label of the edge leaving this vertex
};

// Canonical identity of a Voronoi vertex so both cells compute it identically.
struct VKey {
    int t, a, b, c;  // t: 0 corner, 1 bisector x side, 2 circumcenter
    bool operator<(const VKey& o) const { return std::tie(t, a, b, c) < std::tie(o.t, o.a, o.b, o.c); }
    bool operator==(const VKey& o) const { return t == o.t && a == o.a && b == o.b && c == o.c; }
};

VKey make_key(int site, int lin, int lout)
{
    VKey k{};
    if (lin < 0 && lout < 0) {
        k.t = 0; k.a = std::min(lin, lout); k.b = std::max(lin, lout); k.c = 0;
    } else if (lin < 0 || lout < 0) {
        int s = lin < 0 ? lout : lin, side = lin < 0 ? lin : lout;
        k.t = 1; k.a = std::min(site, s); k.b = std::max(site, s); k.c = side;
    } else {
        int v[3] = {site, lin, lout};
        std::sort(v, v + 3);
        k.t = 2; k.a = v[0]; k.b = v[1]; k.c = v[2];
    }
    return k;
}

void key_coords(const VKey& k, const std::vector<double>& sx, const std::vector<double>& sy,
                double L, double* x, double* y)
{
    if (k.t == 0) {
        bool left = (k.a == SIDE_LEFT || k.b == SIDE_LEFT);
        bool bottom = (k.a == SIDE_BOTTOM || k.b == SIDE_BOTTOM);
        *x = left ? 0.0 : L;
        *y = bottom ? 0.0 : L;
        return;
    }
    if (k.t == 1) {
        double ax = sx[k.a], ay = sy[k.a], bx = sx[k.b], by = sy[k.b];
        double mx = 0.5 * (ax + bx), my = 0.5 * (ay + by), dx = bx - ax, dy = by - ay;
        if (k.c == SIDE_BOTTOM || k.c == SIDE_TOP) {
            double yy = (k.c == SIDE_BOTTOM) ? 0.0 : L;
            *y = yy;
            *x = mx - (yy - my) * dy / dx;
        } else {
            double xx = (k.c == SIDE_LEFT) ? 0.0 : L;
            *x = xx;
            *y = my - (xx - mx) * dx / dy;
        }
    } else {
        double ax = sx[k.a], ay = sy[k.a];
        double bx = sx[k.b] - ax, by = sy[k.b] - ay, cx = sx[k.c] - ax, cy = sy[k.c] - ay;
        double d = 2.0 * (bx * cy - by * cx);
        double b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
        *x = ax + (cy * b2 - by * c2) / d;
        *y = ay + (bx * c2 - cx * b2) / d;
    }
    *x = std::min(std::max(*x, 0.0), L);
    *y = std::min(std::max(*y, 0.0), L);
}

std::vector<CVtx> clip_cell(int i, const std::vector<double>& sx, const std::vector<double>& sy,
                            double L, const std::vector<int>& order)
{
    std::vector<CVtx> poly = {{0, 0, SIDE_BOTTOM}, {L, 0, SIDE_RIGHT}, {L, L, SIDE_TOP}, {0, L, SIDE_LEFT}};
    double six = sx[i], siy = sy[i];
    for (int j : order) {
        if (j == i) continue;
        double dx = sx[j] - six, dy = sy[j] - siy;
        double half = 0.5 * std::sqrt(dx * dx + dy * dy);
        double rmax = 0;
        for (auto& v : poly) rmax = std::max(rmax, std::hypot(v.x - six, v.y - siy));
        if (half > rmax) break;  // sites are visited by distance: nothing further can cut
        double mx = six + 0.5 * dx, my = siy + 0.5 * dy;
        std::vector<CVtx> out;
        size_t n = poly.size();
        for (size_t k = 0; k < n; k++) {
            const CVtx& A = poly[k];
            const CVtx& B = poly[(k + 1) % n];
            double fa = (A.x - mx) * dx + (A.y - my) * dy;
            double fb = (B.x - mx) * dx + (B.y - my) * dy;
            bool ina = fa <= 0, inb = fb <= 0;
            if (ina) out.push_back(A);
            if (ina != inb) {
                double t = fa / (fa - fb);
                CVtx I{A.x + t * (B.x - A.x), A.y + t * (B.y - A.y), ina ? j : A.out};
                out.push_back(I);
            }
        }
        poly.swap(out);
        if (poly.size() < 3) break;
    }
    return poly;
}

bool seg_intersect(double ax, double ay, double bx, double by, double cx, double cy, double dx,
                   double dy)
{
    auto orient = [](double px, double py, double qx, double qy, double rx, double ry) {
        double v = (qx - px) * (ry - py) - (qy - py) * (rx - px);
        return (v > 0) - (v < 0);
    };
    auto onseg = [](double px, double py, double qx, double qy, double rx, double ry) {
        return std::min(px, qx) <= rx && rx <= std::max(px, qx) && std::min(py, qy) <= ry &&
               ry <= std::max(py, qy);
    };
    int o1 = orient(ax, ay, bx, by, cx, cy), o2 = orient(ax, ay, bx, by, dx, dy);
    int o3 = orient(cx, cy, dx, dy, ax, ay), o4 = orient(cx, cy, dx, dy, bx, by);
    if (o1 != o2 && o3 != o4) return true;
    if (o1 == 0 && onseg(ax, ay, bx, by, cx, cy)) return true;
    if (o2 == 0 && onseg(ax, ay, bx, by, dx, dy)) return true;
    if (o3 == 0 && onseg(cx, cy, dx, dy, ax, ay)) return true;
    if (o4 == 0 && onseg(cx, cy, dx, dy, bx, by)) return true;
    return false;
}

bool is_simple(const float* v, int n)
{
    if (n < 3) return false;
    std::vector<float> bx0(n), bx1(n), by0(n), by1(n);
    for (int k = 0; k < n; k++) {
        int l = (k + 1) % n;
        bx0[k] = std::min(v[2 * k], v[2 * l]); bx1[k] = std::max(v[2 * k], v[2 * l]);
        by0[k] = std::min(v[2 * k + 1], v[2 * l + 1]); by1[k] = std::max(v[2 * k + 1], v[2 * l + 1]);
    }
    for (int k = 0; k < n; k++) {
        int k1 = (k + 1) % n;
        for (int l = k + 2; l < n; l++) {
            if (k == 0 && l == n - 1) continue;  // adjacent through the closing edge
            if (bx1[k] < bx0[l] || bx1[l] < bx0[k] || by1[k] < by0[l] || by1[l] < by0[k]) continue;
            int l1 = (l + 1) % n;
            if (seg_intersect(v[2 * k], v[2 * k + 1], v[2 * k1], v[2 * k1 + 1], v[2 * l], v[2 * l + 1],
                              v[2 * l1], v[2 * l1 + 1]))
                return false;
        }
    }
    return true;
}

struct PolySetBuf {
    std::vector<float> xy;
    std::vector<int32_t> off;
};

}  // namespace

struct rj_polys {
    PolySetBuf b;
};

extern "C" int rj_gen_voronoi(const rj_voronoi_params* P, rj_polys** out)
{
    if (!P || !out || P->nsites < 1 || !(P->box > 0)) return -1;
    const int S = P->nsites;
    const double L = P->box;
    for (int attempt = 0; attempt < 8; attempt++) {
        uint64_t sseed = P->seed + 0x1000 * (uint64_t)attempt;
        std::vector<double> sx(S), sy(S);
        for (int i = 0; i < S; i++) {
            sx[i] = u01(sseed, 200, 2 * (uint64_t)i) * L;
            sy[i] = u01(sseed, 200, 2 * (uint64_t)i + 1) * L;
        }
        // Cells by half-plane clipping, visiting other sites by distance.
        std::vector<std::vector<CVtx>> cells(S);
        parallel_for(S * (int64_t)1, [&](int64_t b, int64_t e) {
            std::vector<std::pair<double, int>> d(S);
            std::vector<int> order(S);
            for (int64_t i = b; i < e; i++) {
                for (int j = 0; j < S; j++) d[j] = {std::hypot(sx[j] - sx[i], sy[j] - sy[i]), j};
                std::sort(d.begin(), d.end());
                for (int j = 0; j < S; j++) order[j] = d[j].second;
                cells[i] = clip_cell((int)i, sx, sy, L, order);
            }
        });
        // Canonical vertex keys and the shared-edge consistency check.
        std::vector<std::vector<VKey>> keys(S);
        std::map<std::tuple<VKey, VKey>, int> edge_owner;
        bool ok = true;
        for (int i = 0; i < S && ok; i++) {
            auto& c = cells[i];
            int n = (int)c.size();
            if (n < 3) { ok = false; break; }
            keys[i].resize(n);
            for (int k = 0; k < n; k++) keys[i][k] = make_key(i, c[(k + n - 1) % n].out, c[k].out);
            for (int k = 0; k < n; k++) {
                if (c[k].out < 0) continue;
                edge_owner[{keys[i][k], keys[i][(k + 1) % n]}] = i;
            }
        }
        if (ok) {
            for (auto& kv : edge_owner) {
                auto rev = edge_owner.find({std::get<1>(kv.first), std::get<0>(kv.first)});
                if (rev == edge_owner.end()) { ok = false; break; }
            }
        }
        if (!ok) continue;

        // Subdivide + jitter; edges of non-simple polygons lose their jitter.
        std::map<std::tuple<VKey, VKey>, bool> nojit;
        std::vector<std::vector<float>> fpoly(S);
        for (int pass = 0; pass < 4; pass++) {
            bool all_simple = true;
            std::vector<char> bad(S, 0);
            parallel_for(S, [&](int64_t b, int64_t e) {
                for (int64_t i = b; i < e; i++) {
                    auto& c = cells[i];
                    int n = (int)c.size();
                    std::vector<double> vx, vy;
                    for (int k = 0; k < n; k++) {
                        VKey k1 = keys[i][k], k2 = keys[i][(k + 1) % n];
                        double x1, y1, x2, y2;
                        key_coords(k1, sx, sy, L, &x1, &y1);
                        key_coords(k2, sx, sy, L, &x2, &y2);
                        vx.push_back(x1);
                        vy.push_back(y1);
                        bool fwd = k1 < k2;
                        VKey ka = fwd ? k1 : k2, kb = fwd ? k2 : k1;
                        double ax = fwd ? x1 : x2, ay = fwd ? y1 : y2, bx2 = fwd ? x2 : x1, by2 = fwd ? y2 : y1;
                        double len = std::hypot(bx2 - ax, by2 - ay);
                        uint64_t eh = mix64(((uint64_t)(uint32_t)ka.a << 32) ^ (uint32_t)ka.b) ^
                                      mix64(((uint64_t)(uint32_t)ka.c << 40) ^ ((uint64_t)ka.t << 36) ^
                                            ((uint64_t)(uint32_t)kb.a << 18) ^ (uint32_t)kb.b) ^
                                      mix64(((uint64_t)(uint32_t)kb.c << 8) ^ (uint64_t)kb.t);
                        double sp = P->spacing_lo;
                        if (P->spacing_hi > P->spacing_lo)
                            sp = std::exp(std::log(P->spacing_lo) +
                                          u01(P->seed, 300, eh) * (std::log(P->spacing_hi) - std::log(P->spacing_lo)));
                        int64_t kseg = std::max<int64_t>(1, std::llround(len / sp));
                        kseg = std::min<int64_t>(kseg, 100000);
                        bool shared = c[k].out >= 0;
                        bool jit = shared && P->jitter > 0 && !nojit.count({ka, kb});
                        double amp = std::min(P->jitter, sp);
                        double nx = len > 0 ? -(by2 - ay) / len : 0, ny = len > 0 ? (bx2 - ax) / len : 0;
                        std::vector<double> ix, iy;
                        for (int64_t m = 1; m < kseg; m++) {
                            double t = (double)m / (double)kseg;
                            double px = ax + t * (bx2 - ax), py = ay + t * (by2 - ay);
                            if (jit) {
                                double a = std::min(amp, 0.2 * std::min(t, 1.0 - t) * len);
                                double off = (2.0 * u01(P->seed, 301, eh * 1000003ULL + (uint64_t)m) - 1.0) * a;
                                px += off * nx;
                                py += off * ny;
                            }
                            ix.push_back(std::min(std::max(px, 0.0), L));
                            iy.push_back(std::min(std::max(py, 0.0), L));
                        }
                        if (!fwd) {
                            std::reverse(ix.begin(), ix.end());
                            std::reverse(iy.begin(), iy.end());
                        }
                        vx.insert(vx.end(), ix.begin(), ix.end());
                        vy.insert(vy.end(), iy.begin(), iy.end());
                    }
                    std::vector<float> fv;
                    for (size_t k = 0; k < vx.size(); k++) {
                        float fx = (float)vx[k], fy = (float)vy[k];
                        size_t m = fv.size();
                        if (m >= 2 && fv[m - 2] == fx && fv[m - 1] == fy) continue;
                        fv.push_back(fx);
                        fv.push_back(fy);
                    }
                    while (fv.size() >= 4 && fv[0] == fv[fv.size() - 2] && fv[1] == fv[fv.size() - 1]) {
                        fv.pop_back();
                        fv.pop_back();
                    }
                    fpoly[i].swap(fv);
                    if (!is_simple(fpoly[i].data(), (int)fpoly[i].size() / 2)) bad[i] = 1;
                }
            });
            for (int i = 0; i < S; i++) {
                if (!bad[i]) continue;
                all_simple = false;
                int n = (int)cells[i].size();
                for (int k = 0; k < n; k++) {
                    VKey k1 = keys[i][k], k2 = keys[i][(k + 1) % n];
                    nojit[{std::min(k1, k2), std::max(k1, k2)}] = true;
                }
            }
            if (all_simple) {
                rj_polys* r = new rj_polys();
                r->b.off.push_back(0);
                for (int i = 0; i < S; i++) {
                    r->b.xy.insert(r->b.xy.end(), fpoly[i].begin(), fpoly[i].end());
                    r->b.off.push_back((int32_t)(r->b.xy.size() / 2));
                }
                *out = r;
                return 0;
            }
        }
    }
    return -2;
}

extern "C" int rj_polys_size(const rj_polys* p, int32_t* np, int64_t* nverts)
{
    if (!p) return -1;
    if (np) *np = (int32_t)p->b.off.size() - 1;
    if (nverts) *nverts = (int64_t)p->b.xy.size() / 2;
    return 0;
}

extern "C" int rj_polys_copy(const rj_polys* p, float* xy, int32_t* off)
{
    if (!p) return -1;
    if (xy) std::memcpy(xy, p->b.xy.data(), p->b.xy.size() * sizeof(float));
    if (off) std::memcpy(off, p->b.off.data(), p->b.off.size() * sizeof(int32_t));
    return 0;
}

extern "C" void rj_polys_free(rj_polys* p) { delete p; }
