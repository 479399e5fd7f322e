// rj_impl.hpp -- host-side internals: polygon store, canvas plan, polygon
// scan-conversion bands and the CPUJoin grid index.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "rasterjoin.h"
#include "rj_dev.hpp"

namespace rj {

// A polygon collection in CSR form (PolyHandler::polys / pindexes,
// db/PolyHandler.cpp:49-75) with its bbox.
// A collection on the host: offsets and bounds (the vertices and the edge
// list live on the device only).
struct PolySet {
    int32_t np = 0;
    std::vector<int32_t> off;
    double bbox[4] = {0, 0, 0, 0};  // minx, miny, maxx, maxy (double of the float bounds)
    float bboxf[4] = {0, 0, 0, 0};
    std::vector<float> pbox;        // per-polygon float bounds: minx, miny, maxx, maxy
    // pbox from the vertices xy (2 * off[np] floats); false when a vertex is
    // not finite
    bool derive(const float* xy);
};

// Canvas geometry for a bbox and accuracy (GLFunction.cpp:26-30,
// RasterJoin.cpp:51-78 with a maximum tile side of 8192).
struct CanvasPlan {
    int32_t W = 0, H = 0, splitx = 0, splity = 0;
    double minx = 0, miny = 0, maxx = 0, maxy = 0;
    float minx_f = 0, miny_f = 0, maxx_f = 0, maxy_f = 0;
    float sx_f = 0, sy_f = 0;
    double fxs = 0, fys = 0;  // fixed-point scale: W*256/(maxx-minx)
    double cell = 0;
};
int make_plan(const double bbox[4], double eps_m, double ground_res, CanvasPlan* p);
DevCanvas dev_canvas(const CanvasPlan& p);

// Polygons on one canvas: per-polygon pixel regions and the bands the
// scan-conversion kernels run on. The snapped vertices themselves are
// computed on the device (the snap is monotone, so a region follows from the
// snapped float bounds of its polygon).
struct PolyRaster {
    std::vector<int32_t> region;      // c0, c1, r0, r1 per polygon (inclusive)
    std::vector<Band> bands_sum;      // BAND_SUM / BAND_LABEL
    // BAND_BOUNDS (full region width per band): the bands of at most
    // kBoundsSmallWords LDS words first (nbounds_small of them, launched with
    // lds_bounds_small), then the wider ones (launched with lds_bounds)
    std::vector<Band> bands_bounds;
    int nbounds_small = 0;
    size_t lds_sum = 0, lds_bounds = 0, lds_bounds_small = 0;
    bool bounds_ok = true;            // every region fits a BAND_BOUNDS band
};
// Regions and sum / label bands of every polygon on a canvas.
void rasterize_polys(const PolySet& ps, const CanvasPlan& plan, PolyRaster* out);
// The error-bounds bands (and bounds_ok) from the regions rasterize_polys set.
void bounds_bands(const PolySet& ps, PolyRaster* out);
// Snap one coordinate to the fixed-point canvas grid (round half to even).
int64_t snap_x(const CanvasPlan& plan, float x);
int64_t snap_y(const CanvasPlan& plan, float y);

// CPUJoin grid index (CPUJoin/GridIndexF.cpp:60-155): cell = y*xs + x, built
// on the device (rj_ctx::grid_device); the host keeps its geometry and the
// cells whose candidate list misses a polygon that meets the cell (possible
// through the reference's cell test), which the hybrid join treats as border
// pixels so that its result stays CPUJoin's.
//
// Edge bands (device arrays): horizontal bands of height band_h from miny
// (band_rows of them, band_row(), finer than the grid's rows for polygons
// with many edges per row). For polygon p and band y in [band_lo[p],
// band_hi[p]], the edges (vi, vj) whose y-span [min, max) holds the y of some
// point that band_row maps to y, as float4 (vi.x, vi.y, vj.x, vj.y) in
// edges[band_off[pband[p] + y - band_lo[p]] ..). The crossing test of a point
// in band y toggles only on those edges (every other edge fails its straddle
// test), so testing them gives the full loop's result with far fewer edges.
struct GridIndex {
    int32_t xs = 0, ys = 0;
    float minx = 0, miny = 0, width = 0, height = 0;
    std::vector<int32_t> deficient;
    float band_h = 0;
    int32_t band_rows = 0;
};
// Pixel rectangles (i0, i1, j0, j1) covering the deficient cells, one pixel margin.
void deficient_rects(const GridIndex& g, const CanvasPlan& plan, std::vector<int32_t>* rects);

// Fixed-point aggregate value decoding: sum = hi + lo * 2^-32.
inline double fixed_to_double(int64_t hi, uint64_t lo)
{
    return (double)hi + (double)lo * (1.0 / 4294967296.0);
}

}  // namespace rj
