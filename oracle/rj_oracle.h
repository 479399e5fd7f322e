/* rj_oracle.h -- CPU restatement of the raster-join hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline. The product (raster-join_amd/) never links it.
 *
 * What each group restates (reference paths relative to the reference root):
 *   orc_grid_*      CPUJoin/GridIndexF.cpp:39-175,383-549 (buildGrid, getRegion,
 *                   isInsidePoly, polygonRectIntersectionCPU and helpers)
 *   orc_join_exact  CPUJoin/CPUJoin.cpp:62-122 (per-point getRegion, per-thread
 *                   histograms, reduce) plus the filter/avg semantics of
 *                   RasterJoin/shaders/index.glsl:69-243 with integer compares for
 *                   Uint/Int columns
 *   orc_plan_make   RasterJoin/join/GLFunction.cpp:26-30, RasterJoin.cpp:51-78
 *   orc_raster_join RasterJoin.cpp:91-243, points.vert:138-159, polygon.frag:12-28
 *                   (canvas splat + per-polygon sum of pixel-center coverage)
 *   orc_errorbounds RasterJoinBounds.cpp:306-381, line-bound.frag:219-262
 *
 * Parity pins (what checks each group) are listed in tests/ and in the
 * "Oracle" section of the design notes at the repository root.
 */
#ifndef RJ_ORACLE_H
#define RJ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 16-byte constraint record, same layout as QueryConstraint (common/Common.h:34-39). */
typedef struct orc_constraint {
    int32_t attrib_id; /* index into the cols[] array */
    int32_t type;      /* 0 EQ, 1 LT, 2 LTE, 3 GT, 4 GTE */
    union { int32_t ival; float fval; uint32_t uval; } val;
    int32_t dummy;
} orc_constraint;

/* column types: 1 Uint(u32) 2 Int(i32) 3 Float(f32) 4 Uint stored as u8 */

/* Point input shared by the joins. */
typedef struct orc_points {
    int64_t n;
    const float* xy;            /* 2n floats, interleaved */
    int32_t ncols;
    const void* const* cols;    /* ncols column pointers */
    const int32_t* types;       /* ncols types */
    int32_t ncons;
    const orc_constraint* cons;
    int32_t agg_type;           /* 0 count, 1 avg, 2 sum */
    int32_t agg_col;            /* column index for avg/sum */
} orc_points;

/* ---- CPUJoin grid index ---- */
typedef struct orc_grid orc_grid;
orc_grid* orc_grid_build(const float* xy, const int32_t* off, int32_t np,
                         const float bbox[4], int32_t xs, int32_t ys, int32_t nthreads);
int64_t orc_grid_ncand(const orc_grid* g);
/* cell_off: xs*ys+1 entries (cell index = y*xs + x); cand: ncand entries, sorted per cell */
void orc_grid_export(const orc_grid* g, int32_t* cell_off, int32_t* cand);
void orc_grid_free(orc_grid* g);

/* Exact join (CPUJoin semantics). Outputs np-sized arrays.
 * sum_hi/sum_lo: exact fixed-point sums (value = hi + lo * 2^-32 summed),
 * sum_f64: plain double sums in point order (for the relative-error check).
 * Any output pointer may be NULL. Returns 0 on success. */
int orc_join_exact(const orc_grid* g, const orc_points* pts, int32_t nthreads,
                   int64_t* counts, int64_t* sum_hi, uint64_t* sum_lo, double* sum_f64);

/* Brute force over every polygon (no grid), same float PIP. */
int orc_join_brute(const float* xy, const int32_t* off, int32_t np, const orc_points* pts,
                   int64_t* counts);

/* ---- canvas geometry ---- */
typedef struct orc_plan {
    int32_t W, H, splitx, splity;
    double minx, miny, maxx, maxy;
    float minx_f, miny_f, maxx_f, maxy_f;
    float sx_f, sy_f;       /* pixels per unit */
    double fxs, fys;        /* fixed-point scale: W*256/(maxx-minx) */
    double cell;            /* requested cell size in units */
} orc_plan;

int orc_plan_make(const double bbox[4], double eps_m, double ground_res, orc_plan* p);
/* pixel of a point, -1 when outside the canvas */
int64_t orc_pixel(const orc_plan* p, float x, float y);
/* snap vertices to the 8-bit subpixel canvas grid */
void orc_snap(const orc_plan* p, const float* xy, int64_t nv, int64_t* X, int64_t* Y);

/* Coverage of one polygon on the canvas: out[j*W+i] = 1 when the pixel center is
 * inside. method 0 = per-pixel crossing test, 1 = edge-flag scanline. */
int orc_coverage(const orc_plan* p, const float* xy, int32_t nv, int32_t method, uint8_t* out);

/* Raster join: canvas splat then per-polygon sums. canvas_cnt (W*H u32) may be
 * passed in to receive the canvas; it may be NULL. */
int orc_raster_join(const orc_plan* p, const float* pxy, const int32_t* off, int32_t np,
                    const orc_points* pts, int64_t* counts, int64_t* sum_hi, uint64_t* sum_lo,
                    uint32_t* canvas_cnt);

/* Boundary pixels of one polygon (touched by an edge, dilated by r pixels). */
int orc_boundary(const orc_plan* p, const float* xy, int32_t nv, int32_t r, uint8_t* out);

/* Area of one ring inside each pixel: out[j*W+i] = signed area in units of
 * 2^-24 of a pixel (positive for a counter-clockwise ring), exact up to one unit per edge
 * crossing the pixel's row. */
int orc_area(const orc_plan* p, const float* xy, int32_t nv, int32_t* out);

/* Errorbounds: bounds = [fn | fp | fn1 | fp1] (4*np int64) from a count canvas. */
int orc_errorbounds(const orc_plan* p, const float* pxy, const int32_t* off, int32_t np,
                    const uint32_t* canvas_cnt, int64_t* bounds);

#ifdef __cplusplus
}
#endif
#endif
