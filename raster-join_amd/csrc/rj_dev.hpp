// rj_dev.hpp -- plain argument structs shared by the host engine and the HIP
// kernels, and the kernel launch entry points (defined in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rasterjoin.h"

namespace rj {

constexpr int kSub = 256;          // subpixel units per pixel of the fixed-point canvas grid
constexpr int kSubShift = 8;
constexpr uint32_t kLabelNone = 0xFFFFFFFFu;    // pixel center in no polygon
constexpr uint32_t kLabelBorder = 0xFFFFFFFEu;  // pixel handled by the exact test
constexpr uint32_t kLabelMixed = 0xFFFFFFFDu;   // tile map: the tile holds several labels
constexpr uint32_t kLabelNoneHalo = 0xFFFFFFFCu;  // hybrid map: no polygon, a touched pixel next to it
constexpr uint32_t kLabelMaxId = 0xFFFFFFF0u;
// hybrid maps in the halo mode flag the labels of pixels next to a touched
// pixel with this bit (16-bit maps: 0x8000, so ids < 0x7FF0 there)
constexpr uint32_t kHaloFlag16 = 0x8000u, kHaloFlag32 = 0x40000000u;
constexpr int kTileShiftDefault = 4;              // label tiles of 16 x 16 pixels (3..5 allowed)
constexpr int kLabelRowAlign = 32;               // label-map rows are padded to 32 pixels

// Columns of a point table, as seen by a kernel.
struct DevCols {
    const void* col[RJ_MAX_COLUMNS];
    int32_t type[RJ_MAX_COLUMNS];
    // bounds of the table's locations (minx, miny, maxx, maxy); box_ok: every
    // location is finite and inside them
    float box[4];
    int32_t box_ok;
    uint32_t in_range;  // bit c: float column c is finite and in [-2^31, 2^31) on every row
};

constexpr int kMaxSlots = 5;  // distinct columns one query may read (GLFunction.cpp:250-253)

// Order-preserving 32-bit key of a typed column value: a < b as the type
// compares them (Uint and u8 unsigned, Int signed, Float as IEEE floats)
// iff key(a) < key(b) unsigned. Float NaNs fall outside [key(-inf),
// key(+inf)] (negative NaNs below, positive above); -0 and +0 are adjacent.
__host__ __device__ inline uint32_t col_key(int type, uint32_t bits)
{
    if (type == RJ_ATTR_FLOAT) return (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);
    if (type == RJ_ATTR_INT) return bits ^ 0x80000000u;
    return bits;
}

// Filters and aggregate of one query. The columns it reads are numbered as
// slots; the aggregate names its slot. The constraints on a slot's column
// (a conjunction) become one key range: a row passes when slot_lo[s] <=
// col_key(type, value) <= slot_hi[s] for every slot (lo > hi: no row).
struct DevQuery {
    int32_t agg_type;  // RJ_AGG_*
    int32_t nslots;
    int32_t slot_col[kMaxSlots];
    uint32_t slot_lo[kMaxSlots];
    uint32_t slot_hi[kMaxSlots];
    int32_t agg_slot;  // -1 for count
    int32_t plain;     // 1: no slot has constraints
};

// Canvas geometry (same numbers as the host plan).
struct DevCanvas {
    int32_t W, H;
    int32_t LW;         // row stride of label maps (W rounded up to kLabelRowAlign)
    float minx, miny, maxx, maxy;
    float sx, sy;
};

// CPUJoin grid index on the device (cell = y*xs + x), with the row bands of
// the polygons' edges (the edges a point of a grid row can cross).
struct DevGrid {
    int32_t xs, ys;
    float minx, miny, width, height;
    const int32_t* cell_off;
    const int32_t* cand;
    float band_h;             // edge bands: height and count (band_row)
    int32_t band_rows;
    const int32_t* band_lo;   // per polygon: first and last band with edges
    const int32_t* band_hi;
    const int32_t* pband;     // per polygon: its first band
    const int32_t* band_off;  // per band: first edge
    const float4* edges;      // (vi.x, vi.y, vj.x, vj.y)
};

// Edge band of a y (GridIndex): the same float operations on the host and
// the device; monotone in t, clamped to [0, rows).
__host__ __device__ inline int band_row(float t, float miny, float band_h, int rows)
{
    const float fy = (t - miny) / band_h;
    if (!(fy >= 0.0f)) return 0;
    if (!(fy < (float)rows)) return rows - 1;
    return (int)fy;
}

// Polygon collection on the device.
struct DevPolys {
    int32_t np;
    const float* xy;       // float vertices (CSR), 2 per vertex
    const int32_t* off;    // np+1
    const int32_t* X;      // snapped canvas coordinates (subpixel units)
    const int32_t* Y;
    const int32_t* region; // 4 per polygon: c0, c1, r0, r1 (inclusive pixel bbox)
};

// Per-polygon accumulators.
struct DevResult {
    unsigned long long* cnt;
    long long* hi;
    unsigned long long* lo;
};

// One band of rows of one polygon for the scan-conversion kernels.
struct Band {
    int32_t poly;
    int32_t r0;      // first row
    int32_t nrows;
    int32_t wpr;     // 32-bit words per row of the band
    int32_t c0, c1;  // first and last column (inclusive)
};

enum BandMode { BAND_SUM = 0, BAND_LABEL = 1, BAND_BOUNDS = 2, BAND_LABEL_ATOMIC = 3 };
// Options of a band launch: BAND_SUM adds the canvas sums too (kBandSums);
// BAND_BOUNDS leaves the polygons' counts to a label pass that wrote them
// (kBandNoCounts).
constexpr int kBandSums = 1, kBandNoCounts = 2;
// Error-bounds bands of at most this many LDS words run in a launch of their own.
constexpr int kBoundsSmallWords = 4096;
// PT_HYBRID_QUEUE: the hybrid pass that appends its border points to a queue
// for k_exact instead of testing them itself. PT_LABEL_BOUNDS: a count pass
// over the error-bounds label map that also counts the points of its flagged
// (boundary) pixels per pixel into the count canvas.
enum PointMode { PT_INDEX = 0, PT_LABEL = 1, PT_LABEL_HYBRID = 2, PT_SPLAT = 3, PT_HYBRID_QUEUE = 4,
                 PT_LABEL_BOUNDS = 5 };

// Workgroups of a point pass when not set: spatially ordered tables take one
// workgroup per kPassRowsPerWorkgroup rows within these bounds (short row
// ranges per workgroup balance the tail; ordered rows touch few polygons, so
// the per-workgroup flush stays small), other tables the minimum (each
// workgroup flushes about every polygon).
constexpr int kPassWorkgroupsMin = 4096, kPassWorkgroupsMax = 49152, kPassRowsPerWorkgroup = 12288;
constexpr int kSplatRowsPerWorkgroup = 4096;  // canvas splats (PT_SPLAT)

// Arguments of one point pass.
struct PointArgs {
    const float* xy;
    int64_t begin, end;   // rows [begin, end)
    int64_t start;        // begin rounded down to the row group size
    int64_t per_block;    // rows per workgroup (multiple of the rows per iteration)
    DevCols cols;
    DevQuery q;
    DevCanvas cv;
    DevGrid g;
    DevPolys P;
    const void* label;    // label map, 16- or 32-bit ids
    int label16;
    const uint32_t* tiles; // per tile of 2^tshift pixels square: its single label, or kLabelMixed
    int tiles_w;           // tiles per canvas row (LW >> tshift)
    int tshift;
    // hybrid: bitmap of the pixels a snapped edge touches (W bits per row in
    // halo_wpr words); null when the border is dilated instead. A point of an
    // untouched pixel within halo_d pixels of a side it shares with a touched
    // pixel takes the exact test.
    const uint32_t* halo;
    int halo_wpr;
    float halo_d;
    // PT_HYBRID_QUEUE: border points (x, y, aggregate bits) go to queue[0 ..
    // qcap); *qcount counts them (a count above qcap means some were dropped)
    float4* queue;
    uint32_t* qcount;
    uint32_t qcap;
    DevResult res;
    int lds_np;           // LDS sums indexed by polygon when np <= lds_np, else hashed
    uint32_t* cnt;        // count canvas (PT_SPLAT; PT_LABEL_BOUNDS: boundary pixels only)
    unsigned long long* chi;
    unsigned long long* clo;
    int cap;              // workgroups of the pass (0: from wg_rows)
    int wg_rows;          // cap 0: one workgroup per wg_rows rows, kPassWorkgroupsMin..Max
                          // (wg_rows 0: kPassWorkgroupsMin)
    // rows [in_b, in_e) pass every constraint of the query (whole time
    // buckets inside its time window, no other constraint): they run as a
    // query without constraints
    int64_t in_b, in_e;
};

// ---- fused passes over many queries (batch.hip) ----

constexpr int kBatchMax = 64;    // queries of one fused pass (one bit each in a segment mask)
constexpr int kBatchSegs = 2 * kBatchMax + 1;
// Query classes of a fused pass, each run by its own persistent kernel (a
// kernel compiled for one class needs fewer registers than one for all):
// label-map raster queries, hybrid queries with a border queue, per-point
// grid (index) queries, hybrid queries testing border points inline.
enum BatchClass { BC_LABEL = 0, BC_HYBRID = 1, BC_INDEX = 2, BC_HYBRID_INLINE = 3, BC_COUNT = 4 };
// Device counters of a pass, after its block: the chunk counter of class
// BC_LABEL, the exact-tile counter, the border-queue count of every query,
// then the chunk counters of the other classes.
constexpr int kBatchCounters = 2 + kBatchMax + (BC_COUNT - 1);
__host__ __device__ constexpr int batch_chunk_counter(int cls) { return cls == 0 ? 0 : 2 + kBatchMax + cls - 1; }
constexpr int kExactTile = 4096;  // queued border points per exact-pass work item

// Rows [r0, r1) read by the queries of `mask`; split into chunks of
// chunk_rows rows numbered from chunk0.
struct BatchSeg {
    int64_t r0, r1;
    unsigned long long mask;
    int64_t chunk0;
};

// Everything a fused pass reads, copied to the device in one piece: the
// segments and, per query, its point-pass arguments (end = the table's rows)
// and mode (PT_LABEL, PT_LABEL_HYBRID or PT_INDEX).
struct BatchBlock {
    int32_t nseg;
    int32_t nq;
    int32_t lds_cap;      // queries of at most lds_cap polygons sum in LDS
    int32_t pad;
    int64_t nchunks;
    int64_t chunk_rows;
    unsigned long long* res;  // result words of all queries, cleared before the pass
    int64_t res_words;
    BatchSeg seg[kBatchSegs];
    unsigned long long cls[BC_COUNT];  // members of each class (bit j: query j)
    // members that read the same 16-bit label map form a group: a chunk's
    // labels are computed once for the group into LDS (key_off: byte offset
    // of that buffer in the dynamic LDS) and each member reads them there
    int32_t key_off;
    int32_t ngroups;
    int32_t grp[kBatchMax];                // group of each member, -1: none
    unsigned long long gmask[kBatchMax];   // members of each group
    int32_t mode[kBatchMax];
    PointArgs q[kBatchMax];
};
// rows of a chunk whose group keys fit the LDS key buffer
constexpr int64_t kGroupChunkMax = 16384;

// Per-polygon sums of a point pass in LDS: indexed by polygon when np <=
// lds_cap, else kHashSlots hashed slots with a 4-byte tag each.
constexpr int kHashSlots = 512;
inline size_t acc_lds_bytes(int np, int lds_cap, bool avg)
{
    const bool hash = np > lds_cap;
    const size_t slots = hash ? kHashSlots : (size_t)np;
    return (avg ? 3 : 1) * slots * 8 + (hash ? kHashSlots * 4 : 0);
}

// Threads' chunk totals of the error-bounds kernel's row prefix sums.
constexpr int kBandScan = 256;

// LDS bytes a band of `nrows` x `wpr` words needs in each mode.
inline size_t band_lds_bytes(int mode, int nrows, int wpr)
{
    size_t w = (size_t)nrows * wpr;
    if (mode == BAND_BOUNDS)  // toggles, touched rows (+2), area sums (32 per word), scan totals
        return 4 * (w + (size_t)(nrows + 2) * wpr + (size_t)32 * w + kBandScan);
    return 4 * w;
}

// ---- launchers (kernels.hip) ----
void launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v);
// n 64-bit words from src to dst (device addresses; pinned host memory included)
struct CopyWords {
    const unsigned long long* src;
    unsigned long long* dst;
    size_t n;
};
void launch_copy_words(hipStream_t s, CopyWords r0, CopyWords r1 = CopyWords{nullptr, nullptr, 0});
// prev[k]: the vertex before k in its polygon's ring (the last for the first)
void launch_prev(hipStream_t s, const int32_t* off, int np, int64_t nv, int32_t* prev);
void launch_snap(hipStream_t s, const float* xy, int64_t nv, double minx, double miny, double fxs, double fys,
                 int32_t* X, int32_t* Y);
void launch_band(hipStream_t s, int mode, const Band* bands, int nbands, const DevPolys& P,
                 const DevCanvas& c, const uint32_t* cnt, const long long* chi,
                 const unsigned long long* clo, DevResult res, long long* bounds, void* label,
                 int label16, unsigned long long* aux, size_t lds_bytes, int opt);
// Tile map (tiles of 2^tshift pixels square, tshift 3..5); `labeled` (may be
// null) receives the count of labeled pixels.
void launch_tile_labels(hipStream_t s, const void* label, int label16, const DevCanvas& c, int tshift,
                        uint32_t* tiles, unsigned long long* labeled);
void launch_halo_flags(hipStream_t s, const uint32_t* mask, const DevCanvas& c, void* label, int label16);
void launch_touch(hipStream_t s, const int32_t* X, const int32_t* Y, const int32_t* prev,
                  int64_t nverts, const DevCanvas& c, uint32_t* mask_bits);
void launch_mark_rects(hipStream_t s, const int32_t* rects, int nrects, const DevCanvas& c,
                       uint32_t* mask_bits);
void launch_dilate_to_label(hipStream_t s, const uint32_t* mask_bits, const DevCanvas& c, int r,
                            void* label, int label16);
// mode PT_*; a.start and a.per_block are filled in by the launcher.
void launch_points(hipStream_t s, int mode, PointArgs a);
// Exact test of the queued border points of a PT_HYBRID_QUEUE pass (the same
// arguments), summed per workgroup in LDS.
void launch_exact(hipStream_t s, const PointArgs& a);
// Fused pass: clears the results named by the block and the work counter,
// then `nwg` persistent workgroups take chunks from the counter until none is
// left; rows = rows per thread (4 or 8). The block is read-only on the device.
// lds: dynamic LDS bytes of the per-polygon sums. Queries in PT_HYBRID_QUEUE
// mode then get their queued border points tested by a second persistent
// kernel. counters: kBatchCounters words. One persistent kernel per class
// with members (bit c of `classes`); rows[c] = its rows per thread (4 or 8).
void launch_batch(hipStream_t s, const BatchBlock* dev_block, uint32_t* counters, const int* rows, int nwg,
                  size_t lds, unsigned classes);

}  // namespace rj
