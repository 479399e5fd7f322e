// batch.hip -- fused point passes: up to 64 queries share one scan of a
// resident point table.
//
// The host splits the union of the queries' row ranges into segments whose
// set of reading queries is constant (a 64-bit mask each) and the segments
// into chunks of consecutive rows. A fixed number of persistent workgroups
// take chunks from a counter; for a chunk, a workgroup runs every query of the
// segment's mask (of its kernel's class: label-map raster, hybrid or index
// queries run in separate kernels, each compiled for its own modes only) over
// the chunk's rows, one query after the other, with the
// single-query loop (point_rows): its filters, its key (label map, label map +
// exact test on border pixels, or the CPUJoin grid + exact test) and its
// per-polygon sums. A chunk is small (a few thousand rows of a spatially
// ordered table): the first query reads its rows from HBM, the next ones find
// them in cache, and each query's label-map and tile gathers stay within the
// chunk's small area of its maps. So the points are read from HBM about once
// per pass however many queries read them, and the maps of different queries
// do not compete for the cache within a row group.
#include "point_dev.hpp"

namespace rj {
namespace {

// The block through the constant address space: it does not change during a
// pass, so the compiler may keep a query's arguments in scalar registers
// across the loop's global atomics.
typedef const __attribute__((address_space(4))) BatchBlock ConstBlock;

// One query over one chunk, with the single-query loop and its LDS sums
// (acc_init / acc_flush per chunk).
template <int MODE, bool AVG, int R>
__device__ __forceinline__ void query_rows(const PointArgs& a, int64_t r0, int64_t r1, unsigned long long* sacc,
                                           int lds_cap)
{
    Acc A;
    __syncthreads();  // the previous query's sums have been flushed
    acc_init(A, sacc, a.P.np, lds_cap, AVG, a.res);
    Run run{-1, 0, 0, 0};
    // a chunk inside the rows that pass every constraint skips the filters
    point_rows<MODE, AVG, R>(a, r0 - r0 % R, r0, r1, A, run, r0 >= a.in_b && r1 <= a.in_e);
    if (MODE == PT_HYBRID_QUEUE) border_finish(a, run);
    if (run.ck >= 0) run_flush<MODE, AVG>(a, A, run.ck, run.cc, run.ch, run.cl);
    acc_flush(A, AVG);
}

// Group keys of one chunk (rows [r0, r1)): every row's label in the group's
// map as 16-bit keys in LDS: the polygon id, 0xFFFE for a hybrid border
// point, 0xFFFF for none. No filter applies here: the members apply their own.
constexpr uint16_t kKeyBorder = 0xFFFE, kKeyNone = 0xFFFF;

template <bool HYB, int R>
__device__ __forceinline__ void group_keys(const PointArgs& a, int64_t r0, int64_t r1, uint16_t* skey)
{
    const int64_t stride = (int64_t)blockDim.x * R;
    const int64_t start = r0 - r0 % R;
    const bool all_in = rows_inside(a.cols, a.cv);
    float nx[R], ny[R];
    load_xy<R>(a.xy, start + (int64_t)threadIdx.x * R, r1, a.end, nx, ny);
    for (int64_t base = start; base < r1; base += stride) {
        const int64_t i0 = base + (int64_t)threadIdx.x * R;
        float px[R], py[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            px[r] = nx[r];
            py[r] = ny[r];
        }
        if (i0 >= r1) continue;
        int pix[R], ixs[R], iys[R];
        bool ok[R];
#pragma unroll
        for (int r = 0; r < R; r++) ok[r] = i0 + r >= r0 && i0 + r < r1;
        row_pixels<R>(a.cv, all_in, px, py, ok, pix, ixs, iys);
        uint32_t lab[R];
        row_labels<HYB, R>(a, px, py, pix, ixs, iys, lab, [&] { load_xy<R>(a.xy, i0 + stride, r1, a.end, nx, ny); });
#pragma unroll
        for (int r = 0; r < R; r++)
            if (i0 + r >= r0 && i0 + r < r1)
                skey[i0 + r - r0] = lab[r] < kLabelMaxId ? (uint16_t)lab[r] : (lab[r] == kLabelBorder ? kKeyBorder : kKeyNone);
    }
}

// One member of a group over the chunk: its filters and aggregate column,
// the keys from LDS. A hybrid member queues its border rows (their x,y are
// read only for the waves that have one).
template <int MODE, bool AVG, int R>
__device__ __forceinline__ void member_rows(const PointArgs& a, int64_t r0, int64_t r1, const uint16_t* skey,
                                            const Acc& A, Run& run)
{
    const DevQuery& q = a.q;
    const int agg_t = AVG ? agg_kind(a.cols, q) : kAggUnsigned;
    const int64_t stride = (int64_t)blockDim.x * R;
    // a chunk inside the rows that pass every constraint skips the filters
    const bool plain = q.plain || (r0 >= a.in_b && r1 <= a.in_e);
    for (int64_t base = r0 - r0 % R; base < r1; base += stride) {
        const int64_t i0 = base + (int64_t)threadIdx.x * R;
        // interior: the workgroup's whole row group lies in the chunk
        const bool interior = base >= r0 && base + stride <= r1;
        if (!interior && i0 >= r1) continue;
        const bool full = interior || (i0 + R <= a.end);
        bool ok[R];
        if (interior) {
#pragma unroll
            for (int r = 0; r < R; r++) ok[r] = true;
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) ok[r] = (i0 + r >= r0) && (i0 + r < r1);
        }
        uint32_t vraw[R];
#pragma unroll
        for (int r = 0; r < R; r++) vraw[r] = 0;
        if (plain) {
            if (AVG) load_rows<R>(a.cols, q.slot_col[q.agg_slot], i0, full, a.end, vraw);
        } else {
#pragma unroll
            for (int s = 0; s < kMaxSlots; s++) {
                if (s >= q.nslots) break;
                uint32_t v[R];
                const int col = q.slot_col[s];
                load_rows<R>(a.cols, col, i0, full, a.end, v);
                slot_filter<R>(a.cols.type[col], q.slot_lo[s], q.slot_hi[s], v, ok);
                if (AVG && s == q.agg_slot) {
#pragma unroll
                    for (int r = 0; r < R; r++) vraw[r] = v[r];
                }
            }
        }
        int key[R];
        bool need[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t k16 = ok[r] ? skey[i0 + r - r0] : kKeyNone;
            key[r] = k16 < 0xFFF0u ? (int)k16 : -1;
            need[r] = MODE == PT_HYBRID_QUEUE && k16 == kKeyBorder;
        }
        if (MODE == PT_HYBRID_QUEUE) {
            bool any = false;
#pragma unroll
            for (int r = 0; r < R; r++) any = any || need[r];
            if (__ballot(any)) {
                float px[R], py[R];
                load_xy<R>(a.xy, i0, r1, a.end, px, py);
                queue_border<R>(a, px, py, vraw, need, run);
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int k = key[r];
            if (k < 0) continue;
            if (k != run.ck) {
                if (run.ck >= 0) run_flush<MODE, AVG>(a, A, run.ck, run.cc, run.ch, run.cl);
                run.ck = k;
                run.cc = 0;
                run.ch = run.cl = 0;
            }
            run.cc++;
            if (AVG) {
                unsigned long long h, l;
                agg_bits(agg_t, vraw[r], h, l);
                run.ch += h;
                run.cl += l;
            }
        }
    }
}

template <int MODE, bool AVG, int R>
__device__ __forceinline__ void member_query(const PointArgs& a, int64_t r0, int64_t r1, const uint16_t* skey,
                                             unsigned long long* sacc, int lds_cap)
{
    Acc A;
    __syncthreads();  // the previous query's sums have been flushed
    acc_init(A, sacc, a.P.np, lds_cap, AVG, a.res);
    Run run{-1, 0, 0, 0};
    member_rows<MODE, AVG, R>(a, r0, r1, skey, A, run);
    if (MODE == PT_HYBRID_QUEUE) border_finish(a, run);
    if (run.ck >= 0) run_flush<MODE, AVG>(a, A, run.ck, run.cc, run.ch, run.cl);
    acc_flush(A, AVG);
}

// The chunk passes every constraint of the member (its rows run as a query
// without constraints).
__device__ __forceinline__ bool chunk_plain(const PointArgs& a, int64_t r0, int64_t r1)
{
    return a.q.plain || (r0 >= a.in_b && r1 <= a.in_e);
}

// The aggregate column a member reads (-1: count only).
__device__ __forceinline__ int agg_column(const PointArgs& a)
{
    return a.q.agg_type == RJ_AGG_COUNT ? -1 : a.q.slot_col[a.q.agg_slot];
}

// One pass of the chunk for the members in `share` (label-map members of one
// group with per-polygon sums in LDS, for all of which the chunk passes every
// constraint: their sums over it are the same), with the arguments of `lead`;
// the sums go to every member's results, counts only to count members.
template <bool AVG, int R>
__device__ __forceinline__ void shared_query(ConstBlock* CB, int lead, unsigned long long share, int64_t r0,
                                             int64_t r1, const uint16_t* skey, unsigned long long* sacc,
                                             int lds_cap)
{
    const PointArgs& a = *(const PointArgs*)(&CB->q[lead]);
    Acc A;
    __syncthreads();  // the previous query's sums have been flushed
    acc_init(A, sacc, a.P.np, lds_cap, AVG, a.res);
    Run run{-1, 0, 0, 0};
    member_rows<PT_LABEL, AVG, R>(a, r0, r1, skey, A, run);
    if (run.ck >= 0) run_flush<PT_LABEL, AVG>(a, A, run.ck, run.cc, run.ch, run.cl);
    __syncthreads();
    for (int k = threadIdx.x; k < A.lds_np; k += blockDim.x) {
        const unsigned long long c = A.cnt[k];
        if (!c) continue;
        unsigned long long mm = share;
        while (mm) {
            const int j = __ffsll((long long)mm) - 1;
            mm &= mm - 1;
            const PointArgs& aj = *(const PointArgs*)(&CB->q[j]);
            const DevResult& g = aj.res;
            atomicAdd(&g.cnt[k], c);
            if (AVG && aj.q.agg_type != RJ_AGG_COUNT) {
                atomicAdd((unsigned long long*)&g.hi[k], A.hi[k]);
                atomicAdd(&g.lo[k], A.lo[k]);
            }
        }
    }
}

template <int R, int CLS>
__global__ __launch_bounds__(kBlock) void k_batch(const BatchBlock* __restrict__ B, uint32_t* __restrict__ counters)
{
    uint32_t* next = counters + batch_chunk_counter(CLS);
    ConstBlock* CB = (ConstBlock*)B;
    __shared__ int64_t s_chunk0[kBatchSegs];
    __shared__ uint32_t s_take;
    extern __shared__ unsigned long long sacc[];  // per-polygon sums of the running query
    const int lds_cap = B->lds_cap;
    const int nseg = B->nseg;
    const int64_t nchunks = B->nchunks;
    const int64_t chunk_rows = B->chunk_rows;
    for (int k = threadIdx.x; k < nseg; k += blockDim.x) s_chunk0[k] = B->seg[k].chunk0;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_take = atomicAdd(next, 1u);
        __syncthreads();
        const uint32_t c = __builtin_amdgcn_readfirstlane(s_take);
        if ((int64_t)c >= nchunks) break;  // every workgroup ends here once the chunks are taken
        // the segment holding chunk c: the last one with chunk0 <= c
        int lo = 0, hi = nseg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_chunk0[mid] <= (int64_t)c) lo = mid;
            else hi = mid - 1;
        }
        const int sg = __builtin_amdgcn_readfirstlane(lo);
        const BatchSeg& S = B->seg[sg];
        const int64_t r0 = S.r0 + ((int64_t)c - S.chunk0) * chunk_rows;
        const int64_t r1 = min(S.r1, r0 + chunk_rows);
        // uniform across the wave: the query arguments are read with scalar loads
        const unsigned long long sm = S.mask & CB->cls[CLS];
        unsigned long long m = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(sm >> 32)) << 32) |
                               (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)sm);
        while (m) {
            const int q = __ffsll((long long)m) - 1;
            if (CLS == BC_LABEL || CLS == BC_HYBRID) {
                // the members of q's group in this chunk share its labels
                constexpr int GM = CLS == BC_HYBRID ? PT_HYBRID_QUEUE : PT_LABEL;
                const int g = CB->grp[q];
                const unsigned long long gm = g >= 0 ? (m & CB->gmask[g]) : 0ull;
                if (__popcll(gm) >= 2) {
                    m &= ~gm;
                    uint16_t* skey = reinterpret_cast<uint16_t*>(reinterpret_cast<char*>(sacc) + B->key_off);
                    __syncthreads();  // the previous group's members have read their keys
                    group_keys<CLS == BC_HYBRID, R>(*(const PointArgs*)(&CB->q[q]), r0, r1, skey);
                    unsigned long long mm = gm;
                    const PointArgs& aq = *(const PointArgs*)(&CB->q[q]);
                    if (CLS == BC_LABEL && aq.P.np <= lds_cap) {
                        // members the chunk passes every constraint of share one pass: with
                        // the aggregate column of the first such avg member (count members
                        // take its counts), else counts only
                        unsigned long long plain = 0;
                        int col = -1, lead = -1;
                        for (unsigned long long t = gm; t; t &= t - 1) {
                            const int j = __ffsll((long long)t) - 1;
                            const PointArgs& aj = *(const PointArgs*)(&CB->q[j]);
                            if (!chunk_plain(aj, r0, r1)) continue;
                            plain |= 1ull << j;
                            if (col < 0 && agg_column(aj) >= 0) {
                                col = agg_column(aj);
                                lead = j;
                            }
                        }
                        unsigned long long share = 0;
                        for (unsigned long long t = plain; t; t &= t - 1) {
                            const int j = __ffsll((long long)t) - 1;
                            const int cj = agg_column(*(const PointArgs*)(&CB->q[j]));
                            if (cj < 0 || cj == col) share |= 1ull << j;
                        }
                        if (lead < 0 && share) lead = __ffsll((long long)share) - 1;
                        if (__popcll(share) >= 2) {
                            mm &= ~share;
                            if (col >= 0) shared_query<true, R>(CB, lead, share, r0, r1, skey, sacc, lds_cap);
                            else shared_query<false, R>(CB, lead, share, r0, r1, skey, sacc, lds_cap);
                        }
                    }
                    while (mm) {
                        const int j = __ffsll((long long)mm) - 1;
                        mm &= mm - 1;
                        const PointArgs& aj = *(const PointArgs*)(&CB->q[j]);
                        const bool avg = aj.q.agg_type != RJ_AGG_COUNT;
                        if (avg) member_query<GM, true, R>(aj, r0, r1, skey, sacc, lds_cap);
                        else member_query<GM, false, R>(aj, r0, r1, skey, sacc, lds_cap);
                    }
                    continue;
                }
            }
            m &= m - 1;
            const PointArgs& a = *(const PointArgs*)(&CB->q[q]);
            const bool avg = a.q.agg_type != RJ_AGG_COUNT;
            if (CLS == BC_LABEL) {
                if (avg) query_rows<PT_LABEL, true, R>(a, r0, r1, sacc, lds_cap);
                else query_rows<PT_LABEL, false, R>(a, r0, r1, sacc, lds_cap);
            } else if (CLS == BC_INDEX) {
                if (avg) query_rows<PT_INDEX, true, R>(a, r0, r1, sacc, lds_cap);
                else query_rows<PT_INDEX, false, R>(a, r0, r1, sacc, lds_cap);
            } else if (CLS == BC_HYBRID_INLINE) {
                if (avg) query_rows<PT_LABEL_HYBRID, true, R>(a, r0, r1, sacc, lds_cap);
                else query_rows<PT_LABEL_HYBRID, false, R>(a, r0, r1, sacc, lds_cap);
            } else {
                if (avg) query_rows<PT_HYBRID_QUEUE, true, R>(a, r0, r1, sacc, lds_cap);
                else query_rows<PT_HYBRID_QUEUE, false, R>(a, r0, r1, sacc, lds_cap);
            }
        }
    }
}

// The queued border points of one work item (one query's queue, entries
// [i0, i1)): exact test and LDS sums, as k_exact.
template <bool AVG>
__device__ __forceinline__ void exact_tile(const PointArgs& a, uint32_t i0, uint32_t i1, unsigned long long* sacc,
                                           int lds_cap)
{
    Acc A;
    __syncthreads();  // the previous item's sums have been flushed
    acc_init(A, sacc, a.P.np, lds_cap, AVG, a.res);
    const DevQuery& q = a.q;
    const int agg_t = AVG ? agg_kind(a.cols, q) : kAggUnsigned;
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const float4 e = a.queue[i];
        const uint32_t v = __float_as_uint(e.z);
        const int key = exact_first<AVG, true, 4>(a.g, a.P, A, e.x, e.y, agg_t, v);
        if (key >= 0) {
            unsigned long long h = 0, l = 0;
            if (AVG) agg_bits(agg_t, v, h, l);
            acc_add(A, key, 1ull, h, l, AVG);
        }
    }
    acc_flush(A, AVG);
}

// Second kernel of a fused pass: the border queues of its PT_HYBRID_QUEUE
// queries in items of kExactTile points, taken from a counter.
__global__ __launch_bounds__(kBlock) void k_exact_batch(const BatchBlock* __restrict__ B,
                                                       uint32_t* __restrict__ counters)
{
    __shared__ int64_t s_item0[kBatchMax + 1];
    __shared__ uint32_t s_take;
    extern __shared__ unsigned long long sacc[];
    ConstBlock* CB = (ConstBlock*)B;
    const int nq = B->nq;
    const int lds_cap = B->lds_cap;
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int j = 0; j < nq; j++) {
            s_item0[j] = t;
            if (B->mode[j] == PT_HYBRID_QUEUE) {
                const uint32_t n = min(counters[2 + j], B->q[j].qcap);
                t += (n + kExactTile - 1) / kExactTile;
            }
        }
        s_item0[nq] = t;
    }
    __syncthreads();
    const int64_t nitems = s_item0[nq];
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_take = atomicAdd(&counters[1], 1u);
        __syncthreads();
        const uint32_t c = __builtin_amdgcn_readfirstlane(s_take);
        if ((int64_t)c >= nitems) break;  // every workgroup ends here once the items are taken
        int j = 0;
        while (j + 1 < nq && s_item0[j + 1] <= (int64_t)c) j++;
        j = __builtin_amdgcn_readfirstlane(j);
        const PointArgs& a = *(const PointArgs*)(&CB->q[j]);
        const uint32_t n = min(counters[2 + j], a.qcap);
        const uint32_t i0 = (uint32_t)((int64_t)c - s_item0[j]) * kExactTile;
        const uint32_t i1 = min(n, i0 + kExactTile);
        if (a.q.agg_type != RJ_AGG_COUNT) exact_tile<true>(a, i0, i1, sacc, lds_cap);
        else exact_tile<false>(a, i0, i1, sacc, lds_cap);
    }
}

// Zero the result words of the pass and its counters.
__global__ void k_batch_clear(const BatchBlock* __restrict__ B, uint32_t* __restrict__ counters)
{
    unsigned long long* res = B->res;
    const int64_t n = B->res_words;
    if (blockIdx.x == 0 && threadIdx.x < kBatchCounters) counters[threadIdx.x] = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        res[i] = 0;
}

}  // namespace

template <int CLS>
static void launch_class(hipStream_t s, const BatchBlock* dev_block, uint32_t* counters, int rows, int nwg,
                         size_t lds)
{
    if (rows == 4)
        hipLaunchKernelGGL((k_batch<4, CLS>), dim3(nwg), dim3(kBlock), lds, s, dev_block, counters);
    else
        hipLaunchKernelGGL((k_batch<8, CLS>), dim3(nwg), dim3(kBlock), lds, s, dev_block, counters);
}

void launch_batch(hipStream_t s, const BatchBlock* dev_block, uint32_t* counters, const int* rows, int nwg,
                  size_t lds, unsigned classes)
{
    hipLaunchKernelGGL(k_batch_clear, dim3(256), dim3(kBlock), 0, s, dev_block, counters);
    if (classes & (1u << BC_LABEL)) launch_class<BC_LABEL>(s, dev_block, counters, rows[BC_LABEL], nwg, lds);
    if (classes & (1u << BC_HYBRID)) launch_class<BC_HYBRID>(s, dev_block, counters, rows[BC_HYBRID], nwg, lds);
    if (classes & (1u << BC_INDEX)) launch_class<BC_INDEX>(s, dev_block, counters, rows[BC_INDEX], nwg, lds);
    if (classes & (1u << BC_HYBRID_INLINE))
        launch_class<BC_HYBRID_INLINE>(s, dev_block, counters, rows[BC_HYBRID_INLINE], nwg, lds);
    if (classes & (1u << BC_HYBRID))
        hipLaunchKernelGGL(k_exact_batch, dim3(nwg), dim3(kBlock), lds, s, dev_block, counters);
}

}  // namespace rj
