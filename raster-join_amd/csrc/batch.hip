// batch.hip -- fused point passes: up to 64 queries share one scan of a
// resident point table.
//
// The host splits the union of the queries' row ranges into segments whose
// set of reading queries is constant (a 64-bit mask each) and the segments
// into chunks. A fixed number of persistent workgroups take chunks from a
// counter. Per row group a thread loads x,y and every column the batch uses
// once, keeps them in registers and runs each query of the segment's mask on
// them: its filters, its key (label map, label map + exact test on border
// pixels, or the CPUJoin grid + exact test) and its per-polygon sums. So the
// points are read from HBM once per pass however many queries read them.
//
// Per query and row group a lane folds its rows into a run (key, count, sum)
// flushed on key changes; the last run is combined across the wave when all
// lanes agree on the key (the common case on a spatially ordered table) and
// added with one atomic.
#include "point_dev.hpp"

namespace rj {
namespace {

// All lanes hold run (ck, cc, ch, cl), ck < 0 for none: one atomic per wave
// when every lane with a run has the same key, else one per lane.
__device__ __forceinline__ void flush_wave(const Acc& A, int ck, unsigned long long cc, unsigned long long ch,
                                           unsigned long long cl, bool avg)
{
    const unsigned long long act = __ballot(ck >= 0);
    if (!act) return;
    const int lead = __ffsll((long long)act) - 1;
    const int k0 = __shfl(ck, lead, kWave);
    if (__all(ck < 0 || ck == k0)) {
        cc = wave_sum(cc);
        if (avg) {
            ch = wave_sum(ch);
            cl = wave_sum(cl);
        }
        if ((int)__lane_id() == lead) acc_add(A, k0, cc, ch, cl, avg);
    } else if (ck >= 0) {
        acc_add(A, ck, cc, ch, cl, avg);
    }
}

// One query on one row group held in registers.
template <int R>
__device__ __forceinline__ void run_query(const BatchQuery& Q, const float (&px)[R], const float (&py)[R],
                                          const uint32_t (&v)[kBatchSlots][R], const bool (&in)[R])
{
    bool ok[R];
#pragma unroll
    for (int r = 0; r < R; r++) ok[r] = in[r];
    for (int k = 0; k < Q.ncons; k++) {
        const int slot = Q.con_slot[k], t = Q.con_type[k], op = Q.con_op[k];
        const uint32_t cv = Q.con_val[k];
#pragma unroll
        for (int s = 0; s < kBatchSlots; s++) {
            if (s != slot) continue;
#pragma unroll
            for (int r = 0; r < R; r++) ok[r] = ok[r] && cmp_bits(t, op, v[s][r], cv);
        }
    }
    const bool avg = Q.avg != 0;
    uint32_t vraw[R];
#pragma unroll
    for (int r = 0; r < R; r++) vraw[r] = 0;
    if (avg) {
#pragma unroll
        for (int s = 0; s < kBatchSlots; s++) {
            if (s != Q.agg_slot) continue;
#pragma unroll
            for (int r = 0; r < R; r++) vraw[r] = v[s][r];
        }
    }
    Acc A;
    A.cnt = A.hi = A.lo = nullptr;
    A.lds_np = 0;
    A.g = Q.res;
    const int agg_t = Q.agg_type;
    int key[R];
    if (Q.mode == PT_INDEX) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            key[r] = -1;
            if (ok[r])
                key[r] = avg ? exact_first<true>(Q.g, Q.P, A, px[r], py[r], agg_t, vraw[r])
                             : exact_first<false>(Q.g, Q.P, A, px[r], py[r], agg_t, vraw[r]);
        }
    } else {
        int ixs[R], iys[R];
        uint32_t lab[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int pix = pixel_xy(Q.cv, px[r], py[r], ixs[r], iys[r]);
            lab[r] = (ok[r] && pix >= 0) ? Q.tiles[(iys[r] >> Q.tshift) * Q.tiles_w + (ixs[r] >> Q.tshift)]
                                         : kLabelNone;
        }
#pragma unroll
        for (int r = 0; r < R; r++)
            if (lab[r] == kLabelMixed) lab[r] = label_at(Q.label, Q.label16, (int64_t)iys[r] * Q.cv.LW + ixs[r]);
#pragma unroll
        for (int r = 0; r < R; r++) {
            key[r] = lab[r] < kLabelMaxId ? (int)lab[r] : -1;
            if (Q.mode == PT_LABEL_HYBRID && lab[r] == kLabelBorder)
                key[r] = avg ? exact_first<true>(Q.g, Q.P, A, px[r], py[r], agg_t, vraw[r])
                             : exact_first<false>(Q.g, Q.P, A, px[r], py[r], agg_t, vraw[r]);
        }
    }
    int ck = -1;
    unsigned long long cc = 0, ch = 0, cl = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int k = key[r];
        if (k < 0) continue;
        if (k != ck) {
            if (ck >= 0) acc_add(A, ck, cc, ch, cl, avg);
            ck = k;
            cc = ch = cl = 0;
        }
        cc++;
        if (avg) {
            unsigned long long h, l;
            agg_bits(agg_t, vraw[r], h, l);
            ch += h;
            cl += l;
        }
    }
    flush_wave(A, ck, cc, ch, cl, avg);
}

template <int R>
__global__ __launch_bounds__(kBlock) void k_batch(const BatchBlock* __restrict__ B, uint32_t* __restrict__ next,
                                                  const float* __restrict__ xy, int64_t nrows)
{
    __shared__ int64_t s_chunk0[kBatchSegs];
    __shared__ uint32_t s_take;
    const int nseg = B->nseg;
    const int nslots = B->nslots;
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
        // uniform across the wave: the query descriptors are then read with
        // scalar loads
        const unsigned long long mask =
            ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(S.mask >> 32)) << 32) |
            (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)S.mask);
        for (int64_t base = r0 - r0 % R; base < r1; base += (int64_t)kBlock * R) {
            const int64_t i0 = base + (int64_t)threadIdx.x * R;
            const bool any = i0 < r1;
            const bool full = i0 + R <= nrows;
            float px[R], py[R];
            uint32_t v[kBatchSlots][R];
            bool in[R];
#pragma unroll
            for (int r = 0; r < R; r++) {
                px[r] = py[r] = 0.0f;
                in[r] = any && (i0 + r >= r0) && (i0 + r < r1);
            }
#pragma unroll
            for (int s = 0; s < kBatchSlots; s++)
#pragma unroll
                for (int r = 0; r < R; r++) v[s][r] = 0;
            if (any) {
                load_xy<R>(xy, i0, r1, nrows, px, py);
#pragma unroll
                for (int s = 0; s < kBatchSlots; s++)
                    if (s < nslots) load_col<R>(B->col[s], B->type[s], i0, full, nrows, v[s]);
            }
            unsigned long long m = mask;
            while (m) {
                const int q = __ffsll((long long)m) - 1;
                m &= m - 1;
                run_query<R>(B->q[q], px, py, v, in);
            }
        }
    }
}

// Zero the result words of the pass and its work counter.
__global__ void k_batch_clear(const BatchBlock* __restrict__ B, uint32_t* __restrict__ next)
{
    unsigned long long* res = B->res;
    const int64_t n = B->res_words;
    if (blockIdx.x == 0 && threadIdx.x == 0) *next = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        res[i] = 0;
}

}  // namespace

void launch_batch(hipStream_t s, const BatchBlock* dev_block, uint32_t* counter, const float* xy, int64_t nrows,
                  int rows, int nwg)
{
    hipLaunchKernelGGL(k_batch_clear, dim3(256), dim3(kBlock), 0, s, dev_block, counter);
    if (rows == 4)
        hipLaunchKernelGGL(k_batch<4>, dim3(nwg), dim3(kBlock), 0, s, dev_block, counter, xy, nrows);
    else
        hipLaunchKernelGGL(k_batch<8>, dim3(nwg), dim3(kBlock), 0, s, dev_block, counter, xy, nrows);
}

}  // namespace rj
