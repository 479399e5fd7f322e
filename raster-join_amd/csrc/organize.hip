// organize.hip -- spatial organization of a resident point table: rows are
// reordered by the Morton key of their location (16 bits per axis over the
// table's bbox), so the rows one workgroup reads are close on the canvas and
// the splat / label gather stay in cache. With a time column the key is
// (time bucket, Morton key): each bucket is a contiguous row range, so a time
// window reads only its buckets (the time filter itself stays exact). Join
// results do not depend on the row order.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "rj_dev.hpp"

namespace rj {
namespace {

__device__ __forceinline__ uint32_t spread16(uint32_t v)
{
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__global__ void k_morton(const float2* __restrict__ xy, int64_t n, float minx, float miny, float sx,
                         float sy, uint32_t* keys, uint32_t* idx)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 p = xy[i];
    float fx = (p.x - minx) * sx, fy = (p.y - miny) * sy;
    fx = fminf(fmaxf(fx, 0.0f), 65535.0f);
    fy = fminf(fmaxf(fy, 0.0f), 65535.0f);
    if (!(fx == fx)) fx = 0.0f;
    if (!(fy == fy)) fy = 0.0f;
    keys[i] = spread16((uint32_t)fx) | (spread16((uint32_t)fy) << 1);
    idx[i] = (uint32_t)i;
}

template <class T>
__global__ void k_gather(const T* __restrict__ src, const uint32_t* __restrict__ perm, int64_t n,
                         T* __restrict__ dst)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

__global__ void k_key64(const float2* __restrict__ xy, const uint32_t* __restrict__ time, int64_t n, float minx,
                        float miny, float sx, float sy, uint32_t t0, uint32_t span, uint32_t nb,
                        unsigned long long* keys, uint32_t* idx)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 p = xy[i];
    float fx = (p.x - minx) * sx, fy = (p.y - miny) * sy;
    fx = fminf(fmaxf(fx, 0.0f), 65535.0f);
    fy = fminf(fmaxf(fy, 0.0f), 65535.0f);
    if (!(fx == fx)) fx = 0.0f;
    if (!(fy == fy)) fy = 0.0f;
    const uint32_t t = time[i];
    uint32_t b = t < t0 ? 0u : (t - t0) / span;
    if (b >= nb) b = nb - 1;
    keys[i] = ((unsigned long long)b << 32) | (spread16((uint32_t)fx) | (spread16((uint32_t)fy) << 1));
    idx[i] = (uint32_t)i;
}

// off[b] = first row of bucket b (b = 0..nb), from the sorted keys.
__global__ void k_bucket_off(const unsigned long long* __restrict__ keys, int64_t n, uint32_t nb, int64_t* off)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t b = (int64_t)(keys[i] >> 32);
    const int64_t bp = i ? (int64_t)(keys[i - 1] >> 32) : -1;
    for (int64_t k = bp + 1; k <= b; k++) off[k] = i;
    if (i == n - 1)
        for (int64_t k = b + 1; k <= (int64_t)nb; k++) off[k] = n;
}

int bits_for(uint32_t v)
{
    int b = 0;
    while (b < 32 && (v >> b) != 0) b++;
    return b;
}

// flag = 1 when some value of the float column is not finite or lies outside
// [-2^31, 2^31).
__global__ void k_float_range(const float* __restrict__ v, int64_t n, uint32_t* flag)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    bool bad = false;
    const int64_t n4 = n / 4;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        const float4 f = reinterpret_cast<const float4*>(v)[i];
        bad |= !(f.x >= -2147483648.0f && f.x < 2147483648.0f) | !(f.y >= -2147483648.0f && f.y < 2147483648.0f) |
               !(f.z >= -2147483648.0f && f.z < 2147483648.0f) | !(f.w >= -2147483648.0f && f.w < 2147483648.0f);
    }
    for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        bad |= !(v[i] >= -2147483648.0f && v[i] < 2147483648.0f);
    if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

}  // namespace

// Bit c of *mask: float column c is finite and in [-2^31, 2^31) on every row
// (its aggregate decode then needs no range tests). flags: ncols words of
// device memory. Returns a hipError_t.
int float_columns_in_range(hipStream_t s, int64_t n, void* const* cols, const int32_t* types, int ncols,
                           uint32_t* flags, uint32_t* mask)
{
    *mask = 0;
    hipError_t e = hipMemsetAsync(flags, 0, (size_t)ncols * 4, s);
    for (int c = 0; c < ncols && e == hipSuccess; c++) {
        if (types[c] != RJ_ATTR_FLOAT || n == 0) continue;
        const int64_t want = (n / 4 + 255) / 256;
        const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, 4096));
        k_float_range<<<blocks, 256, 0, s>>>(static_cast<const float*>(cols[c]), n, flags + c);
        e = hipGetLastError();
    }
    uint32_t h[RJ_MAX_COLUMNS] = {0};
    if (e == hipSuccess && ncols > 0) e = hipMemcpyAsync(h, flags, (size_t)ncols * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return (int)e;
    for (int c = 0; c < ncols; c++)
        if (types[c] == RJ_ATTR_FLOAT && (n == 0 || !h[c])) *mask |= 1u << c;
    return 0;
}

// Sorts the rows of xy (+ columns) by Morton key in place, or, with a time
// column, by (time bucket, Morton key) and writes the nb+1 bucket offsets to
// off_dev. scratch must hold organize_scratch_bytes(n) bytes. Returns a
// hipError_t.
size_t organize_scratch_bytes(int64_t n)
{
    size_t temp32 = 0, temp64 = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, temp32, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 32);
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, temp64, (unsigned long long*)nullptr,
                                             (unsigned long long*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (int)n, 0, 64);
    size_t kv = (size_t)n * (8 * 2 + 4 * 2);  // keys in/out + values in/out
    size_t col = (size_t)n * 8;               // the widest column (xy)
    return kv + col + std::max(temp32, temp64) + 1024;
}

int organize_rows(hipStream_t s, int64_t n, float* xy, void** cols, const int32_t* types, int ncols,
                  const float bbox[4], const uint32_t* time, uint32_t t0, uint32_t span, uint32_t nb,
                  int64_t* off_dev, void* scratch)
{
    if (n < 1) return 0;
    char* base = (char*)scratch;
    unsigned long long* k_in = (unsigned long long*)base;   // 32-bit keys use the front half
    unsigned long long* k_out = k_in + n;
    uint32_t* v_in = (uint32_t*)(k_out + n);
    uint32_t* v_out = v_in + n;
    char* colbuf = (char*)(v_out + n);
    char* temp = colbuf + (size_t)n * 8;
    size_t temp_bytes = 0;
    const float sx = 65535.0f / fmaxf(bbox[2] - bbox[0], 1e-30f);
    const float sy = 65535.0f / fmaxf(bbox[3] - bbox[1], 1e-30f);
    const int blocks = (int)((n + 255) / 256);
    hipError_t e;
    if (time) {
        const int end_bit = 32 + bits_for(nb - 1);
        hipLaunchKernelGGL(k_key64, dim3(blocks), dim3(256), 0, s, (const float2*)xy, time, n, bbox[0], bbox[1], sx,
                           sy, t0, span, nb, k_in, v_in);
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, k_in, k_out, v_in, v_out, (int)n, 0, end_bit, s);
        e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, k_in, k_out, v_in, v_out, (int)n, 0, end_bit, s);
        if (e != hipSuccess) return (int)e;
        hipLaunchKernelGGL(k_bucket_off, dim3(blocks), dim3(256), 0, s, k_out, n, nb, off_dev);
    } else {
        uint32_t* k32_in = (uint32_t*)k_in;
        uint32_t* k32_out = (uint32_t*)k_out;
        hipLaunchKernelGGL(k_morton, dim3(blocks), dim3(256), 0, s, (const float2*)xy, n, bbox[0], bbox[1], sx,
                           sy, k32_in, v_in);
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, k32_in, k32_out, v_in, v_out, (int)n, 0, 32, s);
        e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, k32_in, k32_out, v_in, v_out, (int)n, 0, 32, s);
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(k_gather<float2>, dim3(blocks), dim3(256), 0, s, (const float2*)xy, v_out, n,
                       (float2*)colbuf);
    e = hipMemcpyAsync(xy, colbuf, (size_t)n * 8, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return (int)e;
    for (int c = 0; c < ncols; c++) {
        if (types[c] == RJ_ATTR_UINT8) {
            hipLaunchKernelGGL(k_gather<uint8_t>, dim3(blocks), dim3(256), 0, s, (const uint8_t*)cols[c],
                               v_out, n, (uint8_t*)colbuf);
            e = hipMemcpyAsync(cols[c], colbuf, (size_t)n, hipMemcpyDeviceToDevice, s);
        } else {
            hipLaunchKernelGGL(k_gather<uint32_t>, dim3(blocks), dim3(256), 0, s, (const uint32_t*)cols[c],
                               v_out, n, (uint32_t*)colbuf);
            e = hipMemcpyAsync(cols[c], colbuf, (size_t)n * 4, hipMemcpyDeviceToDevice, s);
        }
        if (e != hipSuccess) return (int)e;
    }
    return (int)hipGetLastError();
}

}  // namespace rj
