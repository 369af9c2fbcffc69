// gcoo_gen.hip -- counter-based synthetic input generators (device side).
//
// The benchmark inputs are generated in HBM instead of on the host: the
// reference fills host arrays with glibc rand() (src/utils.cpp:5-19) and
// copies them over PCIe (src/main.cu:37-38).  The definition of every value
// is in include/gcoo_spdm.h (gcoo_fill_*): a SplitMix64 finaliser over a
// per-element counter, so any element can be recomputed on the host.
#include "gcoo_internal.h"

namespace gcoo {
namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void fill_kernel(
    float *__restrict__ dst, long long n, uint64_t key, float scale,
    uint64_t keep_threshold, int sparse)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        const uint64_t h = mix64(key + (uint64_t)(i + 1) * kGamma);
        const int num = 2 * (int)(h >> 40) - (1 << 24) + 1;  // odd, |num| < 2^24
        float v = (float)num * 0x1p-24f * scale;
        if (sparse && (h & ((1ull << 40) - 1)) < keep_threshold) v = 0.0f;
        dst[i] = v;
    }
}

// IEEE round-to-nearest-even FP32 -> BF16; NaN stays a quiet NaN.
__device__ __forceinline__ unsigned short bf16_rne(unsigned u)
{
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))
        return (unsigned short)((u >> 16) | 0x0040u);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__global__ __launch_bounds__(256) void f32_to_bf16_kernel(const float *__restrict__ x,
                                                          unsigned short *__restrict__ y,
                                                          long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        y[i] = bf16_rne(__float_as_uint(x[i]));
}

__global__ __launch_bounds__(256) void seg_to_bf16_kernel(const int2 *__restrict__ seg,
                                                          int2 *__restrict__ seg16, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int2 e = seg[i];
        seg16[i] = make_int2(e.x, (int)bf16_rne((unsigned)e.y));
    }
}

unsigned blocks_for(long long n)
{
    long long b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (unsigned)b;
}

}  // namespace

int launch_f32_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s)
{
    if (n <= 0) return GCOO_OK;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, s, x,
                       reinterpret_cast<unsigned short *>(y), (long long)n);
    GCOO_CHECK_LAUNCH("f32_to_bf16_kernel");
    return GCOO_OK;
}

int launch_seg_to_bf16(const int2 *seg, int2 *seg16, int64_t n, hipStream_t s)
{
    if (n <= 0) return GCOO_OK;
    hipLaunchKernelGGL(seg_to_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, s, seg, seg16,
                       (long long)n);
    GCOO_CHECK_LAUNCH("seg_to_bf16_kernel");
    return GCOO_OK;
}

int launch_fill(float *dst, int64_t n, uint64_t seed, uint64_t stream_id,
                float scale, uint64_t keep_threshold, bool sparse, hipStream_t s)
{
    if (n <= 0) return GCOO_OK;
    const uint64_t key = [&] {
        uint64_t z = seed * kGamma + stream_id;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }();
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, dst,
                       (long long)n, key, scale, keep_threshold, sparse ? 1 : 0);
    GCOO_CHECK_LAUNCH("fill_kernel");
    return GCOO_OK;
}

}  // namespace gcoo
