// calib.hip - measurement calibrators used by bench.py: the sustained
// v_sad_u8 rate of the device (the integer peak the full search is priced
// against) and a plain copy kernel whose bytes are known exactly (the scale
// factor for the FETCH_SIZE / WRITE_SIZE counters).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bc250_calib.h"

// 8 independent v_sad_u8 chains per thread; inline asm keeps the compiler
// from folding the loop-invariant operands.
__global__ __launch_bounds__(256) void sad_peak_kernel(uint32_t *out, int iters, uint32_t seed)
{
    uint32_t a[8], b[8], acc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        a[k] = seed * (threadIdx.x + 1u) + 0x01234567u * (k + 1);
        b[k] = seed ^ (0x89abcdefu * (k + 3));
        acc[k] = 0;
    }
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int r = 0; r < 16; r++)
#pragma unroll
            for (int k = 0; k < 8; k++) asm volatile("v_sad_u8 %0, %1, %2, %0" : "+v"(acc[k]) : "v"(a[k]), "v"(b[k]));
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) s += acc[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

extern "C" int bc250_calib_sad_peak(uint32_t *out, int blocks, int iters, void *stream)
{
    if (blocks <= 0 || iters <= 0) return -1;
    sad_peak_kernel<<<blocks, 256, 0, (hipStream_t)stream>>>(out, iters, 0x9e3779b9u);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void copy_u32_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int bc250_calib_copy(const void *src, void *dst, size_t bytes, void *stream)
{
    if (bytes % 16) return -1;
    copy_u32_kernel<<<2048, 256, 0, (hipStream_t)stream>>>((const uint4 *)src, (uint4 *)dst, bytes / 16);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
