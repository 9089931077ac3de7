// metrics.hip -- error of a BC1 level against its FP32 source on the device:
// D3D10 decode of each block (BlockDXT.cpp:43-98, the palette icbc evaluates
// with, icbc.h:2557-2585) and the sum over texels of the squared R, G, B
// differences in 8-bit units, as bc1enc.cpp:73-81 measures it.  Per-block
// sums are reduced in a fixed order, so the result is deterministic.
#include "ntx_common.h"

namespace {

__device__ __forceinline__ void expand(uint32_t c, int &r, int &g, int &b)
{
    const int r5 = (c >> 11) & 31, g6 = (c >> 5) & 63, b5 = c & 31;
    r = (r5 << 3) | (r5 >> 2);
    g = (g6 << 2) | (g6 >> 4);
    b = (b5 << 3) | (b5 >> 2);
}

__global__ __launch_bounds__(256) void bc1_sse_kernel(const float *__restrict__ src, int w, int h, int bw, int n,
                                                      const uint2 *__restrict__ blocks, double *__restrict__ partial)
{
    __shared__ double s_sum[256];
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    double sse = 0.0;
    if (b < n) {
        const uint2 blk = blocks[b];
        const uint32_t c0 = blk.x & 0xffffu, c1 = blk.x >> 16;
        int p[4][3];
        expand(c0, p[0][0], p[0][1], p[0][2]);
        expand(c1, p[1][0], p[1][1], p[1][2]);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (c0 > c1) {
                p[2][k] = (2 * p[0][k] + p[1][k]) / 3;
                p[3][k] = (2 * p[1][k] + p[0][k]) / 3;
            } else {
                p[2][k] = (p[0][k] + p[1][k]) / 2;
                p[3][k] = 0;
            }
        }
        const int bx = b % bw, by = b / bw;
        const size_t plane = size_t(w) * size_t(h);
        for (int t = 0; t < 16; t++) {
            const int x = 4 * bx + (t & 3), y = 4 * by + (t >> 2);
            if (x >= w || y >= h) continue;
            const size_t o = size_t(y) * size_t(w) + size_t(x);
            const int idx = (blk.y >> (2 * t)) & 3;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double d = double(p[idx][k]) - double(src[o + k * plane]) * 255.0;
                sse += d * d;
            }
        }
    }
    s_sum[threadIdx.x] = sse;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s_sum[0];
}

__global__ __launch_bounds__(256) void sum_kernel(const double *__restrict__ partial, int n, double *__restrict__ out)
{
    __shared__ double s_sum[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += partial[i];
    s_sum[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s_sum[0];
}

}  // namespace

namespace ntx {

// partial: (nblocks + 255) / 256 doubles; out: one double.
hipError_t launch_bc1_sse(const float *d_src, int w, int h, const void *d_blocks, double *d_partial, double *d_out,
                          hipStream_t stream)
{
    const int bw = (w + 3) / 4, n = bw * ((h + 3) / 4);
    const int groups = (n + 255) / 256;
    hipLaunchKernelGGL(bc1_sse_kernel, dim3(groups), dim3(256), 0, stream, d_src, w, h, bw, n,
                       static_cast<const uint2 *>(d_blocks), d_partial);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, stream, d_partial, groups, d_out);
    return hipGetLastError();
}

int bc1_sse_partials(int w, int h) { return (((w + 3) / 4) * ((h + 3) / 4) + 255) / 256; }

}  // namespace ntx
