// metrics.hip -- decoding and error metrics on the device.  The error of a
// BC1 level against its FP32 source:
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

// ------------------------------------------------------------------ decode --
// Surface::setImage2D (Surface.cpp:908-1118) for BC1/BC3/BC4/BC5: each block
// decoded to 8-bit RGBA as nv::BlockDXT1 / BlockDXT5 / BlockATI1 / BlockATI2
// do (BlockDXT.cpp:43-98, 200-217, 336-432, 475-566), then float(c) / 255 per
// channel into the planes.  d3d9 only changes the BC4/BC5 interpolation
// (the reference passes false to the BC1/BC3 colour decoder for both D3D
// decoders).  One thread per texel.

// BlockDXT1::evaluatePalette with d3d9 = false, entry idx, as r, g, b, a.
__device__ __forceinline__ uint32_t bc1_texel(uint32_t colours, uint32_t bits, int t)
{
    const uint32_t c0 = colours & 0xffffu, c1 = colours >> 16;
    const uint32_t idx = (bits >> (2 * t)) & 3u;
    int e0[3], e1[3];
    expand(c0, e0[0], e0[1], e0[2]);
    expand(c1, e1[0], e1[1], e1[2]);
    int v[3];
    uint32_t a = 255u;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (idx == 0) v[k] = e0[k];
        else if (idx == 1) v[k] = e1[k];
        else if (c0 > c1) v[k] = idx == 2 ? (2 * e0[k] + e1[k]) / 3 : (2 * e1[k] + e0[k]) / 3;
        else v[k] = idx == 2 ? (e0[k] + e1[k]) / 2 : 0;
    }
    if (!(c0 > c1) && idx == 3) a = 0u;
    return uint32_t(v[0]) | uint32_t(v[1]) << 8 | uint32_t(v[2]) << 16 | a << 24;
}

// AlphaBlockDXT5::evaluatePalette + index (BlockDXT.cpp:336-416), texel t.
__device__ __forceinline__ uint32_t alpha_texel(uint2 blk, int t, int d3d9)
{
    const int a0 = int(blk.x & 0xffu), a1 = int((blk.x >> 8) & 0xffu);
    const uint64_t u = uint64_t(blk.x) | uint64_t(blk.y) << 32;
    const int i = int((u >> (16 + 3 * t)) & 7u);
    if (i == 0) return uint32_t(a0);
    if (i == 1) return uint32_t(a1);
    if (a0 > a1) return uint32_t(((8 - i) * a0 + (i - 1) * a1 + (d3d9 ? 3 : 0)) / 7);
    if (i == 6) return 0u;
    if (i == 7) return 255u;
    return uint32_t(((6 - i) * a0 + (i - 1) * a1 + (d3d9 ? 2 : 0)) / 5);
}

__global__ __launch_bounds__(256) void decode_kernel(int format, int d3d9, const uint8_t *__restrict__ blocks, int w, int h,
                                                     int bw, float *__restrict__ dst, const float *__restrict__ div255)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const size_t b = size_t(y >> 2) * size_t(bw) + size_t(x >> 2);
    const int t = (y & 3) * 4 + (x & 3);
    uint32_t rgba;
    if (format == NTX_FORMAT_BC1) {
        const uint2 blk = reinterpret_cast<const uint2 *>(blocks)[b];
        rgba = bc1_texel(blk.x, blk.y, t);
    } else if (format == NTX_FORMAT_BC3) {
        const uint4 blk = reinterpret_cast<const uint4 *>(blocks)[b];
        rgba = (bc1_texel(blk.z, blk.w, t) & 0x00ffffffu) | alpha_texel(make_uint2(blk.x, blk.y), t, 0) << 24;
    } else if (format == NTX_FORMAT_BC4) {
        const uint32_t v = alpha_texel(reinterpret_cast<const uint2 *>(blocks)[b], t, d3d9);
        rgba = v | v << 8 | v << 16 | 0xff000000u;
    } else {  // BC5: x -> R, y -> G, B = 0
        const uint4 blk = reinterpret_cast<const uint4 *>(blocks)[b];
        rgba = alpha_texel(make_uint2(blk.x, blk.y), t, d3d9) | alpha_texel(make_uint2(blk.z, blk.w), t, d3d9) << 8 |
               0xff000000u;
    }
    const size_t plane = size_t(w) * size_t(h), o = size_t(y) * size_t(w) + size_t(x);
    dst[o] = div255[rgba & 0xffu];
    dst[o + plane] = div255[(rgba >> 8) & 0xffu];
    dst[o + 2 * plane] = div255[(rgba >> 16) & 0xffu];
    dst[o + 3 * plane] = div255[rgba >> 24];
}

// ------------------------------------------------------------------ errors --
// nv::rmsColorError / rmsAlphaError (ErrorMetric.cpp:13-69): per texel the
// float terms of the reference, summed in double; the per-thread and
// per-workgroup sums are combined in a fixed order (the reference adds them
// one texel after another, so the double total may differ in its last bits).
__global__ __launch_bounds__(256) void rms_kernel(const float *__restrict__ ref, const float *__restrict__ img, size_t n,
                                                  int mode, double *__restrict__ partial)
{
    __shared__ double s_sum[256];
    double acc = 0.0;
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
        if (mode == 2) {
            const float a = img[i + 3 * n] - ref[i + 3 * n];
            acc += double(a * a);
        } else {
            const float r = ref[i] - img[i], g = ref[i + n] - img[i + n], b = ref[i + 2 * n] - img[i + 2 * n];
            float a = 1.0f;
            if (mode == 1) {
                const float a0 = ref[i + 3 * n];
                a = a0 * a0;
            }
            acc += double((r * r) * a);
            acc += double((g * g) * a);
            acc += double((b * b) * a);
        }
    }
    s_sum[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s_sum[0];
}

}  // namespace

namespace ntx {

hipError_t launch_decode(int format, int d3d9, const void *d_blocks, int w, int h, float *d_dst, const float *d_div255,
                         hipStream_t stream)
{
    if (w <= 0 || h <= 0) return hipSuccess;
    const dim3 grid(unsigned((w + 63) / 64), unsigned((h + 3) / 4));
    hipLaunchKernelGGL(decode_kernel, grid, dim3(256), 0, stream, format, d3d9, static_cast<const uint8_t *>(d_blocks), w,
                       h, (w + 3) / 4, d_dst, d_div255);
    return hipGetLastError();
}

int rms_partials() { return 1024; }

// mode 0: rmsColorError, 1: the same weighted by the reference's alpha^2,
// 2: rmsAlphaError.  partial: rms_partials() doubles; out: the sum.
hipError_t launch_rms(const float *d_ref, const float *d_img, size_t n, int mode, double *d_partial, double *d_out,
                      hipStream_t stream)
{
    hipLaunchKernelGGL(rms_kernel, dim3(unsigned(rms_partials())), dim3(256), 0, stream, d_ref, d_img, n, mode, d_partial);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, stream, d_partial, rms_partials(), d_out);
    return hipGetLastError();
}

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
