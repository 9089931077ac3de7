// gamma.hip -- colour-space steps of the input pipeline: FloatImage::toLinear
// / toGamma on the R, G, B planes of one or more planar FP32 RGBA images
// (FloatImage.cpp:258-297, Gamma.cpp:311-441), and the normal-map
// renormalisation of generated mip levels.
//
// One launch covers a list of levels: each level is 4 planes of n floats, the
// first 3n (R, G, B) are transformed and the last n (A) copied.  A lane owns
// four consecutive floats of a level (one float when a pointer is not 16-byte
// aligned), so the whole launch streams HBM at 16 bytes in and out per float4.
#include "gamma.hpp"
#include "ntx_common.h"
#include "ntx_context.h"

namespace {

using u64 = unsigned long long;

// mode 0: powf(x, 11/5) table form; 1: powf(x, 5/11) table form;
// 2: powf(max(0, x), power), evaluated as exp2(power * log2(x)) in double and
// rounded once: the correctly rounded powf but for inputs within about 2^-28
// of a rounding boundary.  The host C library's powf, which the reference
// calls, is not correctly rounded for about 5 in 10^4 inputs; there the two
// differ by one unit in the last place.
__device__ __forceinline__ float gamma_apply(float x, int mode, float power, const float *tab)
{
    if (mode == 0) return ntx_gamma::pow_11_5(x, tab);
    if (mode == 1) return ntx_gamma::pow_5_11(x, tab);
    return float(exp2(double(power) * log2(double(ntx_gamma::max0(x)))));
}

template <bool VEC>
__global__ __launch_bounds__(256) void gamma_kernel(ntx::GammaJob job, int mode, float power,
                                                    const ntx::Tables *__restrict__ tab)
{
    __shared__ float s_tab[512];
    if (mode != 2) {
        const float *t = mode == 0 ? tab->pow_11_5 : tab->pow_5_11;
        for (int i = threadIdx.x; i < 512; i += blockDim.x) s_tab[i] = t[i];
    }
    __syncthreads();
    const u64 total = job.start[job.levels];
    const u64 stride = u64(gridDim.x) * blockDim.x;
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
        int l = 0;
        while (i >= job.start[l + 1]) l++;
        const u64 j = i - job.start[l];
        const u64 rgb = 3ull * job.texels[l];  // floats of R, G, B
        if (VEC) {
            float4 v = reinterpret_cast<const float4 *>(job.src[l])[j];
            const u64 f = 4 * j;
            if (f + 0 < rgb) v.x = gamma_apply(v.x, mode, power, s_tab);
            if (f + 1 < rgb) v.y = gamma_apply(v.y, mode, power, s_tab);
            if (f + 2 < rgb) v.z = gamma_apply(v.z, mode, power, s_tab);
            if (f + 3 < rgb) v.w = gamma_apply(v.w, mode, power, s_tab);
            reinterpret_cast<float4 *>(job.dst[l])[j] = v;
        } else {
            float v = job.src[l][j];
            if (j < rgb) v = gamma_apply(v, mode, power, s_tab);
            job.dst[l][j] = v;
        }
    }
}

// The normal-map step Compressor::process applies to every generated level
// of a normal map (Context.cpp:329-336): Surface::expandNormals() with its
// default scale 2 and bias -2 * 127 / 255 (Surface.cpp:2959-2963,
// FloatImage.cpp:231-242), normalizeNormalMap() = FloatImage::normalize(0)
// with normalizeSafe(v, 0, 0) (FloatImage.cpp:201-219, Vector.inl:549-595),
// then packNormals() = scaleBias(0.5, 0.5) (Surface.cpp:2952-2956).
// expand_pack = 0: normalizeNormalMap alone (the KTX branch of
// Compressor::process, Context.cpp:446-449), on the values as they are.
__global__ __launch_bounds__(256) void normal_kernel(float *__restrict__ d, u64 n, int expand_pack)
{
    const float bias = -2.0f * 127.0f / 255.0f;
    const u64 stride = u64(gridDim.x) * blockDim.x;
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        float x = d[i], y = d[i + n], z = d[i + 2 * n];
        if (expand_pack) {
            x = 2.0f * x + bias;
            y = 2.0f * y + bias;
            z = 2.0f * z + bias;
        }
        const float l = sqrtf(x * x + y * y + z * z);
        if (fabsf(l) <= 0.0f) {
            x = y = z = 0.0f;
        } else {
            const float s = 1.0f / l;
            x = x * s; y = y * s; z = z * s;
        }
        if (expand_pack) {
            x = 0.5f * x + 0.5f;
            y = 0.5f * y + 0.5f;
            z = 0.5f * z + 0.5f;
        }
        d[i] = x;
        d[i + n] = y;
        d[i + 2 * n] = z;
    }
}

}  // namespace

namespace ntx {

hipError_t launch_normals(float *d_rgba, size_t n, bool expand_pack, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    size_t groups = (n + 255) / 256;
    if (groups > 16384) groups = 16384;
    hipLaunchKernelGGL(normal_kernel, dim3(unsigned(groups)), dim3(256), 0, stream, d_rgba, u64(n), expand_pack ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_gamma(const GammaJob &job_in, int mode, float power, const Tables *d_tab, hipStream_t stream)
{
    GammaJob job = job_in;
    bool vec = true;
    for (int l = 0; l < job.levels; l++)
        vec = vec && (reinterpret_cast<uintptr_t>(job.src[l]) & 15) == 0 && (reinterpret_cast<uintptr_t>(job.dst[l]) & 15) == 0;
    // Lanes per level: one per float4 (n per level) or per float (4n).
    job.start[0] = 0;
    for (int l = 0; l < job.levels; l++) job.start[l + 1] = job.start[l] + (vec ? job.texels[l] : 4 * job.texels[l]);
    const unsigned long long total = job.start[job.levels];
    if (total == 0) return hipSuccess;
    unsigned long long groups = (total + 255) / 256;
    if (groups > 16384) groups = 16384;
    if (vec)
        hipLaunchKernelGGL(gamma_kernel<true>, dim3(unsigned(groups)), dim3(256), 0, stream, job, mode, power, d_tab);
    else
        hipLaunchKernelGGL(gamma_kernel<false>, dim3(unsigned(groups)), dim3(256), 0, stream, job, mode, power, d_tab);
    return hipGetLastError();
}

}  // namespace ntx
