// gamma.hpp -- the reference's gamma transforms, shared by the device kernels
// and the host table builders so both evaluate the same float operations.
//
// FloatImage::toLinear / toGamma (FloatImage.cpp:258-284) use, for gamma
// 2.2f, fast forms of powf(x, 11/5) and powf(x, 5/11) (Gamma.cpp:311-357):
// the float's sign-and-exponent field indexes a 512-entry table of
// 2^(e * p), and a quartic in the mantissa m in [1, 2) approximates m^p; the
// result is their product.  Any other gamma goes through
// powf(max(0, x), gamma) or powf(max(0, x), 1 / gamma) (FloatImage.cpp:286-297).
// Every product and sum is rounded on its own (no fused multiply-add), as in
// the reference's scalar and SSE2 forms alike.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace ntx_gamma {

// Table entry k (sign-and-exponent field k) of 2^((k - 127) * num / den):
// 0 for k == 0 (zero and subnormal inputs) and for negative inputs (k >= 256),
// +inf for k == 255 (inf and NaN), else the double power rounded to float.
inline float table_entry(int k, double num, double den)
{
    if (k == 0 || k >= 256) return 0.0f;
    if (k == 255) return INFINITY;
    return float(std::pow(2.0, double(k - 127) * num / den));
}

// Mantissa of x with a zero exponent: a float in [1, 2).
__host__ __device__ inline float unit_mantissa(uint32_t u)
{
    union {
        uint32_t u;
        float f;
    } m;
    m.u = (u & ((1u << 23) - 1u)) | (127u << 23);
    return m.f;
}

__host__ __device__ inline uint32_t float_bits(float x)
{
    union {
        float f;
        uint32_t u;
    } m;
    m.f = x;
    return m.u;
}

// powf(x, 5/11): linear -> gamma 2.2.  table = the 5/11 exponent table.
__host__ __device__ inline float pow_5_11(float x, const float *table)
{
    const uint32_t u = float_bits(x);
    const float m = unit_mantissa(u);
    const float pe = table[u >> 23];
    const float pm = (((-0.0110083047f * m + 0.0905038750f) * m - 0.324697506f) * m + 0.876040946f) * m + 0.369160989f;
    return pe * pm;
}

// powf(x, 11/5): gamma 2.2 -> linear.  table = the 11/5 exponent table.
__host__ __device__ inline float pow_11_5(float x, const float *table)
{
    const uint32_t u = float_bits(x);
    const float m = unit_mantissa(u);
    const float pe = table[u >> 23];
    const float pm = (((-0.00916587552f * m + 0.119315466f) * m + 1.01847068f) * m - 0.158338739f) * m + 0.0297184721f;
    return pe * pm;
}

// nv::max(0.0f, x) = (x < 0) ? 0 : x (nvcore/Utils.h:161-187; NaN passes).
__host__ __device__ inline float max0(float x) { return (x < 0.0f) ? 0.0f : x; }

// nv::equal(gamma, 1.0f), relative NV_EPSILON (nvmath/nvmath.h:52, 139-143):
// Surface::toLinear and Surface::toGamma do nothing for such a gamma
// (Surface.cpp:1470-1487).
inline bool is_unit_gamma(float gamma)
{
    const float ag = std::fabs(gamma);
    return std::fabs(gamma - 1.0f) <= 0.0001f * (ag > 1.0f ? ag : 1.0f);
}

}  // namespace ntx_gamma
