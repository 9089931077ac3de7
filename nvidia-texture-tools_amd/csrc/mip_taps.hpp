// mip_taps.hpp -- per-texel mipmap filter arithmetic shared by the mip
// kernels (mip.hip) and the batch kernel (batch.hip).  Same tap order as
// NVTT's single-threaded FloatImage filters and no fused multiply-add, so
// every output is bit-identical to the reference's.
#pragma once

#include "ntx_common.h"

namespace ntx_mip {

// FloatImage::fastDownSample (FloatImage.cpp:559-737) for one output texel
// of one plane s (W x H source).  For a 1D source (W == 1 or H == 1) x is
// the output index along the longer axis and y is ignored.
__device__ __forceinline__ float box_texel(const float *s, int W, int H, int x, int y)
{
    if (W == 1 || H == 1) {
        const int n = (W * H) / 2;  // outputs along the longer axis
        const int i = x;
        const float *p = s + 2 * size_t(i);
        if ((W * H) & 1) {
            const float scale = 1.0f / float(2 * n + 1);
            const float w0 = float(n - i), w1 = float(n), w2 = float(1 + i);
            return scale * ((w0 * p[0] + w1 * p[1]) + w2 * p[2]);
        }
        return 0.5f * (p[0] + p[1]);
    }
    const int w = W / 2, h = H / 2;
    const float *p = s + size_t(2 * y) * size_t(W) + 2 * size_t(x);
    if ((W & 1) == 0 && (H & 1) == 0) return 0.25f * (((p[0] + p[1]) + p[W]) + p[W + 1]);
    if ((W & 1) && (H & 1)) {
        const float scale = 1.0f / float(W * H);
        const float v0 = float(h - y), v1 = float(h), v2 = float(1 + y);
        const float w0 = float(w - x), w1 = float(w), w2 = float(1 + x);
        float f = 0.0f;
        f += v0 * ((w0 * p[0] + w1 * p[1]) + w2 * p[2]);
        f += v1 * ((w0 * p[W] + w1 * p[W + 1]) + w2 * p[W + 2]);
        f += v2 * ((w0 * p[2 * W] + w1 * p[2 * W + 1]) + w2 * p[2 * W + 2]);
        return f * scale;
    }
    if (W & 1) {
        const float scale = 1.0f / float(2 * W);
        const float w0 = float(w - x), w1 = float(w), w2 = float(1 + x);
        float f = 0.0f;
        f += w0 * (p[0] + p[W]);
        f += w1 * (p[1] + p[W + 1]);
        f += w2 * (p[2] + p[W + 2]);
        return f * scale;
    }
    const float scale = 1.0f / float(2 * H);
    const float v0 = float(h - y), v1 = float(h), v2 = float(1 + y);
    float f = 0.0f;
    f += v0 * (p[0] + p[1]);
    f += v1 * (p[W] + p[W + 1]);
    f += v2 * (p[2 * W] + p[2 * W + 1]);
    return f * scale;
}

// FloatImage.h:298-330 wrap helpers.
__device__ __forceinline__ int wrap_index(int x, int n, int wm)
{
    if (x >= 0 && x < n) return x;
    if (wm == NTX_WRAP_CLAMP) return x < 0 ? 0 : n - 1;
    if (wm == NTX_WRAP_REPEAT) return x >= 0 ? x % n : (x + 1) % n + n - 1;
    if (n == 1) return 0;
    x = abs(x);
    while (x >= n) x = abs(n + n - x - 2);
    return x;
}

// FloatImage::resize (FloatImage.cpp:761-808) for one output texel (x, y) of
// plane s: for each of the wy source rows the Y taps read, the applyKernelX
// value of that row at output column x (FloatImage.cpp:1115-1133), summed in
// tap order; then the applyKernelY sum over those values
// (FloatImage.cpp:1136-1176).  Each X value is computed exactly as the
// reference's temporary image holds it.
__device__ __forceinline__ float poly_texel(const float *s, int W, int H, int x, int y, const float *kx, const int *lx,
                                            int wx, const float *ky, const int *ly, int wy, int wrap)
{
    const float *kxr = kx + size_t(x) * wx;
    const float *kyr = ky + size_t(y) * wy;
    const int l = lx[x], top = ly[y];
    const bool xin = l >= 0 && l + wx <= W;
    float sum = 0.0f;
    for (int j = 0; j < wy; ++j) {
        const float *row = s + size_t(wrap_index(top + j, H, wrap)) * size_t(W);
        float t = 0.0f;
        if (xin) {
            for (int i = 0; i < wx; ++i) t += kxr[i] * row[l + i];
        } else {
            for (int i = 0; i < wx; ++i) t += kxr[i] * row[wrap_index(l + i, W, wrap)];
        }
        sum += kyr[j] * t;
    }
    return sum;
}

}  // namespace ntx_mip
