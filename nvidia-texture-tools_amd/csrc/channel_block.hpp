// channel_block.hpp -- device code shared by the BC3/BC4/BC5 kernels
// (bc345.hip) and the batch kernel (batch.hip): the 8-bit block gather of
// ColorBlock::init (ColorBlock.cpp:80-110: truncating float->u8, edge blocks
// filled by repeating pixels) and QuickCompress::compressDXT5A on one channel
// (QuickCompressDXT.cpp:772-821).
#pragma once

#include "ntx_common.h"

namespace ntx_bc4 {


// nv::max / nv::min (nvcore/Utils.h:160-188).
__device__ __forceinline__ float nvmaxf(float a, float b) { return (b < a) ? a : b; }
__device__ __forceinline__ float nvminf(float a, float b) { return (a < b) ? a : b; }

__device__ __forceinline__ uint32_t to_u8(float v)
{
    return uint32_t(uint8_t(255 * nvminf(nvmaxf(v, 0.0f), 1.0f)));
}

// Texel (e, i) of block (bx, by) with pixel repetition at the edges.
__device__ __forceinline__ size_t block_texel(int w, int h, int bx, int by, int t)
{
    const int x = 4 * bx, y = 4 * by;
    const int bw = min(w - x, 4), bh = min(h - y, 4);
    const int e = t & 3, i = t >> 2;
    return size_t(y + i % bh) * size_t(w) + size_t(x + e % bw);
}

// ------------------------------------------------------------- BC4 channel --

// BlockDXT.cpp:336-378 palettes (d3d9 = false).
__device__ __forceinline__ void alpha_palette(uint32_t a0, uint32_t a1, uint32_t pal[8])
{
    pal[0] = a0;
    pal[1] = a1;
    if (a0 > a1) {
#pragma unroll
        for (int k = 1; k <= 6; k++) pal[1 + k] = ((7 - k) * a0 + k * a1) / 7;
    } else {
#pragma unroll
        for (int k = 1; k <= 4; k++) pal[1 + k] = ((5 - k) * a0 + k * a1) / 5;
        pal[6] = 0;
        pal[7] = 255;
    }
}

// The 16 channel values of a block are held packed, texel 4q + b in byte b
// of word q.  The texel loops below run four texels per trip with the trip
// itself not unrolled, which keeps register use low; every sum still runs
// in texel order 0..15.

// computeAlphaIndices (QuickCompressDXT.cpp:505-536): nearest palette entry,
// first on ties; rewrites the 48 index bits and returns the squared error.
__device__ __forceinline__ uint32_t alpha_indices(const uint32_t vq[4], uint64_t &blk)
{
    uint32_t pal[8];
    alpha_palette(uint32_t(blk & 0xff), uint32_t((blk >> 8) & 0xff), pal);
    uint32_t total = 0;
    uint64_t bits = 0;
    uint32_t w0 = vq[0], w1 = vq[1], w2 = vq[2], w3 = vq[3];
#pragma unroll 1
    for (int q = 0; q < 4; q++) {
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int v = int((w0 >> (8 * b)) & 0xffu);
            uint32_t be = 256 * 256, bp = 8;
#pragma unroll
            for (uint32_t p = 0; p < 8; p++) {
                const int d = int(pal[p]) - v;
                const uint32_t e = uint32_t(d * d);
                if (e < be) { be = e; bp = p; }
            }
            total += be;
            bits |= uint64_t(bp) << (16 + 3 * (4 * q + b));
        }
        w0 = w1; w1 = w2; w2 = w3;
    }
    blk = (blk & 0xffffull) | bits;
    return total;
}

// optimizeAlpha8 (QuickCompressDXT.cpp:538-592): least-squares endpoints for
// the current indices; truncating conversion, NaN -> 0 through nv::max.
__device__ __forceinline__ void optimize_alpha8(const uint32_t vq[4], uint64_t &blk)
{
    float a2 = 0, b2 = 0, ab = 0, ax = 0, bx = 0;
    uint32_t w0 = vq[0], w1 = vq[1], w2 = vq[2], w3 = vq[3];
#pragma unroll 1
    for (int q = 0; q < 4; q++) {
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const float v = float((w0 >> (8 * b)) & 0xffu);
            const uint32_t idx = uint32_t(blk >> (16 + 3 * (4 * q + b))) & 7u;
            const float alpha = (idx < 2) ? 1.0f - float(idx) : (8.0f - float(idx)) / 7.0f;
            const float beta = 1 - alpha;
            a2 += alpha * alpha;
            b2 += beta * beta;
            ab += alpha * beta;
            ax += alpha * v;
            bx += beta * v;
        }
        w0 = w1; w1 = w2; w2 = w3;
    }
    const float factor = 1.0f / (a2 * b2 - ab * ab);
    const float a = (ax * b2 - bx * ab) * factor;
    const float b = (bx * a2 - ax * ab) * factor;
    uint32_t a0 = uint32_t(nvminf(nvmaxf(a, 0.0f), 255.0f));
    uint32_t a1 = uint32_t(nvminf(nvmaxf(b, 0.0f), 255.0f));
    uint64_t bits = blk & ~0xffffull;
    if (a0 < a1) {
        const uint32_t t = a0; a0 = a1; a1 = t;
        uint64_t nb = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t idx = uint32_t(bits >> (16 + 3 * i)) & 7u;
            nb |= uint64_t(idx < 2 ? 1 - idx : 9 - idx) << (16 + 3 * i);
        }
        bits = nb;
    } else if (a0 == a1) {
        bits = 0;
    }
    blk = bits | uint64_t(a0) | uint64_t(a1) << 8;
}

// QuickCompress::compressDXT5A (QuickCompressDXT.cpp:772-821) on one channel.
__device__ __forceinline__ uint64_t compress_channel(const uint32_t vq[4])
{
    uint32_t mx = 0, mn = 255;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t v = (vq[i >> 2] >> (8 * (i & 3))) & 0xffu;
        mx = v > mx ? v : mx;
        mn = v < mn ? v : mn;
    }
    // One value: both endpoints are it, every texel takes palette entry 0
    // (first on ties) with error 0, and the refinement below stops at its
    // first step (no error is below 0): the same block, without the work.
    if (mx == mn) return uint64_t(mx) | uint64_t(mx) << 8;
    const uint32_t d = (mx - mn) / 34;
    uint64_t blk = uint64_t((mx - d) & 0xff) | uint64_t((mn + d) & 0xff) << 8;
    uint32_t besterr = alpha_indices(vq, blk);
    uint64_t best = blk;
    if (besterr == 0) return best;  // likewise: the first refinement step stops
#pragma unroll 1
    for (int it = 0; it < 8; it++) {
        optimize_alpha8(vq, blk);
        const uint32_t err = alpha_indices(vq, blk);
        if (err >= besterr) break;
        // sameIndices (QuickCompressDXT.cpp:643-647) ORs ~0xFFFF into both
        // words, so only the endpoint bytes are compared.
        if ((blk & 0xffffull) == (best & 0xffffull)) { best = blk; break; }
        besterr = err;
        best = blk;
    }
    return best;
}

// The block's 16 values of one plane, packed as compress_channel takes them
// (ColorBlock::init gather: truncating conversion, edge pixels repeated).
__device__ __forceinline__ void gather_channel(const float *plane, int w, int h, int bx, int by, uint32_t vq[4])
{
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t word = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) word |= to_u8(plane[block_texel(w, h, bx, by, 4 * q + b)]) << (8 * b);
        vq[q] = word;
    }
}

}  // namespace ntx_bc4
