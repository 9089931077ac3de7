// bc345.hip -- BC3 / BC4 / BC5 encoders for gfx950, bit-exact with NVTT's
// CPU path at Quality_Normal:
//   alpha / BC4 channel: QuickCompress::compressDXT5A (QuickCompressDXT.cpp:772-821),
//   BC3 colour: nvsquish WeightedClusterFit scalar build (weightedclusterfit.cpp:39-106, 476-589)
//     or OptimalCompress::compressDXT1 for single-colour blocks (CompressorDX9.cpp:158-176).
// Input is gathered as ColorBlock::init does (ColorBlock.cpp:80-110):
// truncating float->u8 conversion, edge blocks filled by repeating pixels.
#include "bucket.hpp"
#include "channel_block.hpp"

#include <cfloat>

namespace {

using namespace ntx_bc4;

// One thread per (block, channel job).  Channel plane: 0 = R, 1 = G, 3 = A.
__global__ __launch_bounds__(256) void channel_kernel(const float *__restrict__ src, int w, int h, int bw, int nblocks,
                                                     int njobs, int ch0, int ch1, int stride, int off1,
                                                     uint8_t *__restrict__ out, int urgent)
{
    if (urgent) __builtin_amdgcn_s_setprio(3);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblocks * njobs) return;
    const int blk = t / njobs, job = t - blk * njobs;
    const int ch = job == 0 ? ch0 : ch1;
    const int bx = blk % bw, by = blk / bw;
    const float *plane = src + size_t(ch) * size_t(w) * size_t(h);
    uint32_t vq[4];
    gather_channel(plane, w, h, bx, by, vq);
    const uint64_t r = compress_channel(vq);
    uint2 *dst = reinterpret_cast<uint2 *>(out + size_t(blk) * stride + (job == 0 ? 0 : off1));
    *dst = make_uint2(uint32_t(r), uint32_t(r >> 32));
}

// ---------------------------------------------------------- BC3 colour ----

// std::min / std::max as used by squish's Vec3 Min/Max (maths.h).
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

struct SqCand {
    float ax, ay, az, bx, by, bz;
};

// colourblock.cpp:30-54 FloatToInt / FloatTo565.
__device__ __forceinline__ int float_to_int(float a, int limit)
{
    int i = int(a + 0.5f);
    return i < 0 ? 0 : (i > limit ? limit : i);
}

// ------------------------------------------ BC3 colour, blocks bucketed by colour count --
//
// Pass 1 (bc3_classify_kernel, one thread per block) finishes single-colour
// blocks and appends every other block to the list of its distinct-colour
// count k (2..16).  Pass 2 (bc3_search_kernel, one block per lane) walks the
// lists from k = 16 down: the lanes of a wave share k, so they run the
// reference's nested (c0, c1, c2) loops with identical trip counts and the
// reference's own running sums, with no cross-lane work.

using namespace ntx_bucket;

__device__ __forceinline__ void bc3_single_colour(const ntx::Tables *__restrict__ tab, uint32_t r8, uint32_t g8, uint32_t b8,
                                                  uint2 *dst)
{
    // OptimalCompress::compressDXT1 (OptimalCompressDXT.cpp:254-269).
    uint32_t c0 = uint32_t(tab->omatch5[r8] & 0xff) << 11 | uint32_t(tab->omatch6[g8] & 0xff) << 5 | uint32_t(tab->omatch5[b8] & 0xff);
    uint32_t c1 = uint32_t(tab->omatch5[r8] >> 8) << 11 | uint32_t(tab->omatch6[g8] >> 8) << 5 | uint32_t(tab->omatch5[b8] >> 8);
    uint32_t idx = 0xaaaaaaaau;
    if (c0 < c1) { const uint32_t t = c0; c0 = c1; c1 = t; idx ^= 0x55555555u; }
    *dst = make_uint2(c0 | c1 << 16, idx);
}

__global__ __launch_bounds__(64) void zero_counts_kernel(int *counts)
{
    for (int i = threadIdx.x; i < kListsOffset; i += 64) counts[i] = 0;
}

// BC3 pass 1, one thread per block (CompressorDX9.cpp:145-176): the alpha
// half by QuickCompress::compressDXT5A (as channel_kernel does for BC4), then
// the colour: single-colour blocks are finished with OptimalCompress, every
// other block is appended to the list of its distinct-colour count k and its
// 8-bit texels are stored for the search as one 64-byte line per block,
// R | G << 8 | B << 16 | A << 24 (A = 255 unless alpha weights the fit).  A
// finished block is written as one 16-byte store when the output is 16-byte
// aligned; inner blocks read their rows as 16-byte loads when the rows are
// aligned (rows16).
// kN: Format_BC3n (CompressorDXT5n::compressBlock, CompressorDX9.cpp:179-222
// at Normal/Production): the alpha block encodes the red channel
// (swizzle(4, 1, 5, 0), then QuickCompress::compressDXT5A), the colour block
// the texels (255, G, 0, A) (swizzle(4, 1, 5, 3)) with metric (0, 1, 0),
// given by the host.  A single-green block takes compressDXT1G(g)
// (OptimalCompressDXT.cpp:294-311), which is compressDXT1 of (255, g, 0):
// the OMatch5 entries of 255 and 0 are (31, 31) and (0, 0).
template <bool kN>
__global__ __launch_bounds__(256) void bc3_front_kernel(const float *__restrict__ src, int w, int h, int bw, int nblocks,
                                                        const ntx::Tables *__restrict__ tab, uint8_t *__restrict__ out,
                                                        int wide, int *__restrict__ counts, int *__restrict__ lists, int cap,
                                                        int weight_by_alpha, uint4 *__restrict__ packed, int urgent,
                                                        int rows16)
{
    if (urgent) __builtin_amdgcn_s_setprio(3);
    __shared__ int s_cnt[kCountSlots], s_base[kCountSlots];
    const int t = threadIdx.x;
    if (t < kCountSlots) s_cnt[t] = 0;
    __syncthreads();
    const int blk = blockIdx.x * blockDim.x + t;
    int k = 0, local = -1;
    if (blk < nblocks) {
        const int bx = blk % bw, by = blk / bw;
        const size_t plane = size_t(w) * size_t(h);
        uint32_t av[4];
        uint32_t rgb[16];
        if (rows16 && 4 * bx + 4 <= w && 4 * by + 4 <= h) {
            // Inner block, 16-byte aligned rows: each row of each plane is one
            // 16-byte load (the same texels ColorBlock::init gathers).
            const float *p0 = src + size_t(4 * by) * size_t(w) + size_t(4 * bx);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float *p = p0 + size_t(r) * size_t(w);
                const float4 vr = *reinterpret_cast<const float4 *>(p);
                const float4 vg = *reinterpret_cast<const float4 *>(p + plane);
                const float4 vb = *reinterpret_cast<const float4 *>(p + 2 * plane);
                const float4 va = *reinterpret_cast<const float4 *>(p + 3 * plane);
                rgb[4 * r + 0] = to_u8(vr.x) | to_u8(vg.x) << 8 | to_u8(vb.x) << 16;
                rgb[4 * r + 1] = to_u8(vr.y) | to_u8(vg.y) << 8 | to_u8(vb.y) << 16;
                rgb[4 * r + 2] = to_u8(vr.z) | to_u8(vg.z) << 8 | to_u8(vb.z) << 16;
                rgb[4 * r + 3] = to_u8(vr.w) | to_u8(vg.w) << 8 | to_u8(vb.w) << 16;
                av[r] = to_u8(va.x) | to_u8(va.y) << 8 | to_u8(va.z) << 16 | to_u8(va.w) << 24;
            }
        } else {
            gather_channel(src + 3 * plane, w, h, bx, by, av);
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const size_t o = block_texel(w, h, bx, by, i);
                rgb[i] = to_u8(src[o]) | to_u8(src[o + plane]) << 8 | to_u8(src[o + 2 * plane]) << 16;
            }
        }
        uint64_t alpha;
        if constexpr (kN) {
            uint32_t xr[4];  // byte b of xr[q]: red of texel 4q + b
#pragma unroll
            for (int q = 0; q < 4; q++)
                xr[q] = (rgb[4 * q] & 0xffu) | (rgb[4 * q + 1] & 0xffu) << 8 | (rgb[4 * q + 2] & 0xffu) << 16 |
                        (rgb[4 * q + 3] & 0xffu) << 24;
#pragma unroll
            for (int i = 0; i < 16; i++) rgb[i] = 0xffu | (rgb[i] & 0xff00u);
            alpha = compress_channel(xr);
        } else {
            alpha = compress_channel(av);
        }
        uint8_t *dst = out + size_t(blk) * 16;
        bool single = true;
#pragma unroll
        for (int i = 1; i < 16; i++) single = single && rgb[i] == rgb[0];
        if (single) {
            uint2 colour;
            bc3_single_colour(tab, rgb[0] & 0xff, (rgb[0] >> 8) & 0xff, rgb[0] >> 16, &colour);
            if (wide) {
                *reinterpret_cast<uint4 *>(dst) = make_uint4(uint32_t(alpha), uint32_t(alpha >> 32), colour.x, colour.y);
            } else {
                reinterpret_cast<uint2 *>(dst)[0] = make_uint2(uint32_t(alpha), uint32_t(alpha >> 32));
                reinterpret_cast<uint2 *>(dst)[1] = colour;
            }
        } else {
            *reinterpret_cast<uint2 *>(dst) = make_uint2(uint32_t(alpha), uint32_t(alpha >> 32));
#pragma unroll
            for (int i = 0; i < 16; i++) {
                bool d = true;
#pragma unroll
                for (int j = 0; j < i; j++) d = d && rgb[j] != rgb[i];
                k += d ? 1 : 0;
            }
            local = atomicAdd(&s_cnt[k], 1);
            uint4 *line = packed + size_t(blk) * 4;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t a = weight_by_alpha ? av[q] : 0xffffffffu;  // byte b of av[q] = alpha of texel 4q + b
                line[q] = make_uint4(rgb[4 * q] | (a & 0xffu) << 24, rgb[4 * q + 1] | ((a >> 8) & 0xffu) << 24,
                                     rgb[4 * q + 2] | ((a >> 16) & 0xffu) << 24, rgb[4 * q + 3] | (a >> 24) << 24);
            }
        }
    }
    append_to_lists(s_cnt, s_base, counts, lists, cap, k, local, blk);
}

// BC1a pass 1, one thread per block (CompressorDXT1a::compressBlock,
// CompressorDX9.cpp:83-111): a block whose 16 texels share one RGB takes
// OptimalCompress::compressDXT1a (OptimalCompressDXT.cpp:271-292: without
// transparent texels compressDXT1's OMatch endpoints, otherwise the
// OMatchAlpha endpoints with col0 <= col1 and indices 0xaaaaaaaa | the
// alpha-0 mask); any other block is appended to the list of its number k of
// distinct RGB values among the texels of non-zero alpha (the squish kDxt1
// colour set) with its texels as R | G << 8 | B << 16 | A << 24.  A block
// whose texels are all transparent but not of one RGB has an empty colour
// set; the fit's single Compress3 candidate then has zero weights, its
// endpoints clamp to 0 with error 0, and WriteColourBlock3 gives endpoints
// (0, 0) with every index 3, which is written here directly.
__global__ __launch_bounds__(256) void bc1a_front_kernel(const float *__restrict__ src, int w, int h, int bw,
                                                         int nblocks, const ntx::Tables *__restrict__ tab,
                                                         uint8_t *__restrict__ out, int *__restrict__ counts,
                                                         int *__restrict__ lists, int cap, uint4 *__restrict__ packed)
{
    __shared__ int s_cnt[kCountSlots], s_base[kCountSlots];
    const int t = threadIdx.x;
    if (t < kCountSlots) s_cnt[t] = 0;
    __syncthreads();
    const int blk = blockIdx.x * blockDim.x + t;
    int k = 0, local = -1;
    if (blk < nblocks) {
        const int bx = blk % bw, by = blk / bw;
        const size_t plane = size_t(w) * size_t(h);
        uint32_t px[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const size_t o = block_texel(w, h, bx, by, i);
            px[i] = to_u8(src[o]) | to_u8(src[o + plane]) << 8 | to_u8(src[o + 2 * plane]) << 16 |
                    to_u8(src[o + 3 * plane]) << 24;
        }
        uint32_t mask = 0;
        bool single = true;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            if ((px[i] >> 24) == 0) mask |= 3u << (2 * i);
            single = single && ((px[i] ^ px[0]) & 0xffffffu) == 0;
        }
        uint2 *dst = reinterpret_cast<uint2 *>(out + size_t(blk) * 8);
        if (single) {
            const uint32_t r8 = px[0] & 0xff, g8 = (px[0] >> 8) & 0xff, b8 = (px[0] >> 16) & 0xff;
            if (mask == 0) {
                bc3_single_colour(tab, r8, g8, b8, dst);
            } else {
                uint32_t c0 = uint32_t(tab->omatcha5[r8] & 0xff) << 11 | uint32_t(tab->omatcha6[g8] & 0xff) << 5 |
                              uint32_t(tab->omatcha5[b8] & 0xff);
                uint32_t c1 = uint32_t(tab->omatcha5[r8] >> 8) << 11 | uint32_t(tab->omatcha6[g8] >> 8) << 5 |
                              uint32_t(tab->omatcha5[b8] >> 8);
                if (c0 > c1) { const uint32_t tt = c0; c0 = c1; c1 = tt; }
                *dst = make_uint2(c0 | c1 << 16, 0xaaaaaaaau | mask);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                bool d = (px[i] >> 24) != 0;
#pragma unroll
                for (int j = 0; j < i; j++) d = d && !((px[j] >> 24) != 0 && ((px[j] ^ px[i]) & 0xffffffu) == 0);
                k += d ? 1 : 0;
            }
            if (k == 0) {
                *dst = make_uint2(0u, 0xffffffffu);
            } else {
                local = atomicAdd(&s_cnt[k], 1);
                uint4 *line = packed + size_t(blk) * 4;
#pragma unroll
                for (int q = 0; q < 4; q++) line[q] = make_uint4(px[4 * q], px[4 * q + 1], px[4 * q + 2], px[4 * q + 3]);
            }
        }
    }
    append_to_lists(s_cnt, s_base, counts, lists, cap, k, local, blk);
}

// The Compress4 loop body (weightedclusterfit.cpp:507-541) on running sums,
// in two parts: the cluster weights' terms (alpha2, beta2, alphabeta and
// the inverse determinant), then the endpoints and the error.
__device__ __forceinline__ float4 sq_weights(float x0w, float x1w, float x2w, float sumw)
{
    const float w3 = ((sumw - x0w) - x1w) - x2w;
    const float alpha2 = (x0w + x1w * (4.0f / 9.0f)) + x2w * (1.0f / 9.0f);
    const float beta2 = (w3 + x2w * (4.0f / 9.0f)) + x1w * (1.0f / 9.0f);
    const float ab = (x1w + x2w) * (2.0f / 9.0f);
    const float factor = 1.0f / (alpha2 * beta2 - ab * ab);
    return make_float4(alpha2, beta2, ab, factor);
}

__device__ __forceinline__ float sq_eval_t(float4 x0, float4 x1, float4 x2, float4 xsum, float4 t, float mx, float my,
                                           float mz, SqCand *out)
{
    const float alpha2 = t.x, beta2 = t.y, ab = t.z, factor = t.w;
    const float axr = (x0.x + x1.x * (2.0f / 3.0f)) + x2.x * (1.0f / 3.0f);
    const float axg = (x0.y + x1.y * (2.0f / 3.0f)) + x2.y * (1.0f / 3.0f);
    const float axb = (x0.z + x1.z * (2.0f / 3.0f)) + x2.z * (1.0f / 3.0f);
    const float bxr = xsum.x - axr, bxg = xsum.y - axg, bxb = xsum.z - axb;
    float ar = (axr * beta2 - bxr * ab) * factor;
    float ag = (axg * beta2 - bxg * ab) * factor;
    float abl = (axb * beta2 - bxb * ab) * factor;
    float br = (bxr * alpha2 - axr * ab) * factor;
    float bg = (bxg * alpha2 - axg * ab) * factor;
    float bb = (bxb * alpha2 - axb * ab) * factor;
    ar = smin(1.0f, smax(0.0f, ar)); ag = smin(1.0f, smax(0.0f, ag)); abl = smin(1.0f, smax(0.0f, abl));
    br = smin(1.0f, smax(0.0f, br)); bg = smin(1.0f, smax(0.0f, bg)); bb = smin(1.0f, smax(0.0f, bb));
    ar = floorf(31.0f * ar + 0.5f) * (1.0f / 31.0f);
    ag = floorf(63.0f * ag + 0.5f) * (1.0f / 63.0f);
    abl = floorf(31.0f * abl + 0.5f) * (1.0f / 31.0f);
    br = floorf(31.0f * br + 0.5f) * (1.0f / 31.0f);
    bg = floorf(63.0f * bg + 0.5f) * (1.0f / 63.0f);
    bb = floorf(31.0f * bb + 0.5f) * (1.0f / 31.0f);
    const float e1r = ((ar * ar) * alpha2 + (br * br) * beta2) + 2.0f * ((((ar * br) * ab) - ar * axr) - br * bxr);
    const float e1g = ((ag * ag) * alpha2 + (bg * bg) * beta2) + 2.0f * ((((ag * bg) * ab) - ag * axg) - bg * bxg);
    const float e1b = ((abl * abl) * alpha2 + (bb * bb) * beta2) + 2.0f * ((((abl * bb) * ab) - abl * axb) - bb * bxb);
    if (out) {
        out->ax = ar; out->ay = ag; out->az = abl;
        out->bx = br; out->by = bg; out->bz = bb;
    }
    return e1r * mx + e1g * my + e1b * mz;
}

__device__ __forceinline__ float sq_eval(float4 x0, float4 x1, float4 x2, float4 xsum, float mx, float my, float mz,
                                         SqCand *out)
{
    return sq_eval_t(x0, x1, x2, xsum, sq_weights(x0.w, x1.w, x2.w, xsum.w), mx, my, mz, out);
}

// The scalar Compress3 loop body (weightedclusterfit.cpp:392-430): palette
// {a, b, (a + b) / 2}, cluster weights w0 (a), w1 (the midpoint), w2 (b).
__device__ __forceinline__ float sq_eval3(float4 x0, float4 x1, float4 xsum, float mx, float my, float mz, SqCand *out)
{
    const float w2 = (xsum.w - x0.w) - x1.w;
    const float alpha2 = x0.w + x1.w * 0.25f;
    const float beta2 = w2 + x1.w * 0.25f;
    const float ab = x1.w * 0.25f;
    const float factor = 1.0f / (alpha2 * beta2 - ab * ab);
    const float axr = x0.x + x1.x * 0.5f, axg = x0.y + x1.y * 0.5f, axb = x0.z + x1.z * 0.5f;
    const float bxr = xsum.x - axr, bxg = xsum.y - axg, bxb = xsum.z - axb;
    float ar = (axr * beta2 - bxr * ab) * factor;
    float ag = (axg * beta2 - bxg * ab) * factor;
    float abl = (axb * beta2 - bxb * ab) * factor;
    float br = (bxr * alpha2 - axr * ab) * factor;
    float bg = (bxg * alpha2 - axg * ab) * factor;
    float bb = (bxb * alpha2 - axb * ab) * factor;
    ar = smin(1.0f, smax(0.0f, ar)); ag = smin(1.0f, smax(0.0f, ag)); abl = smin(1.0f, smax(0.0f, abl));
    br = smin(1.0f, smax(0.0f, br)); bg = smin(1.0f, smax(0.0f, bg)); bb = smin(1.0f, smax(0.0f, bb));
    ar = floorf(31.0f * ar + 0.5f) * (1.0f / 31.0f);
    ag = floorf(63.0f * ag + 0.5f) * (1.0f / 63.0f);
    abl = floorf(31.0f * abl + 0.5f) * (1.0f / 31.0f);
    br = floorf(31.0f * br + 0.5f) * (1.0f / 31.0f);
    bg = floorf(63.0f * bg + 0.5f) * (1.0f / 63.0f);
    bb = floorf(31.0f * bb + 0.5f) * (1.0f / 31.0f);
    const float e1r = ((ar * ar) * alpha2 + (br * br) * beta2) + 2.0f * ((((ar * br) * ab) - ar * axr) - br * bxr);
    const float e1g = ((ag * ag) * alpha2 + (bg * bg) * beta2) + 2.0f * ((((ag * bg) * ab) - ag * axg) - bg * bxg);
    const float e1b = ((abl * abl) * alpha2 + (bb * bb) * beta2) + 2.0f * ((((abl * bb) * ab) - abl * axb) - bb * bxb);
    if (out) {
        out->ax = ar; out->ay = ag; out->az = abl;
        out->bx = br; out->by = bg; out->bz = bb;
    }
    return e1r * mx + e1g * my + e1b * mz;
}

__device__ __forceinline__ float4 add4(float4 a, float4 b)
{
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// ColourSet minimal set, weighted PCA and the stable sort of
// WeightedClusterFit (weightedclusterfit.cpp:39-106) for a block whose 8-bit
// texels (ColorBlock::init's gather, stored by the classify pass) are
// `texels`.  Leaves the weighted points in sorted order in column `lane` of
// s_p (rows from `count` on are zero) and returns the point count, the sum
// of the weighted points, the sorted order (op) and each texel's point (pidx).
// dxt1 (the kDxt1 colour set, colourset.cpp:44-51): texels of alpha 0 stay
// out of the set (pidx -1, index 3 when written) and set `transparent`.
__device__ __forceinline__ int bc3_prepare(float4 (*s_p)[64], int lane, const uint4 *__restrict__ texels,
                                           int weight_by_alpha, float cwx, float cwy, float cwz, float4 &xsum_out,
                                           int op[16], int pidx[16], bool dxt1, bool &transparent)
{
    uint32_t rgb[16];
    float wt[16];
    bool clear[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint4 v = texels[q];
        const uint32_t t4[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
        for (int j = 0; j < 4; j++) {
            rgb[4 * q + j] = t4[j] & 0xffffffu;
            wt[4 * q + j] = weight_by_alpha ? float((t4[j] >> 24) + 1) / 256.0f : 1.0f;
            clear[4 * q + j] = dxt1 && (t4[j] >> 24) == 0;
        }
    }

    // ColourSet minimal set (colourset.cpp:35-101): a texel whose RGB bytes
    // repeat an earlier texel of the set adds its weight to that point, in
    // texel order.
    int count = 0;
    transparent = false;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (clear[i]) {
            pidx[i] = -1;
            transparent = true;
            continue;
        }
        int p = count;
        bool d = true;
#pragma unroll
        for (int j = i - 1; j >= 0; j--)
            if (!clear[j] && rgb[j] == rgb[i]) { d = false; p = pidx[j]; }
        pidx[i] = p;
        if (d) {
            s_p[p][lane] = make_float4(float(rgb[i] & 0xff) / 255.0f, float((rgb[i] >> 8) & 0xff) / 255.0f,
                                       float(rgb[i] >> 16) / 255.0f, wt[i]);
            count++;
        } else {
            s_p[p][lane].w += wt[i];
        }
    }

    // ComputeWeightedCovariance (maths.cpp:32-59), metric on the deltas.
    float4 P[16];
#pragma unroll
    for (int p = 0; p < 16; p++) P[p] = s_p[p][lane];
    float total = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
#pragma unroll
    for (int p = 0; p < 16; p++)
        if (p < count) {
            total += P[p].w;
            cx += P[p].x * P[p].w; cy += P[p].y * P[p].w; cz += P[p].z * P[p].w;
        }
    const float tinv = 1.0f / total;
    cx *= tinv; cy *= tinv; cz *= tinv;
    float m0 = 0, m1 = 0, m2 = 0, m3 = 0, m4 = 0, m5 = 0;
#pragma unroll
    for (int p = 0; p < 16; p++)
        if (p < count) {
            const float ax = (P[p].x - cx) * cwx, ay = (P[p].y - cy) * cwy, az = (P[p].z - cz) * cwz;
            const float bxw = ax * P[p].w, byw = ay * P[p].w, bzw = az * P[p].w;
            m0 += ax * bxw; m1 += ax * byw; m2 += ax * bzw;
            m3 += ay * byw; m4 += ay * bzw; m5 += az * bzw;
        }
    // ComputePrincipleComponent, scalar build (maths.cpp:109-131).
    float vx, vy, vz;
    {
        const float r0 = m0 * m0 + m1 * m1 + m2 * m2;
        const float r1 = m1 * m1 + m3 * m3 + m4 * m4;
        const float r2 = m2 * m2 + m4 * m4 + m5 * m5;
        if (r0 > r1 && r0 > r2) { vx = m0; vy = m1; vz = m2; }
        else if (r1 > r2) { vx = m1; vy = m3; vz = m4; }
        else { vx = m2; vy = m4; vz = m5; }
        for (int it = 0; it < 8; it++) {
            const float x = vx * m0 + vy * m1 + vz * m2;
            const float y = vx * m1 + vy * m3 + vz * m4;
            const float z = vx * m2 + vy * m4 + vz * m5;
            const float norm = smax(smax(x, y), z);
            const float iv = 1.0f / norm;
            if (norm == 0.0f) { vx = vy = vz = 0.0f; break; }
            vx = x * iv; vy = y * iv; vz = z * iv;
        }
    }
    // Stable insertion sort by projection (weightedclusterfit.cpp:62-77),
    // written as the network it executes: each pass bubbles element i down
    // while it is strictly smaller than its left neighbour.
    float dps[16];
#pragma unroll
    for (int p = 0; p < 16; p++) {
        dps[p] = P[p].x * vx + P[p].y * vy + P[p].z * vz;
        op[p] = p;
    }
#pragma unroll
    for (int i = 1; i < 16; i++) {
        bool go = i < count;
#pragma unroll
        for (int j = i; j > 0; j--) {
            go = go && dps[j] < dps[j - 1];
            if (go) {
                const float td = dps[j]; dps[j] = dps[j - 1]; dps[j - 1] = td;
                const int to = op[j]; op[j] = op[j - 1]; op[j - 1] = to;
            }
        }
    }
    // Weighted points in sorted order (weightedclusterfit.cpp:92-104).
    float4 wp[16];
    float4 xsum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float4 q = s_p[op[i]][lane];
        wp[i] = make_float4(q.w * q.x, q.w * q.y, q.w * q.z, q.w);
        if (i < count) xsum = add4(xsum, wp[i]);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) s_p[i][lane] = i < count ? wp[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    s_p[16][lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);

    xsum_out = xsum;
    return count;
}

// Endpoints of the winning partition from the same running sums, indices
// by sorted position through the minimal set, WriteColourBlock4
// (colourblock.cpp:108-138).
__device__ __forceinline__ void bc3_write(float4 (*s_p)[64], int lane, const int op[16], const int pidx[16], float4 xsum,
                                          float msx, float msy, float msz, float besterror, int b0, int b1, int b2,
                                          uint2 *dst)
{
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (!(besterror < FLT_MAX)) {
        *dst = make_uint2(0u, 0u);  // the reference leaves the block unwritten here
        return;
    }
    if (b2 < 0) {
        // Compress3's winner (c0, c1) = (b0, b1): sorted positions below b0
        // take index 0, the next b1 index 2, the rest 1, texels outside the
        // set 3; WriteColourBlock3 (colourblock.cpp:75-106) orders the
        // endpoints a <= b, swapping indices 0 and 1.
        float4 y0 = zero, y1 = zero;
        for (int i = 0; i < b0; i++) y0 = add4(y0, s_p[i][lane]);
        for (int i = b0; i < b0 + b1; i++) y1 = add4(y1, s_p[i][lane]);
        SqCand c;
        sq_eval3(y0, y1, xsum, msx, msy, msz, &c);
        int ca = (float_to_int(31.0f * c.ax, 31) << 11) | (float_to_int(63.0f * c.ay, 63) << 5) | float_to_int(31.0f * c.az, 31);
        int cb = (float_to_int(31.0f * c.bx, 31) << 11) | (float_to_int(63.0f * c.by, 63) << 5) | float_to_int(31.0f * c.bz, 31);
        uint32_t bits = 0;
#pragma unroll
        for (int t = 0; t < 16; t++) {
            uint32_t idx = 3u;
            if (pidx[t] >= 0) {
                int pos = 0;
#pragma unroll
                for (int i = 0; i < 16; i++) pos += (op[i] == pidx[t]) ? i : 0;
                idx = pos < b0 ? 0u : pos < b0 + b1 ? 2u : 1u;
                if (ca > cb) idx = idx < 2u ? idx ^ 1u : idx;
            }
            bits |= idx << (2 * t);
        }
        if (ca > cb) { const int t = ca; ca = cb; cb = t; }
        *dst = make_uint2(uint32_t(ca) | uint32_t(cb) << 16, bits);
        return;
    }
    // The winner's endpoints from the same running sums.
    float4 y0 = zero, y1 = zero, y2 = zero;
    for (int i = 0; i < b0; i++) y0 = add4(y0, s_p[i][lane]);
    for (int i = b0; i < b0 + b1; i++) y1 = add4(y1, s_p[i][lane]);
    for (int i = b0 + b1; i < b0 + b1 + b2; i++) y2 = add4(y2, s_p[i][lane]);
    SqCand c;
    sq_eval(y0, y1, y2, xsum, msx, msy, msz, &c);

    // Indices by sorted position (remap through the minimal set), then
    // WriteColourBlock4 (colourblock.cpp:108-138).
    int ca = (float_to_int(31.0f * c.ax, 31) << 11) | (float_to_int(63.0f * c.ay, 63) << 5) | float_to_int(31.0f * c.az, 31);
    int cb = (float_to_int(31.0f * c.bx, 31) << 11) | (float_to_int(63.0f * c.by, 63) << 5) | float_to_int(31.0f * c.bz, 31);
    uint32_t bits = 0;
#pragma unroll
    for (int t = 0; t < 16; t++) {
        int pos = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) pos += (op[i] == pidx[t]) ? i : 0;
        uint32_t idx = pos < b0 ? 0u : pos < b0 + b1 ? 2u : pos < b0 + b1 + b2 ? 3u : 1u;
        if (ca < cb) idx = (idx ^ 1u) & 3u;
        else if (ca == cb) idx = 0u;
        bits |= idx << (2 * t);
    }
    if (ca < cb) { const int t = ca; ca = cb; cb = t; }
    *dst = make_uint2(uint32_t(ca) | uint32_t(cb) << 16, bits);
}

// dxt1 (BC1a, the kDxt1 fit, colourfit.cpp:41-57): Compress3 first
// (weightedclusterfit.cpp:372-474), then Compress4 unless a texel is
// transparent, which replaces it only with a strictly lower error; lists
// from k = 1.
__global__ __launch_bounds__(64) void bc3_search_kernel(const uint4 *__restrict__ packed, int weight_by_alpha,
                                                        float cwx, float cwy, float cwz, int stride, int offset,
                                                        uint8_t *__restrict__ out, int *__restrict__ counts,
                                                        const int *__restrict__ lists, int cap,
                                                        const ntx::Tables *__restrict__ tab, int dxt1)
{
    // Column `lane` holds this lane's points; row 16 stays zero (the
    // reference's loops read one element past the end and discard it).
    __shared__ float4 s_p[17][64];
    const int lane = threadIdx.x;

    // Work item -> (colour count k, position in list k), k = 16 first.  A
    // list never holds more than cap entries, nor indices at or above cap.
    int item = blockIdx.x * blockDim.x + lane;
    int k = 0, pos_in = 0;
    const int kmin = dxt1 ? 1 : 2;
    for (int b = 16; b >= kmin; b--) {
        const int c = min(counts[b], cap);
        if (k == 0) {
            if (item < c) { k = b; pos_in = item; }
            else item -= c;
        }
    }
    int blk = k != 0 ? lists[size_t(k) * cap + pos_in] : 0;
    const bool live = k != 0 && unsigned(blk) < unsigned(cap);
    int count = 0;
    int op[16], pidx[16];
    bool transparent = false;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 xsum = zero;
    if (live) {
        count = bc3_prepare(s_p, lane, packed + size_t(blk) * 4, weight_by_alpha, cwx, cwy, cwz, xsum, op, pidx,
                            dxt1 != 0, transparent);
    }
    const float msx = cwx * cwx, msy = cwy * cwy, msz = cwz * cwz;

    // Compress3 (kDxt1 only): first strict minimum over (c0, c1).
    float best3 = FLT_MAX;
    int d0 = 0, d1 = 0;
    if (live && dxt1) {
        float4 x0 = zero;
        for (int c0 = 0; c0 <= count; c0++) {
            float4 x1 = zero;
            for (int c1 = 0; c1 <= count - c0; c1++) {
                const float e = sq_eval3(x0, x1, xsum, msx, msy, msz, nullptr);
                if (e < best3) { best3 = e; d0 = c0; d1 = c1; }
                x1 = add4(x1, s_p[c0 + c1][lane]);  // row `count` is zero
            }
            x0 = add4(x0, s_p[c0][lane]);
        }
    }

    // Compress4 (weightedclusterfit.cpp:495-553): first strict minimum in loop order.
    const bool live4 = live && !transparent;
    float besterror = FLT_MAX;
    int b0 = 0, b1 = 0, b2 = 0;
    // A wave whose live lanes all hold 16 points of weight 1 takes the
    // weight terms of each candidate from sq16 (scalar loads shared by the
    // wave); idle lanes run along on whatever their LDS column holds and
    // write nothing.
    const uint64_t live_mask = __builtin_amdgcn_ballot_w64(live4);
    const uint64_t unit_mask = __builtin_amdgcn_ballot_w64(live4 && count == 16 && !weight_by_alpha);
    if (live_mask != 0 && unit_mask == live_mask) {
        const float4 *__restrict__ tw = reinterpret_cast<const float4 *>(tab->sq16);
        int idx = 0;
        float4 x0 = zero;
        for (int c0 = 0; c0 <= 16; c0++) {
            float4 x1 = zero;
            for (int c1 = 0; c1 <= 16 - c0; c1++) {
                float4 x2 = zero;
                for (int c2 = 0; c2 <= 16 - c0 - c1; c2++, idx++) {
                    const float e = sq_eval_t(x0, x1, x2, xsum, tw[idx], msx, msy, msz, nullptr);
                    if (e < besterror) { besterror = e; b0 = c0; b1 = c1; b2 = c2; }
                    x2 = add4(x2, s_p[c0 + c1 + c2][lane]);
                }
                x1 = add4(x1, s_p[c0 + c1][lane]);
            }
            x0 = add4(x0, s_p[c0][lane]);
        }
    } else if (live4 && !weight_by_alpha) {
        // Integer weights (point multiplicities): each candidate's terms come
        // from sqw by the exact cluster weights; the next candidate's entry
        // is loaded while this one is evaluated.
        const float4 *__restrict__ tw = reinterpret_cast<const float4 *>(tab->sqw);
        float4 x0 = zero;
        for (int c0 = 0; c0 <= count; c0++) {
            float4 x1 = zero;
            for (int c1 = 0; c1 <= count - c0; c1++) {
                const int base = (int(x0.w) * 17 + int(x1.w)) * 17;
                float4 x2 = zero;
                float4 t = tw[base];
                for (int c2 = 0; c2 <= count - c0 - c1; c2++) {
                    const float4 pn = s_p[c0 + c1 + c2][lane];  // row `count` is zero
                    const float4 tn = tw[base + int(x2.w + pn.w)];
                    const float e = sq_eval_t(x0, x1, x2, xsum, t, msx, msy, msz, nullptr);
                    if (e < besterror) { besterror = e; b0 = c0; b1 = c1; b2 = c2; }
                    x2 = add4(x2, pn);
                    t = tn;
                }
                x1 = add4(x1, s_p[c0 + c1][lane]);
            }
            x0 = add4(x0, s_p[c0][lane]);
        }
    } else if (live4) {
        float4 x0 = zero;
        for (int c0 = 0; c0 <= count; c0++) {
            float4 x1 = zero;
            for (int c1 = 0; c1 <= count - c0; c1++) {
                float4 x2 = zero;
                for (int c2 = 0; c2 <= count - c0 - c1; c2++) {
                    const float e = sq_eval(x0, x1, x2, xsum, msx, msy, msz, nullptr);
                    if (e < besterror) { besterror = e; b0 = c0; b1 = c1; b2 = c2; }
                    x2 = add4(x2, s_p[c0 + c1 + c2][lane]);
                }
                x1 = add4(x1, s_p[c0 + c1][lane]);
            }
            x0 = add4(x0, s_p[c0][lane]);
        }
    }
    // Lanes that skipped Compress4 but ran along in the wave-uniform loops
    // hold no result of it.
    if (!live4) besterror = FLT_MAX;
    // Compress4 replaces Compress3's block only with a strictly lower error
    // (ColourFit::m_besterror); b2 = -1 marks Compress3's winner.
    if (dxt1 && !(besterror < best3)) {
        besterror = best3;
        b0 = d0;
        b1 = d1;
        b2 = -1;
    }
    if (live)
        bc3_write(s_p, lane, op, pidx, xsum, msx, msy, msz, besterror, b0, b1, b2,
                  reinterpret_cast<uint2 *>(out + size_t(blk) * stride + offset));
    release_counts(counts);
}

// The same search with 2^lpb_shift lanes per block, for images too small to
// fill the device one block per lane: lane `part` of a block scans a
// contiguous slice of the reference's (c0, c1, c2) loop order.  Its running
// sums start as sequential sums from zero, which is what the reference's
// running sums are at that point of the loop, and then advance exactly as
// the nested loops do; the slices' winners are combined preferring the
// earlier slice on equal error, which is the first strict minimum overall.
__global__ __launch_bounds__(64) void bc3_search_split_kernel(const uint4 *__restrict__ packed, int weight_by_alpha,
                                                              float cwx, float cwy, float cwz, int stride, int offset,
                                                              uint8_t *__restrict__ out, int *__restrict__ counts,
                                                              const int *__restrict__ lists, int cap,
                                                              const ntx::Tables *__restrict__ tab, int lpb_shift,
                                                              int urgent)
{
    if (urgent) __builtin_amdgcn_s_setprio(3);
    __shared__ float4 s_p[17][64];
    const int lane = threadIdx.x;
    const int part = lane & ((1 << lpb_shift) - 1);
    int item = (blockIdx.x * blockDim.x + lane) >> lpb_shift;
    int k = 0, pos_in = 0;
    for (int b = 16; b >= 2; b--) {
        const int c = min(counts[b], cap);
        if (k == 0) {
            if (item < c) { k = b; pos_in = item; }
            else item -= c;
        }
    }
    const int listed = k != 0 ? lists[size_t(k) * cap + pos_in] : 0;
    const bool live = k != 0 && unsigned(listed) < unsigned(cap);
    const float msx = cwx * cwx, msy = cwy * cwy, msz = cwz * cwz;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float besterror = FLT_MAX;
    int best = 0;  // c0 | c1 << 8 | c2 << 16
    int blk = 0;
    int op[16], pidx[16];
    float4 xsum = zero;
    if (live) {
        blk = listed;
        bool transparent;
        const int count = bc3_prepare(s_p, lane, packed + size_t(blk) * 4, weight_by_alpha, cwx, cwy, cwz, xsum, op,
                                      pidx, false, transparent);
        const int first = tab->sq_off[count], total = tab->sq_off[count + 1] - first;
        const int begin = (part * total) >> lpb_shift, end = ((part + 1) * total) >> lpb_shift;
        if (begin < end) {
            const uint32_t pk = tab->sq[first + begin];
            int c0 = int(pk & 0xff), c1 = int((pk >> 8) & 0xff), c2 = int((pk >> 16) & 0xff);
            float4 x0 = zero, x1 = zero, x2 = zero;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                if (i < c0) x0 = add4(x0, s_p[i][lane]);
                else if (i < c0 + c1) x1 = add4(x1, s_p[i][lane]);
                else if (i < c0 + c1 + c2) x2 = add4(x2, s_p[i][lane]);
            }
            for (int n = begin; n < end; n++) {
                const float e = sq_eval(x0, x1, x2, xsum, msx, msy, msz, nullptr);
                if (e < besterror) { besterror = e; best = c0 | c1 << 8 | c2 << 16; }
                // Next (c0, c1, c2) in the nested loops' order, with their updates.
                if (c0 + c1 + c2 < count) {
                    x2 = add4(x2, s_p[c0 + c1 + c2][lane]);
                    c2++;
                } else if (c0 + c1 < count) {
                    x1 = add4(x1, s_p[c0 + c1][lane]);
                    c1++;
                    c2 = 0;
                    x2 = zero;
                } else {
                    x0 = add4(x0, s_p[c0][lane]);
                    c0++;
                    c1 = c2 = 0;
                    x1 = x2 = zero;
                }
            }
        }
    }
    // Butterfly over the block's lanes, narrowest first: before the step with
    // offset `off` a lane holds the winner of the off aligned consecutive
    // slices around it, and the partner's group lies entirely before its own
    // in loop order exactly when this lane's `off` bit is set.
    for (int off = 1; off < (1 << lpb_shift); off <<= 1) {
        const float oe = __shfl_xor(besterror, off);
        const int ob = __shfl_xor(best, off);
        if (oe < besterror || (oe == besterror && (part & off))) { besterror = oe; best = ob; }
    }
    release_counts(counts);
    if (!live || part != 0) return;
    bc3_write(s_p, lane, op, pidx, xsum, msx, msy, msz, besterror, best & 0xff, (best >> 8) & 0xff, best >> 16,
              reinterpret_cast<uint2 *>(out + size_t(blk) * stride + offset));
}

}  // namespace

namespace ntx {

// BC4 (R), BC5 (R then G) or the BC3 alpha half (A), one thread per channel block.
hipError_t launch_channels(const float *d_src, int w, int h, int format, void *d_out, hipStream_t stream, bool urgent)
{
    const int bw = (w + 3) / 4, bh = (h + 3) / 4, nblocks = bw * bh;
    if (nblocks == 0) return hipSuccess;
    int njobs = 1, ch0 = 0, ch1 = 0, stride = 8, off1 = 0;
    if (format == NTX_FORMAT_BC5) { njobs = 2; ch0 = 0; ch1 = 1; stride = 16; off1 = 8; }
    else if (format == NTX_FORMAT_BC3) { ch0 = 3; stride = 16; }
    const int threads = nblocks * njobs;
    hipLaunchKernelGGL(channel_kernel, dim3((threads + 255) / 256), dim3(256), 0, stream, d_src, w, h, bw, nblocks,
                       njobs, ch0, ch1, stride, off1, reinterpret_cast<uint8_t *>(d_out), urgent ? 1 : 0);
    return hipGetLastError();
}

// BC3 pass 1: alpha halves, single-colour blocks, bucketing of the rest.
// urgent: the waves raise their issue priority (s_setprio), for small
// latency-bound levels running beside large searches that would otherwise
// take most of each SIMD's issue slots.
// counts: the bc3_lists_offset() ints before the lists, zero on entry
// (cleared by the previous search or by launch_bc3_clear_counts); lists:
// kCountSlots x cap ints.
hipError_t launch_bc3_front(const float *d_src, int w, int h, const Tables *d_tab, void *d_out, int *d_counts,
                            int *d_lists, int cap, int weight_by_alpha, uint32_t *d_packed, bool urgent,
                            hipStream_t stream, bool normal)
{
    const int bw = (w + 3) / 4, bh = (h + 3) / 4, nblocks = bw * bh;
    if (nblocks == 0) return hipSuccess;
    const int wide = (reinterpret_cast<uintptr_t>(d_out) & 15) == 0;
    // Every row of every plane starts 16-byte aligned: float4 row loads.
    const int rows16 = (reinterpret_cast<uintptr_t>(d_src) & 15) == 0 && (w & 3) == 0;
    hipLaunchKernelGGL(normal ? bc3_front_kernel<true> : bc3_front_kernel<false>, dim3((nblocks + 255) / 256), dim3(256),
                       0, stream, d_src, w, h, bw, nblocks, d_tab, reinterpret_cast<uint8_t *>(d_out), wide, d_counts,
                       d_lists, cap, weight_by_alpha, reinterpret_cast<uint4 *>(d_packed), urgent ? 1 : 0, rows16);
    return hipGetLastError();
}

// BC1a (CompressorDXT1a): pass 1 finishes single-colour and empty blocks and
// buckets the others; pass 2 is the kDxt1 fit, one block per lane, into
// 8-byte blocks.  Its last workgroup clears the tallies.
hipError_t launch_bc1a(const float *d_src, int w, int h, int weight_by_alpha, const float *cw, const Tables *d_tab,
                       void *d_out, int *d_counts, int *d_lists, int cap, uint32_t *d_packed, hipStream_t stream)
{
    const int bw = (w + 3) / 4, bh = (h + 3) / 4, nblocks = bw * bh;
    if (nblocks == 0) return hipSuccess;
    uint4 *packed = reinterpret_cast<uint4 *>(d_packed);
    hipLaunchKernelGGL(bc1a_front_kernel, dim3((nblocks + 255) / 256), dim3(256), 0, stream, d_src, w, h, bw, nblocks,
                       d_tab, reinterpret_cast<uint8_t *>(d_out), d_counts, d_lists, cap, packed);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bc3_search_kernel, dim3((nblocks + 63) / 64), dim3(64), 0, stream, packed, weight_by_alpha,
                       cw[0], cw[1], cw[2], 8, 0, reinterpret_cast<uint8_t *>(d_out), d_counts, d_lists, cap, d_tab, 1);
    return hipGetLastError();
}

// Zeroes the tallies and the ticket (a fill kernel, so the sequence stays
// capturable in a graph); needed only for fresh scratch memory or after an
// interrupted sequence.
hipError_t launch_bc3_clear_counts(int *d_counts, hipStream_t stream)
{
    hipLaunchKernelGGL(zero_counts_kernel, dim3(1), dim3(64), 0, stream, d_counts);
    return hipGetLastError();
}

// BC3 colour, pass 2: 2^lpb_shift lanes per block over the bucketed lists
// (one block per lane when lpb_shift is 0); its last workgroup clears the
// tallies.
hipError_t launch_bc3_search(const uint32_t *d_packed, int w, int h, int weight_by_alpha, const float *cw, void *d_out,
                             int *d_counts, const int *d_lists, int cap, const Tables *d_tab, int lpb_shift,
                             bool urgent, hipStream_t stream)
{
    const int bw = (w + 3) / 4, bh = (h + 3) / 4, nblocks = bw * bh;
    if (nblocks == 0) return hipSuccess;
    const uint4 *packed = reinterpret_cast<const uint4 *>(d_packed);
    if (lpb_shift <= 0) {
        hipLaunchKernelGGL(bc3_search_kernel, dim3((nblocks + 63) / 64), dim3(64), 0, stream, packed, weight_by_alpha,
                           cw[0], cw[1], cw[2], 16, 8, reinterpret_cast<uint8_t *>(d_out), d_counts, d_lists, cap,
                           d_tab, 0);
    } else {
        const size_t lanes = size_t(nblocks) << lpb_shift;
        hipLaunchKernelGGL(bc3_search_split_kernel, dim3(unsigned((lanes + 63) / 64)), dim3(64), 0, stream, packed,
                           weight_by_alpha, cw[0], cw[1], cw[2], 16, 8, reinterpret_cast<uint8_t *>(d_out), d_counts,
                           d_lists, cap, d_tab, lpb_shift, urgent ? 1 : 0);
    }
    return hipGetLastError();
}

// Scratch of the BC3 colour passes, in 32-bit words: 32 count slots, the
// kCountSlots lists of up to nblocks entries, then (16-byte aligned) 16
// packed texels per block.
int bc3_lists_offset() { return kListsOffset; }
int bc3_packed_offset(int nblocks) { return (kListsOffset + kCountSlots * nblocks + 3) & ~3; }
int bc3_scratch_ints(int nblocks) { return bc3_packed_offset(nblocks) + 16 * nblocks; }

}  // namespace ntx
