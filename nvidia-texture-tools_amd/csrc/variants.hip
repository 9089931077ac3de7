// variants.hip -- the Production / Highest / Fastest block compressors of
// BC3, BC3n, BC4 and BC5 for gfx950 (Context.cpp:1075-1116 picks them),
// bit-exact with NVTT's CPU code:
//   OptimalCompress::compressDXT5A (OptimalCompressDXT.cpp:512-607): BC4 and
//     BC5 at Production and Highest (ProductionCompressorBC4/BC5,
//     CompressorDX10.cpp:65-87), the BC3 alpha and the BC3n red block at
//     Highest (CompressorDX9.cpp:149-152, 214-216).  One wavefront per block:
//     the lanes split the (alpha0, alpha1) pairs of both search loops.
//   OptimalCompress::compressDXT1G (OptimalCompressDXT.cpp:315-390): the BC3n
//     colour block at Highest (CompressorDX9.cpp:184-186).  One wavefront
//     per block over the (g0, g1) pairs.
//   QuickCompress::compressDXT5 (QuickCompressDXT.cpp:653-713, 772-827): BC3
//     and BC3n at Fastest (FastCompressorDXT5 / DXT5n, CompressorDX9.cpp:67-80).
//     One thread per block.
// Input is gathered as ColorBlock::init does (ColorBlock.cpp:80-110).
#include "channel_block.hpp"

#include <algorithm>
#include <cfloat>

namespace {

using namespace ntx_bc4;

// ---------------------------------------------------------------- helpers --

// Pair k of a triangular search loop whose row m holds m + 1 pairs: k =
// m (m + 1) / 2 + j, 0 <= j <= m.
__device__ __forceinline__ void tri_pair(int k, int &m, int &j)
{
    int r = int((sqrtf(8.0f * float(k) + 1.0f) - 1.0f) * 0.5f);
    if (r * (r + 1) / 2 > k) r--;
    if ((r + 1) * (r + 2) / 2 <= k) r++;
    m = r;
    j = k - r * (r + 1) / 2;
}

// The smaller (error, pair index) over the wavefront: the first strict
// minimum of the sequential loop when every lane visits its pairs in
// increasing index order.
__device__ __forceinline__ void wave_argmin(uint32_t &err, int &k)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t oe = uint32_t(__shfl_xor(int(err), off));
        const int ok = __shfl_xor(k, off);
        if (oe < err || (oe == err && ok < k)) { err = oe; k = ok; }
    }
}

// -------------------------------------------- OptimalCompress::compressDXT5A --

// computeAlphaError (OptimalCompressDXT.cpp:189-216) with unit weights: the
// squared distance of each texel to its nearest palette entry, summed (the
// reference sums these integers in float, exactly).  v: the block's values,
// the same in every lane.
__device__ __forceinline__ uint32_t alpha_error(const uint32_t v[16], uint32_t a0, uint32_t a1)
{
    uint32_t pal[8];
    alpha_palette(a0, a1, pal);
    uint32_t total = 0;
#pragma unroll
    for (int t = 0; t < 16; t++) {
        uint32_t best = 0xffffffffu;
#pragma unroll
        for (int p = 0; p < 8; p++) {
            const int d = int(pal[p]) - int(v[t]);
            best = min(best, uint32_t(d * d));
        }
        total += best;
    }
    return total;
}

// OptimalCompress::compressDXT5A on one channel block, by one wavefront.
// Blocks spanning under 8 values, or under 6 apart from 0 and 255, take
// their extremes (lines 531-542).  Otherwise the reference tries every
// eight-value pair (a0, a1), a0 - a1 > 8, in the range widened by 8, then
// every six-value pair (a1, a0) in the range of the values other than 0 and
// 255 widened by 6, keeping the first strictly lower error; its first error
// seed is read from the output block before the call (line 546), and here
// it is the error of (max, min), the endpoints it keeps when no pair wins.
__device__ uint64_t optimal_channel(const uint32_t vq[4], int lane)
{
    uint32_t v[16];
    uint32_t mina = 255, maxa = 0, mina01 = 255, maxa01 = 0;
#pragma unroll
    for (int t = 0; t < 16; t++) {
        v[t] = (vq[t >> 2] >> (8 * (t & 3))) & 0xffu;
        mina = min(mina, v[t]);
        maxa = max(maxa, v[t]);
        if (v[t] != 0 && v[t] != 255) {
            mina01 = min(mina01, v[t]);
            maxa01 = max(maxa01, v[t]);
        }
    }
    uint32_t e0 = maxa, e1 = mina;
    if (int(maxa) - int(mina) < 8) {
        // e0, e1 as they are
    } else if (int(maxa01) - int(mina01) < 6) {
        e0 = mina01;
        e1 = maxa01;
    } else {
        const uint32_t seed = alpha_error(v, maxa, mina);
        const int lo = mina <= 8 ? 0 : int(mina) - 8, hi = maxa >= 247 ? 255 : int(maxa) + 8;
        const int lo6 = mina01 <= 6 ? 0 : int(mina01) - 6, hi6 = maxa01 >= 249 ? 255 : int(maxa01) + 6;
        const int n1 = max(hi - lo - 9, 0), n2 = max(hi6 - lo6 - 9, 0);
        const int t1 = n1 * (n1 + 1) / 2, total = t1 + n2 * (n2 + 1) / 2;
        uint32_t best = 0xffffffffu;
        int bk = 0x7fffffff;
        for (int base = 0; base < total; base += 64) {
            const int k = base + lane;
            if (k < total) {
                int m, j;
                uint32_t a0, a1;
                if (k < t1) {
                    tri_pair(k, m, j);
                    a0 = uint32_t(lo + 9 + m);
                    a1 = uint32_t(lo + j);
                } else {
                    tri_pair(k - t1, m, j);
                    a0 = uint32_t(lo6 + j);
                    a1 = uint32_t(lo6 + 9 + m);
                }
                const uint32_t e = alpha_error(v, a0, a1);
                if (e < best) { best = e; bk = k; }
            }
        }
        wave_argmin(best, bk);
        if (best < seed) {
            int m, j;
            if (bk < t1) {
                tri_pair(bk, m, j);
                e0 = uint32_t(lo + 9 + m);
                e1 = uint32_t(lo + j);
            } else {
                tri_pair(bk - t1, m, j);
                e0 = uint32_t(lo6 + j);
                e1 = uint32_t(lo6 + 9 + m);
            }
        }
    }
    // computeAlphaIndices (OptimalCompressDXT.cpp:219-246): nearest entry,
    // first on ties, as alpha_indices computes it.
    uint64_t blk = uint64_t(e0) | uint64_t(e1) << 8;
    alpha_indices(vq, blk);
    return blk;
}

// One wavefront per (block, channel job), a grid-stride loop over them.
// Job 0 reads plane ch0 and writes at byte off0 of the block's `stride`
// bytes, job 1 plane ch1 at off1.
__global__ __launch_bounds__(256) void optimal_channel_kernel(const float *__restrict__ src, int w, int h, int bw,
                                                              int nblocks, int njobs, int ch0, int ch1, int stride,
                                                              int off0, int off1, uint8_t *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(int((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int waves = int((gridDim.x * blockDim.x) >> 6);
    const int items = nblocks * njobs;
    for (int item = wave; item < items; item += waves) {
        const int blk = item / njobs, job = item - blk * njobs;
        const int bx = blk % bw, by = blk / bw;
        const float *plane = src + size_t(job ? ch1 : ch0) * size_t(w) * size_t(h);
        uint32_t vq[4];
        gather_channel(plane, w, h, bx, by, vq);
#pragma unroll
        for (int q = 0; q < 4; q++) vq[q] = __builtin_amdgcn_readfirstlane(vq[q]);
        const uint64_t r = optimal_channel(vq, lane);
        if (lane == 0)
            *reinterpret_cast<uint2 *>(out + size_t(blk) * stride + (job ? off1 : off0)) =
                make_uint2(uint32_t(r), uint32_t(r >> 32));
    }
}

// -------------------------------------------- OptimalCompress::compressDXT1G --

__device__ __forceinline__ int expand6(uint32_t g)
{
    g &= 63u;
    return int((g << 2) | (g >> 4));
}

// computeGreenError (OptimalCompressDXT.cpp:71-110): the four-colour palette
// of the expanded greens whatever their order, squared distances.  The
// endpoint fields are 6 bits wide, so a quantised green of 64 is stored as 0.
__device__ __forceinline__ uint32_t green_error(const uint32_t g[16], uint32_t g0, uint32_t g1)
{
    const int p0 = expand6(g0), p1 = expand6(g1);
    const int p2 = (2 * p0 + p1) / 3, p3 = (2 * p1 + p0) / 3;
    uint32_t total = 0;
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const int x = int(g[t]);
        const uint32_t d0 = uint32_t((x - p0) * (x - p0)), d1 = uint32_t((x - p1) * (x - p1));
        const uint32_t d2 = uint32_t((x - p2) * (x - p2)), d3 = uint32_t((x - p3) * (x - p3));
        total += min(min(d0, d1), min(d2, d3));
    }
    return total;
}

// OptimalCompress::compressDXT1G(const ColorBlock &) by one wavefront: a
// single green takes the OMatch6 endpoints (compressDXT1G(uint8),
// :294-311); otherwise the seed (max, min) of the greens quantised as
// (g + 1) >> 2, then every pair g0 > g1 of the range widened by 4, first
// strictly lower error; then BlockDXT1::evaluatePalette (BlockDXT.cpp:43-99)
// of the winner and computeGreenIndices (:112-145).  Red 31, blue 0.
__device__ uint2 optimal_green(const uint32_t g[16], const ntx::Tables *__restrict__ tab, int lane)
{
    uint32_t ming = 63, maxg = 0;
    bool single = true;
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const uint32_t q = (g[t] + 1) >> 2;
        ming = min(ming, q);
        maxg = max(maxg, q);
        single = single && g[t] == g[0];
    }
    if (single) {
        uint32_t c0 = 31u << 11 | uint32_t(tab->omatch6[g[0]] & 0xff) << 5;
        uint32_t c1 = 31u << 11 | uint32_t(tab->omatch6[g[0]] >> 8) << 5;
        uint32_t idx = 0xaaaaaaaau;
        if (c0 < c1) { const uint32_t t = c0; c0 = c1; c1 = t; idx ^= 0x55555555u; }
        return make_uint2(c0 | c1 << 16, idx);
    }
    const uint32_t seed = green_error(g, maxg, ming);
    const int lo = ming <= 4 ? 0 : int(ming) - 4, hi = maxg >= 59 ? 63 : int(maxg) + 4;
    const int n = hi - lo, total = n * (n + 1) / 2;
    uint32_t best = 0xffffffffu;
    int bk = 0x7fffffff;
    for (int base = 0; base < total; base += 64) {
        const int k = base + lane;
        if (k < total) {
            int m, j;
            tri_pair(k, m, j);
            const uint32_t e = green_error(g, uint32_t(lo + 1 + m), uint32_t(lo + j));
            if (e < best) { best = e; bk = k; }
        }
    }
    wave_argmin(best, bk);
    uint32_t b0 = maxg, b1 = ming;
    if (best < seed) {
        int m, j;
        tri_pair(bk, m, j);
        b0 = uint32_t(lo + 1 + m);
        b1 = uint32_t(lo + j);
    }
    const uint32_t c0 = 31u << 11 | (b0 & 63u) << 5, c1 = 31u << 11 | (b1 & 63u) << 5;
    const int p0 = expand6(b0), p1 = expand6(b1);
    int p2, p3;
    if (c0 > c1) {
        p2 = (2 * p0 + p1) / 3;
        p3 = (2 * p1 + p0) / 3;
    } else {
        p2 = (p0 + p1) / 2;
        p3 = 0;
    }
    uint32_t idx = 0;
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const int x = int(g[t]);
        const uint32_t d0 = uint32_t((p0 - x) * (p0 - x)), d1 = uint32_t((p1 - x) * (p1 - x));
        const uint32_t d2 = uint32_t((p2 - x) * (p2 - x)), d3 = uint32_t((p3 - x) * (p3 - x));
        const uint32_t x0 = uint32_t(d1 > d2) & uint32_t(d0 > d2), x1 = uint32_t(d0 > d3) & uint32_t(d1 > d3);
        const uint32_t x2 = uint32_t(d0 > d3) & uint32_t(d2 > d3);
        idx |= (x2 | ((x0 | x1) << 1)) << (2 * t);
    }
    return make_uint2(c0 | c1 << 16, idx);
}

// One wavefront per block: the colour half (bytes 8..15) of BC3n blocks.
__global__ __launch_bounds__(256) void optimal_green_kernel(const float *__restrict__ src, int w, int h, int bw,
                                                            int nblocks, const ntx::Tables *__restrict__ tab,
                                                            uint8_t *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(int((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int waves = int((gridDim.x * blockDim.x) >> 6);
    const float *plane = src + size_t(w) * size_t(h);
    for (int blk = wave; blk < nblocks; blk += waves) {
        uint32_t vq[4];
        gather_channel(plane, w, h, blk % bw, blk / bw, vq);
        uint32_t g[16];
#pragma unroll
        for (int t = 0; t < 16; t++) g[t] = (__builtin_amdgcn_readfirstlane(vq[t >> 2]) >> (8 * (t & 3))) & 0xffu;
        const uint2 r = optimal_green(g, tab, lane);
        if (lane == 0) *reinterpret_cast<uint2 *>(out + size_t(blk) * 16 + 8) = r;
    }
}

// ------------------------------------------------ QuickCompress::compressDXT1 --
// In the reference's float operation order (nvmath Vector3: lerp v1 * s +
// t * v2, dot x + y + z, division by a scalar per component); nv::min / max
// as in Utils.h.

struct F3 {
    float x, y, z;
};

__device__ __forceinline__ float q_clamp(float x, float a, float b) { return nvminf(nvmaxf(x, a), b); }

__device__ __forceinline__ float q_dist(F3 a, F3 b)
{
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return dx * dx + dy * dy + dz * dz;
}

// roundAndExpand (QuickCompressDXT.cpp:122-150).
__device__ __forceinline__ uint32_t q_round_expand(F3 &v)
{
    uint32_t r = uint32_t(floorf(q_clamp(v.x * (31.0f / 255.0f), 0.0f, 31.0f)));
    uint32_t g = uint32_t(floorf(q_clamp(v.y * (63.0f / 255.0f), 0.0f, 63.0f)));
    uint32_t b = uint32_t(floorf(q_clamp(v.z * (31.0f / 255.0f), 0.0f, 31.0f)));
    if (fabsf(v.x - float(((r + 1) << 3) | ((r + 1) >> 2))) < fabsf(v.x - float((r << 3) | (r >> 2)))) r = min(r + 1, 31u);
    if (fabsf(v.y - float(((g + 1) << 2) | ((g + 1) >> 4))) < fabsf(v.y - float((g << 2) | (g >> 4)))) g = min(g + 1, 63u);
    if (fabsf(v.z - float(((b + 1) << 3) | ((b + 1) >> 2))) < fabsf(v.z - float((b << 3) | (b >> 2)))) b = min(b + 1, 31u);
    v.x = float((r << 3) | (r >> 2));
    v.y = float((g << 2) | (g >> 4));
    v.z = float((b << 3) | (b >> 2));
    return r << 11 | g << 5 | b;
}

// computeIndices4 (QuickCompressDXT.cpp:195-229).
__device__ __forceinline__ uint32_t q_indices4(const F3 blk[16], F3 mx, F3 mn)
{
    const float t1 = 1.0f / 3.0f, s1 = 1.0f - t1, t2 = 2.0f / 3.0f, s2 = 1.0f - t2;
    const F3 p2 = { mx.x * s1 + t1 * mn.x, mx.y * s1 + t1 * mn.y, mx.z * s1 + t1 * mn.z };
    const F3 p3 = { mx.x * s2 + t2 * mn.x, mx.y * s2 + t2 * mn.y, mx.z * s2 + t2 * mn.z };
    uint32_t idx = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float d0 = q_dist(mx, blk[i]), d1 = q_dist(mn, blk[i]), d2 = q_dist(p2, blk[i]), d3 = q_dist(p3, blk[i]);
        const uint32_t b0 = d0 > d3, b1 = d1 > d2, b2 = d0 > d2, b3 = d1 > d3, b4 = d2 > d3;
        const uint32_t x0 = b1 & b2, x1 = b0 & b3, x2 = b0 & b4;
        idx |= (x2 | ((x0 | x1) << 1)) << (2 * i);
    }
    return idx;
}

// QuickCompress::compressDXT1 (QuickCompressDXT.cpp:653-713): single RGB ->
// OptimalCompress::compressDXT1; otherwise the colour box with the diagonal
// picked by covariance signs (findMinMaxColorsBox, selectDiagonal,
// insetBBox, :72-117), rounded and expanded, four-colour indices, then one
// least-squares step (optimizeEndPoints4, :406-452).  rgb: R | G << 8 | B << 16.
__device__ uint2 quick_dxt1(const uint32_t rgb[16], const ntx::Tables *__restrict__ tab)
{
    bool single = true;
#pragma unroll
    for (int i = 1; i < 16; i++) single = single && rgb[i] == rgb[0];
    if (single) {
        const uint32_t r8 = rgb[0] & 0xff, g8 = (rgb[0] >> 8) & 0xff, b8 = rgb[0] >> 16;
        uint32_t c0 = uint32_t(tab->omatch5[r8] & 0xff) << 11 | uint32_t(tab->omatch6[g8] & 0xff) << 5 |
                      uint32_t(tab->omatch5[b8] & 0xff);
        uint32_t c1 = uint32_t(tab->omatch5[r8] >> 8) << 11 | uint32_t(tab->omatch6[g8] >> 8) << 5 |
                      uint32_t(tab->omatch5[b8] >> 8);
        uint32_t idx = 0xaaaaaaaau;
        if (c0 < c1) { const uint32_t t = c0; c0 = c1; c1 = t; idx ^= 0x55555555u; }
        return make_uint2(c0 | c1 << 16, idx);
    }
    F3 blk[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
        blk[i] = F3{ float(rgb[i] & 0xff), float((rgb[i] >> 8) & 0xff), float(rgb[i] >> 16) };
    F3 mx = { 0.0f, 0.0f, 0.0f }, mn = { 255.0f, 255.0f, 255.0f };
#pragma unroll
    for (int i = 0; i < 16; i++) {
        mx.x = nvmaxf(mx.x, blk[i].x); mx.y = nvmaxf(mx.y, blk[i].y); mx.z = nvmaxf(mx.z, blk[i].z);
        mn.x = nvminf(mn.x, blk[i].x); mn.y = nvminf(mn.y, blk[i].y); mn.z = nvminf(mn.z, blk[i].z);
    }
    {
        const float cx = (mx.x + mn.x) * 0.5f, cy = (mx.y + mn.y) * 0.5f, cz = (mx.z + mn.z) * 0.5f;
        float covx = 0.0f, covy = 0.0f;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const float tz = blk[i].z - cz;
            covx += (blk[i].x - cx) * tz;
            covy += (blk[i].y - cy) * tz;
        }
        if (covx < 0) { const float t = mx.x; mx.x = mn.x; mn.x = t; }
        if (covy < 0) { const float t = mx.y; mx.y = mn.y; mn.y = t; }
        const float k = (8.0f / 255.0f) / 16.0f;
        const F3 in = { (mx.x - mn.x) / 16.0f - k, (mx.y - mn.y) / 16.0f - k, (mx.z - mn.z) / 16.0f - k };
        const F3 a = { q_clamp(mx.x - in.x, 0.0f, 255.0f), q_clamp(mx.y - in.y, 0.0f, 255.0f),
                       q_clamp(mx.z - in.z, 0.0f, 255.0f) };
        const F3 b = { q_clamp(mn.x + in.x, 0.0f, 255.0f), q_clamp(mn.y + in.y, 0.0f, 255.0f),
                       q_clamp(mn.z + in.z, 0.0f, 255.0f) };
        mx = a;
        mn = b;
    }
    uint32_t k0 = q_round_expand(mx), k1 = q_round_expand(mn);
    if (k0 < k1) {
        const F3 t = mx; mx = mn; mn = t;
        const uint32_t u = k0; k0 = k1; k1 = u;
    }
    uint32_t idx = q_indices4(blk, mx, mn);

    // optimizeEndPoints4: least squares on these indices; a near-singular
    // system (nv::equal(denom, 0), nvmath.h:139-143) keeps the block.
    float a2 = 0.0f, b2 = 0.0f, ab = 0.0f;
    F3 ax = { 0.0f, 0.0f, 0.0f }, bx = { 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t bits = idx >> (2 * i);
        float beta = float(bits & 1);
        if (bits & 2) beta = (1 + beta) / 3.0f;
        const float alpha = 1.0f - beta;
        a2 += alpha * alpha;
        b2 += beta * beta;
        ab += alpha * beta;
        ax.x += blk[i].x * alpha; ax.y += blk[i].y * alpha; ax.z += blk[i].z * alpha;
        bx.x += blk[i].x * beta; bx.y += blk[i].y * beta; bx.z += blk[i].z * beta;
    }
    const float denom = a2 * b2 - ab * ab;
    if (fabsf(denom) <= 0.0001f * nvmaxf(1.0f, nvmaxf(fabsf(denom), 0.0f)))
        return make_uint2(k0 | k1 << 16, idx);
    const float f = 1.0f / denom;
    F3 a = { (ax.x * b2 - bx.x * ab) * f, (ax.y * b2 - bx.y * ab) * f, (ax.z * b2 - bx.z * ab) * f };
    F3 b = { (bx.x * a2 - ax.x * ab) * f, (bx.y * a2 - ax.y * ab) * f, (bx.z * a2 - ax.z * ab) * f };
    a = F3{ q_clamp(a.x, 0.0f, 255.0f), q_clamp(a.y, 0.0f, 255.0f), q_clamp(a.z, 0.0f, 255.0f) };
    b = F3{ q_clamp(b.x, 0.0f, 255.0f), q_clamp(b.y, 0.0f, 255.0f), q_clamp(b.z, 0.0f, 255.0f) };
    k0 = q_round_expand(a);
    k1 = q_round_expand(b);
    if (k0 < k1) {
        const F3 t = a; a = b; b = t;
        const uint32_t u = k0; k0 = k1; k1 = u;
    }
    return make_uint2(k0 | k1 << 16, q_indices4(blk, a, b));
}

// QuickCompress::compressDXT5 (QuickCompressDXT.cpp:823-827), one thread
// per block: QuickCompress::compressDXT5A of the alpha (kN: of the red,
// after swizzle(4, 1, 5, 0), CompressorDX9.cpp:74-80) and compressDXT1 of
// the colour (kN: (255, G, 0)).
template <bool kN>
__global__ __launch_bounds__(256) void bc3_fast_kernel(const float *__restrict__ src, int w, int h, int bw, int nblocks,
                                                       const ntx::Tables *__restrict__ tab, uint8_t *__restrict__ out)
{
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nblocks) return;
    const int bx = blk % bw, by = blk / bw;
    const size_t plane = size_t(w) * size_t(h);
    uint32_t rgb[16], a4[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t word = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const size_t o = block_texel(w, h, bx, by, 4 * q + b);
            const uint32_t r = to_u8(src[o]), g = to_u8(src[o + plane]);
            if (kN) {
                rgb[4 * q + b] = 0xffu | g << 8;
                word |= r << (8 * b);
            } else {
                rgb[4 * q + b] = r | g << 8 | to_u8(src[o + 2 * plane]) << 16;
                word |= to_u8(src[o + 3 * plane]) << (8 * b);
            }
        }
        a4[q] = word;
    }
    const uint64_t alpha = compress_channel(a4);
    const uint2 colour = quick_dxt1(rgb, tab);
    uint2 *dst = reinterpret_cast<uint2 *>(out + size_t(blk) * 16);
    dst[0] = make_uint2(uint32_t(alpha), uint32_t(alpha >> 32));
    dst[1] = colour;
}

}  // namespace

namespace ntx {

// Wavefront-per-block launches: 4 wavefronts per workgroup, a grid-stride
// loop over the blocks, at most `groups` workgroups.
static int wave_grid(int items, int groups)
{
    return std::max(1, std::min((items + 3) / 4, groups));
}

// OptimalCompress::compressDXT5A of the channels `format` encodes with it:
// BC4 red; BC5 red then green; BC3 alpha; BC3n red (the alpha block).
hipError_t launch_optimal_channels(const float *d_src, int w, int h, int format, void *d_out, int groups,
                                   hipStream_t stream)
{
    const int bw = (w + 3) / 4, nblocks = bw * ((h + 3) / 4);
    if (nblocks == 0) return hipSuccess;
    int njobs = 1, ch0 = 0, ch1 = 0, stride = 16;
    switch (format) {
    case NTX_FORMAT_BC4: stride = 8; break;
    case NTX_FORMAT_BC5: njobs = 2; ch1 = 1; break;
    case NTX_FORMAT_BC3: ch0 = 3; break;
    case NTX_FORMAT_BC3N: break;
    default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(optimal_channel_kernel, dim3(wave_grid(nblocks * njobs, groups)), dim3(256), 0, stream, d_src, w,
                       h, bw, nblocks, njobs, ch0, ch1, stride, 0, 8, static_cast<uint8_t *>(d_out));
    return hipGetLastError();
}

// OptimalCompress::compressDXT1G into the colour half of BC3n blocks.
hipError_t launch_optimal_green(const float *d_src, int w, int h, const Tables *d_tab, void *d_out, int groups,
                                hipStream_t stream)
{
    const int bw = (w + 3) / 4, nblocks = bw * ((h + 3) / 4);
    if (nblocks == 0) return hipSuccess;
    hipLaunchKernelGGL(optimal_green_kernel, dim3(wave_grid(nblocks, groups)), dim3(256), 0, stream, d_src, w, h, bw,
                       nblocks, d_tab, static_cast<uint8_t *>(d_out));
    return hipGetLastError();
}

// QuickCompress::compressDXT5 of whole BC3 (normal = false) or BC3n blocks.
hipError_t launch_bc3_fast(const float *d_src, int w, int h, bool normal, const Tables *d_tab, void *d_out,
                           hipStream_t stream)
{
    const int bw = (w + 3) / 4, nblocks = bw * ((h + 3) / 4);
    if (nblocks == 0) return hipSuccess;
    const dim3 grid((nblocks + 255) / 256);
    if (normal)
        hipLaunchKernelGGL(bc3_fast_kernel<true>, grid, dim3(256), 0, stream, d_src, w, h, bw, nblocks, d_tab,
                           static_cast<uint8_t *>(d_out));
    else
        hipLaunchKernelGGL(bc3_fast_kernel<false>, grid, dim3(256), 0, stream, d_src, w, h, bw, nblocks, d_tab,
                           static_cast<uint8_t *>(d_out));
    return hipGetLastError();
}

}  // namespace ntx
