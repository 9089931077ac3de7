// bc345.hip -- BC3 / BC4 / BC5 encoders for gfx950, bit-exact with NVTT's
// CPU path at Quality_Normal:
//   alpha / BC4 channel: QuickCompress::compressDXT5A (QuickCompressDXT.cpp:772-821),
//   BC3 colour: nvsquish WeightedClusterFit scalar build (weightedclusterfit.cpp:39-106, 476-589)
//     or OptimalCompress::compressDXT1 for single-colour blocks (CompressorDX9.cpp:158-176).
// Input is gathered as ColorBlock::init does (ColorBlock.cpp:80-110):
// truncating float->u8 conversion, edge blocks filled by repeating pixels.
#include "ntx_common.h"

#include <cfloat>

namespace {

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

// computeAlphaIndices (QuickCompressDXT.cpp:505-536): nearest palette entry,
// first on ties; rewrites the 48 index bits and returns the squared error.
__device__ __forceinline__ uint32_t alpha_indices(const uint32_t v[16], uint64_t &blk)
{
    uint32_t pal[8];
    alpha_palette(uint32_t(blk & 0xff), uint32_t((blk >> 8) & 0xff), pal);
    uint32_t total = 0;
    uint64_t bits = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        uint32_t be = 256 * 256, bp = 8;
#pragma unroll
        for (uint32_t p = 0; p < 8; p++) {
            const int d = int(pal[p]) - int(v[i]);
            const uint32_t e = uint32_t(d * d);
            if (e < be) { be = e; bp = p; }
        }
        total += be;
        bits |= uint64_t(bp) << (16 + 3 * i);
    }
    blk = (blk & 0xffffull) | bits;
    return total;
}

// optimizeAlpha8 (QuickCompressDXT.cpp:538-592): least-squares endpoints for
// the current indices; truncating conversion, NaN -> 0 through nv::max.
__device__ __forceinline__ void optimize_alpha8(const uint32_t v[16], uint64_t &blk)
{
    float a2 = 0, b2 = 0, ab = 0, ax = 0, bx = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t idx = uint32_t(blk >> (16 + 3 * i)) & 7u;
        const float alpha = (idx < 2) ? 1.0f - float(idx) : (8.0f - float(idx)) / 7.0f;
        const float beta = 1 - alpha;
        a2 += alpha * alpha;
        b2 += beta * beta;
        ab += alpha * beta;
        ax += alpha * float(v[i]);
        bx += beta * float(v[i]);
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

__device__ uint64_t compress_channel(const uint32_t v[16])
{
    uint32_t mx = 0, mn = 255;
#pragma unroll
    for (int i = 0; i < 16; i++) { mx = v[i] > mx ? v[i] : mx; mn = v[i] < mn ? v[i] : mn; }
    const uint32_t d = (mx - mn) / 34;
    uint64_t blk = uint64_t((mx - d) & 0xff) | uint64_t((mn + d) & 0xff) << 8;
    uint32_t besterr = alpha_indices(v, blk);
    uint64_t best = blk;
    for (int it = 0; it < 8; it++) {
        optimize_alpha8(v, blk);
        const uint32_t err = alpha_indices(v, blk);
        if (err >= besterr) break;
        // sameIndices (QuickCompressDXT.cpp:643-647) ORs ~0xFFFF into both
        // words, so only the endpoint bytes are compared.
        if ((blk & 0xffffull) == (best & 0xffffull)) { best = blk; break; }
        besterr = err;
        best = blk;
    }
    return best;
}

// One thread per (block, channel job).  Channel plane: 0 = R, 1 = G, 3 = A.
__global__ __launch_bounds__(256) void channel_kernel(const float *__restrict__ src, int w, int h, int bw, int nblocks,
                                                     int njobs, int ch0, int ch1, int stride, int off1,
                                                     uint8_t *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblocks * njobs) return;
    const int blk = t / njobs, job = t - blk * njobs;
    const int ch = job == 0 ? ch0 : ch1;
    const int bx = blk % bw, by = blk / bw;
    const float *plane = src + size_t(ch) * size_t(w) * size_t(h);
    uint32_t v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = to_u8(plane[block_texel(w, h, bx, by, i)]);
    const uint64_t r = compress_channel(v);
    uint2 *dst = reinterpret_cast<uint2 *>(out + size_t(blk) * stride + (job == 0 ? 0 : off1));
    *dst = make_uint2(uint32_t(r), uint32_t(r >> 32));
}

// ---------------------------------------------------------- BC3 colour ----

constexpr int kGroup = 16;
constexpr int kBlocksPerWG = 16;
constexpr int kTri = 153;  // partial-sum rows s=0..16, lengths 0..16-s

__device__ __forceinline__ int tri_index(int s, int len) { return s * 17 - (s * (s - 1)) / 2 + len; }

struct SqShared {
    float4 part[kBlocksPerWG][kTri];  // left-to-right partial sums of weighted points (x,y,z,w)
    float4 wp[kBlocksPerWG][16];      // weighted points in projection order
    float dps[kBlocksPerWG][16];
    int order[kBlocksPerWG][16];
};

__device__ __forceinline__ float gshfl(float v, int s) { return __shfl(v, s, kGroup); }
__device__ __forceinline__ int gshfl(int v, int s) { return __shfl(v, s, kGroup); }
__device__ __forceinline__ uint32_t gballot(bool p)
{
    const unsigned long long m = __ballot(p);
    return uint32_t(m >> (threadIdx.x & 48)) & 0xffffu;
}

// std::min / std::max as used by squish's Vec3 Min/Max (maths.h).
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

struct SqCand {
    float ax, ay, az, bx, by, bz;
};

// One candidate of the scalar Compress4 loop body (weightedclusterfit.cpp:507-541).
__device__ __forceinline__ float sq_candidate(const float4 *part, uint32_t pk, float4 xsum, float mx, float my, float mz,
                                              SqCand *out)
{
    const int c0 = int(pk & 0xff), c1 = int((pk >> 8) & 0xff), c2 = int((pk >> 16) & 0xff);
    const float4 x0 = part[tri_index(0, c0)];
    const float4 x1 = part[tri_index(c0, c1)];
    const float4 x2 = part[tri_index(c0 + c1, c2)];
    const float w3 = ((xsum.w - x0.w) - x1.w) - x2.w;
    const float alpha2 = (x0.w + x1.w * (4.0f / 9.0f)) + x2.w * (1.0f / 9.0f);
    const float beta2 = (w3 + x2.w * (4.0f / 9.0f)) + x1.w * (1.0f / 9.0f);
    const float ab = (x1.w + x2.w) * (2.0f / 9.0f);
    const float factor = 1.0f / (alpha2 * beta2 - ab * ab);
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
    out->ax = ar; out->ay = ag; out->az = abl;
    out->bx = br; out->by = bg; out->bz = bb;
    return e1r * mx + e1g * my + e1b * mz;
}

// colourblock.cpp:30-54 FloatToInt / FloatTo565.
__device__ __forceinline__ int float_to_int(float a, int limit)
{
    int i = int(a + 0.5f);
    return i < 0 ? 0 : (i > limit ? limit : i);
}

__global__ __launch_bounds__(256) void bc3_color_kernel(const float *__restrict__ src, int w, int h, int bw, int nblocks,
                                                        int weight_by_alpha, float cwx, float cwy, float cwz,
                                                        const ntx::Tables *__restrict__ tab, int stride, int offset,
                                                        uint8_t *__restrict__ out)
{
    __shared__ SqShared sh;
    const int g = threadIdx.x >> 4;
    const int lane = threadIdx.x & 15;
    const int blk = blockIdx.x * kBlocksPerWG + g;
    if (blk >= nblocks) return;
    const int bx = blk % bw, by = blk / bw;
    const size_t plane = size_t(w) * size_t(h);
    const size_t o = block_texel(w, h, bx, by, lane);
    const uint32_t r8 = to_u8(src[o]), g8 = to_u8(src[o + plane]), b8 = to_u8(src[o + 2 * plane]),
                   a8 = to_u8(src[o + 3 * plane]);
    uint2 *dst = reinterpret_cast<uint2 *>(out + size_t(blk) * stride + offset);

    // Single RGB colour (ColorBlock::isSingleColor, ColorBlock.cpp:136-149).
    const uint32_t rgb = r8 | g8 << 8 | b8 << 16;
    const uint32_t rgb0 = uint32_t(gshfl(int(rgb), 0));
    if (gballot(rgb != rgb0) == 0) {
        if (lane == 0) {
            // OptimalCompress::compressDXT1 (OptimalCompressDXT.cpp:254-269).
            uint32_t c0 = uint32_t(tab->omatch5[r8] & 0xff) << 11 | uint32_t(tab->omatch6[g8] & 0xff) << 5 |
                          uint32_t(tab->omatch5[b8] & 0xff);
            uint32_t c1 = uint32_t(tab->omatch5[r8] >> 8) << 11 | uint32_t(tab->omatch6[g8] >> 8) << 5 |
                          uint32_t(tab->omatch5[b8] >> 8);
            uint32_t idx = 0xaaaaaaaau;
            if (c0 < c1) { const uint32_t t = c0; c0 = c1; c1 = t; idx ^= 0x55555555u; }
            *dst = make_uint2(c0 | c1 << 16, idx);
        }
        return;
    }

    // ColourSet with createMinimalSet (colourset.cpp:35-101): texels sharing
    // RGB bytes map to the point of their first occurrence.
    int first = lane;
    for (int j = 15; j >= 0; j--) {
        const uint32_t o2 = uint32_t(gshfl(int(rgb), j));
        if (j < lane && o2 == rgb) first = j;
    }
    const bool distinct = first == lane;
    const uint32_t dmask = gballot(distinct);
    const int count = __builtin_popcount(dmask);
    // Point id of every texel: rank of its first occurrence among distinct texels.
    const int my_point = __builtin_popcount(dmask & ((1u << first) - 1u));
    const float tw = weight_by_alpha ? float(a8 + 1) / 256.0f : 1.0f;
    // Lane p becomes point p: colour of the p-th distinct texel, weights summed in texel order.
    int src_lane = 0;
    {
        uint32_t mm = dmask;
        for (int d = 0; d < lane && mm; d++) mm &= mm - 1u;
        src_lane = mm ? __builtin_ctz(mm) : 0;
    }
    const bool live = lane < count;
    const float px = float(gshfl(int(r8), src_lane)) / 255.0f;
    const float py = float(gshfl(int(g8), src_lane)) / 255.0f;
    const float pz = float(gshfl(int(b8), src_lane)) / 255.0f;
    float pw = 0.0f;
    for (int i = 0; i < 16; i++) {
        const int pi = gshfl(my_point, i);
        const float wi = gshfl(tw, i);
        if (pi == lane) pw += wi;
    }

    // ComputeWeightedCovariance (maths.cpp:32-59), metric on the deltas.
    float total = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    const float qx = px * pw, qy = py * pw, qz = pz * pw;
    for (int j = 0; j < count; j++) {
        total += gshfl(pw, j);
        cx += gshfl(qx, j); cy += gshfl(qy, j); cz += gshfl(qz, j);
    }
    const float tinv = 1.0f / total;
    cx *= tinv; cy *= tinv; cz *= tinv;
    const float ax = (px - cx) * cwx, ay = (py - cy) * cwy, az = (pz - cz) * cwz;
    const float bxw = ax * pw, byw = ay * pw, bzw = az * pw;
    const float k0 = ax * bxw, k1 = ax * byw, k2 = ax * bzw, k3 = ay * byw, k4 = ay * bzw, k5 = az * bzw;
    float m0 = 0, m1 = 0, m2 = 0, m3 = 0, m4 = 0, m5 = 0;
    for (int j = 0; j < count; j++) {
        m0 += gshfl(k0, j); m1 += gshfl(k1, j); m2 += gshfl(k2, j);
        m3 += gshfl(k3, j); m4 += gshfl(k4, j); m5 += gshfl(k5, j);
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
    // Stable insertion sort by projection (weightedclusterfit.cpp:62-77).
    const float dp = px * vx + py * vy + pz * vz;
    const bool has_nan = gballot(live && dp != dp) != 0;
    int rank = lane;
    if (!has_nan) {
        rank = 0;
        for (int j = 0; j < count; j++) {
            const float o2 = gshfl(dp, j);
            rank += (o2 < dp || (o2 == dp && j < lane)) ? 1 : 0;
        }
        if (live) sh.order[g][rank] = lane;
        ntx_wave_sync();
    } else {
        if (live) { sh.dps[g][lane] = dp; sh.order[g][lane] = lane; }
        ntx_wave_sync();
        if (lane == 0) {
            for (int i = 0; i < count; i++)
                for (int j = i; j > 0 && sh.dps[g][j] < sh.dps[g][j - 1]; j--) {
                    const float t = sh.dps[g][j]; sh.dps[g][j] = sh.dps[g][j - 1]; sh.dps[g][j - 1] = t;
                    const int q = sh.order[g][j]; sh.order[g][j] = sh.order[g][j - 1]; sh.order[g][j - 1] = q;
                }
        }
        ntx_wave_sync();
        for (int i = 0; i < count; i++)
            if (sh.order[g][i] == lane) rank = i;
    }
    // Weighted points in sorted order (weightedclusterfit.cpp:92-104).
    const int sp = live ? sh.order[g][lane] : 0;
    const float sw = gshfl(pw, sp);
    const float wx = sw * gshfl(px, sp), wy = sw * gshfl(py, sp), wz = sw * gshfl(pz, sp);
    if (live) sh.wp[g][lane] = make_float4(wx, wy, wz, sw);
    float4 xsum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int j = 0; j < count; j++) {
        xsum.x += gshfl(wx, j); xsum.y += gshfl(wy, j); xsum.z += gshfl(wz, j); xsum.w += gshfl(sw, j);
    }
    ntx_wave_sync();
    // Partial sums: lane s writes row s, T[s][len] = sum of wp[s .. s+len-1]
    // accumulated left to right from zero, as the nested loops do.
    if (lane <= count) {
        float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        sh.part[g][tri_index(lane, 0)] = acc;
        for (int len = 1; lane + len <= count; len++) {
            const float4 v = sh.wp[g][lane + len - 1];
            acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
            sh.part[g][tri_index(lane, len)] = acc;
        }
    }
    if (lane == 0 && count == 16) sh.part[g][tri_index(16, 0)] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    ntx_wave_sync();

    // Compress4 search: candidates in loop order q; the first strict minimum wins.
    const float msx = cwx * cwx, msy = cwy * cwy, msz = cwz * cwz;
    const int qbeg = tab->sq_off[count], qn = tab->sq_off[count + 1] - qbeg;
    float be = FLT_MAX;
    int bq = 0;
    SqCand c;
    for (int q = lane; q < qn; q += kGroup) {
        const float e = sq_candidate(sh.part[g], tab->sq[qbeg + q], xsum, msx, msy, msz, &c);
        if (e < be) { be = e; bq = q; }
    }
    for (int off = 8; off >= 1; off >>= 1) {
        const float oe = gshfl(be, lane ^ off);
        const int oq = gshfl(bq, lane ^ off);
        if (oe < be || (oe == be && oq < bq)) { be = oe; bq = oq; }
    }
    if (!(be < FLT_MAX)) {
        // The reference never writes the block in this case; emit zeros.
        if (lane == 0) *dst = make_uint2(0u, 0u);
        return;
    }
    const uint32_t pk = tab->sq[qbeg + bq];
    sq_candidate(sh.part[g], pk, xsum, msx, msy, msz, &c);
    const int b0 = int(pk & 0xff), b1 = int((pk >> 8) & 0xff), b2 = int((pk >> 16) & 0xff);
    // Index of this texel: by the sorted position of its point.
    const int pos = gshfl(rank, my_point);
    uint32_t idx = pos < b0 ? 0u : pos < b0 + b1 ? 2u : pos < b0 + b1 + b2 ? 3u : 1u;
    // WriteColourBlock4 (colourblock.cpp:108-138).
    int ca = (float_to_int(31.0f * c.ax, 31) << 11) | (float_to_int(63.0f * c.ay, 63) << 5) | float_to_int(31.0f * c.az, 31);
    int cb = (float_to_int(31.0f * c.bx, 31) << 11) | (float_to_int(63.0f * c.by, 63) << 5) | float_to_int(31.0f * c.bz, 31);
    if (ca < cb) { const int t = ca; ca = cb; cb = t; idx = (idx ^ 1u) & 3u; }
    else if (ca == cb) { idx = 0u; }
    uint32_t bits = idx << (2 * lane);
    for (int off = 8; off >= 1; off >>= 1) bits |= uint32_t(gshfl(int(bits), lane ^ off));
    if (lane == 0) *dst = make_uint2(uint32_t(ca) | uint32_t(cb) << 16, bits);
}

}  // namespace

namespace ntx {

// BC4 (R), BC5 (R then G) or the BC3 alpha half (A), one thread per channel block.
hipError_t launch_channels(const float *d_src, int w, int h, int format, void *d_out, hipStream_t stream)
{
    const int bw = (w + 3) / 4, bh = (h + 3) / 4, nblocks = bw * bh;
    if (nblocks == 0) return hipSuccess;
    int njobs = 1, ch0 = 0, ch1 = 0, stride = 8, off1 = 0;
    if (format == NTX_FORMAT_BC5) { njobs = 2; ch0 = 0; ch1 = 1; stride = 16; off1 = 8; }
    else if (format == NTX_FORMAT_BC3) { ch0 = 3; stride = 16; }
    const int threads = nblocks * njobs;
    hipLaunchKernelGGL(channel_kernel, dim3((threads + 255) / 256), dim3(256), 0, stream, d_src, w, h, bw, nblocks,
                       njobs, ch0, ch1, stride, off1, reinterpret_cast<uint8_t *>(d_out));
    return hipGetLastError();
}

hipError_t launch_bc3_color(const float *d_src, int w, int h, int weight_by_alpha, const float *cw, const Tables *d_tab,
                            void *d_out, hipStream_t stream)
{
    const int bw = (w + 3) / 4, bh = (h + 3) / 4, nblocks = bw * bh;
    if (nblocks == 0) return hipSuccess;
    hipLaunchKernelGGL(bc3_color_kernel, dim3((nblocks + kBlocksPerWG - 1) / kBlocksPerWG), dim3(kBlocksPerWG * kGroup), 0,
                       stream, d_src, w, h, bw, nblocks, weight_by_alpha, cw[0], cw[1], cw[2], d_tab, 16, 8,
                       reinterpret_cast<uint8_t *>(d_out));
    return hipGetLastError();
}

}  // namespace ntx
