// bc1_block.hpp -- device code of the BC1 encoder shared by the per-image
// kernel (bc1.hip) and the batch kernel (batch.hip): icbc v1.05
// compress_dxt1 at Quality_Level8 (and Level9: the same plus refine_endpoints;
// Level1: box fit + least squares), the SSE4.1 build (no FMA,
// round-half-even, 4-lane tie-break), icbc.h:3485-3559.
//
// One 4x4 block per 16-lane group.  Lane l of a group owns texel l (gather,
// index selection, error terms) and, after colour reduction, cluster l.  The
// partition search spreads partition p over lane p % 16; the SSE tie-break
// (per-lane strict '<' over p = lane + 4k, then lowest lane, icbc.h:533-545)
// is reproduced by ordering candidates by (error, p & 3, p >> 2).
#pragma once

#include "ntx_common.h"

#include <cfloat>
#include <type_traits>

namespace ntx_bc1 {

constexpr int kGroup = 16;
constexpr int kBlocksPerWG = 16;

struct Bc1Shared {
    // Row 0 is zero; row i + 1 holds the (r, g, b, w) prefix sums over the
    // first i + 1 clusters in projection order, so a cluster-table entry
    // (which stores cluster index + 1, 0 for none) indexes it directly.
    float4 sat[kBlocksPerWG][17];
    float4 col[kBlocksPerWG][16];    // scratch: sorted products for the SAT scan
    float dps[kBlocksPerWG][16];
    int order[kBlocksPerWG][16];
    uint32_t four[976];
    uint32_t three[160];
    float div255[256];
};

// The workgroup's copy of the cluster tables and the k / 255 table.
__device__ __forceinline__ void load_shared_tables(Bc1Shared &sh, const ntx::Tables *tab, int t, int nthreads)
{
    for (int i = t; i < 976; i += nthreads) sh.four[i] = tab->four[i];
    for (int i = t; i < 160; i += nthreads) sh.three[i] = tab->three[i];
    for (int i = t; i < 256; i += nthreads) sh.div255[i] = tab->div255[i];
}

__device__ __forceinline__ float sse_sat(float a)
{
    // minps(maxps(a, 0), 1): NaN -> 0 (icbc.h:474-478).
    float t = (a > 0.0f) ? a : 0.0f;
    return (t < 1.0f) ? t : 1.0f;
}
__device__ __forceinline__ float round5(float x) { return rintf(x * 31.0f) * (1.0f / 31.0f); }
__device__ __forceinline__ float round6(float x) { return rintf(x * 63.0f) * (1.0f / 63.0f); }

__device__ __forceinline__ float gshfl(float v, int src) { return __shfl(v, src, kGroup); }
__device__ __forceinline__ int gshfl(int v, int src) { return __shfl(v, src, kGroup); }

// Lane I of this group's 16-lane row, read by every lane of the row (DPP
// row_newbcast, one VALU op).  Every branch around these reads is uniform
// over a group, so the source lane is always active.
template <int I>
__device__ __forceinline__ float gbcast(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + I, 0xf, 0xf, false));
}
// The value of the lane R places further on in the row, cyclically (DPP row_ror).
template <int R>
__device__ __forceinline__ int gror(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x120 + R, 0xf, 0xf, false);
}
template <int R>
__device__ __forceinline__ float gror(float v)
{
    return __int_as_float(gror<R>(__float_as_int(v)));
}

// f(std::integral_constant<int, i>) for i = I .. N-1, in order.
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// Mask of the lanes of this group for which pred holds, bit i = lane i.
__device__ __forceinline__ uint32_t gballot(bool pred)
{
    const unsigned long long m = __ballot(pred);
    const int base = threadIdx.x & 48;  // group offset inside the wave
    return uint32_t(m >> base) & 0xffffu;
}

struct Cand {
    float ax, ay, az, bx, by, bz;
};

// The cluster-weight terms of one least-squares candidate (alpha2, beta2,
// alphabeta, 1 / (alpha2 beta2 - alphabeta^2)), icbc.h:2501-2511 (four
// colours: x0, x1, x2 = prefix sums at c0 <= c1 <= c2) or :2201-2208 (three
// colours: x0, x1 at c0 <= c1; x2 unused).  Only the .w fields are read.  The
// host builds the same terms for integer weights (ntx_tables.cpp) with these
// operations.
template <bool FOUR>
__device__ __forceinline__ float4 cluster_terms(float4 x0, float4 x1, float4 x2, float sumw)
{
    float alpha2, beta2, ab;
    if (FOUR) {
        const float w3 = sumw - x2.w;
        const float w2 = x2.w - x1.w;
        const float w1 = x1.w - x0.w;
        const float w0 = x0.w;
        alpha2 = w2 * (1.0f / 9.0f) + (w1 * (4.0f / 9.0f) + w0);
        beta2 = w1 * (1.0f / 9.0f) + (w2 * (4.0f / 9.0f) + w3);
        ab = (w1 + w2) * (2.0f / 9.0f);
    } else {
        const float w2 = sumw - x1.w;
        const float w1 = x1.w - x0.w;
        ab = w1 * 0.25f;
        alpha2 = x0.w + ab;
        beta2 = w2 + ab;
    }
    return make_float4(alpha2, beta2, ab, 1.0f / (alpha2 * beta2 - ab * ab));
}

// The rest of the candidate (icbc.h:2513-2538 / :2210-2236): endpoints,
// snap to 5:6:5, error under the squared metric (mx, my, mz).  t holds
// cluster_terms of the same prefix sums.
template <bool FOUR>
__device__ __forceinline__ float eval_terms(float4 x0, float4 x1, float4 x2, float4 sum, float4 t, float mx, float my,
                                            float mz, Cand *out)
{
    const float alpha2 = t.x, beta2 = t.y, ab = t.z, factor = t.w;
    float axr, axg, axb;
    if (FOUR) {
        const float x2r = x2.x - x1.x, x2g = x2.y - x1.y, x2b = x2.z - x1.z;
        const float x1r = x1.x - x0.x, x1g = x1.y - x0.y, x1b = x1.z - x0.z;
        axr = x2r * (1.0f / 3.0f) + (x1r * (2.0f / 3.0f) + x0.x);
        axg = x2g * (1.0f / 3.0f) + (x1g * (2.0f / 3.0f) + x0.y);
        axb = x2b * (1.0f / 3.0f) + (x1b * (2.0f / 3.0f) + x0.z);
    } else {
        const float x1r = x1.x - x0.x, x1g = x1.y - x0.y, x1b = x1.z - x0.z;
        axr = x0.x + x1r * 0.5f;
        axg = x0.y + x1g * 0.5f;
        axb = x0.z + x1b * 0.5f;
    }
    const float bxr = sum.x - axr, bxg = sum.y - axg, bxb = sum.z - axb;
    float ar = (axr * beta2 - bxr * ab) * factor;
    float ag = (axg * beta2 - bxg * ab) * factor;
    float abl = (axb * beta2 - bxb * ab) * factor;
    float br = (bxr * alpha2 - axr * ab) * factor;
    float bg = (bxg * alpha2 - axg * ab) * factor;
    float bb = (bxb * alpha2 - axb * ab) * factor;
    ar = round5(sse_sat(ar)); ag = round6(sse_sat(ag)); abl = round5(sse_sat(abl));
    br = round5(sse_sat(br)); bg = round6(sse_sat(bg)); bb = round5(sse_sat(bb));
    const float e2r = (ar * (br * ab - axr) - br * bxr) * 2.0f;
    const float e2g = (ag * (bg * ab - axg) - bg * bxg) * 2.0f;
    const float e2b = (abl * (bb * ab - axb) - bb * bxb) * 2.0f;
    const float e1r = (ar * ar) * alpha2 + ((br * br) * beta2 + e2r);
    const float e1g = (ag * ag) * alpha2 + ((bg * bg) * beta2 + e2g);
    const float e1b = (abl * abl) * alpha2 + ((bb * bb) * beta2 + e2b);
    if (out) {
        out->ax = ar; out->ay = ag; out->az = abl;
        out->bx = br; out->by = bg; out->bz = bb;
    }
    return e1r * mx + e1g * my + e1b * mz;
}

// One least-squares candidate, icbc.h:2501-2538 (four colours) or
// :2201-2236 (three colours).  sat is the zero-led table: entry 0 gives the
// zeros icbc uses for an absent cluster, and the FLT_MAX entries past the
// colour count are read exactly as the SSE build reads them.
template <bool FOUR>
__device__ __forceinline__ float eval_candidate(const float4 *sat, uint32_t packed, float4 sum, float mx, float my,
                                                float mz, Cand *out)
{
    const float4 x0 = sat[packed & 0xff];
    const float4 x1 = sat[(packed >> 8) & 0xff];
    const float4 x2 = FOUR ? sat[(packed >> 16) & 0xff] : x1;
    return eval_terms<FOUR>(x0, x1, x2, sum, cluster_terms<FOUR>(x0, x1, x2, sum.w), mx, my, mz, out);
}

// (error, p & 3, p >> 2) ordering: true if candidate (e, p) beats (be, bp).
__device__ __forceinline__ bool key_less(float e, int p, float be, int bp)
{
    if (e < be) return true;
    if (be < e) return false;
    if (!(e == be)) return false;  // NaN never wins
    const int lp = p & 3, lb = bp & 3;
    if (lp != lb) return lp < lb;
    return p < bp;
}

// Partition search over [0, roundup4(total)) for the block of this group.
template <bool FOUR>
__device__ void cluster_fit(const float4 *sat, const uint32_t *table, int total, int lane, float4 sum, float mx,
                            float my, float mz, Cand *best)
{
    const int limit = (total + 3) & ~3;
    float be = FLT_MAX;
    int bp = 0;
    Cand c;
    for (int p = lane; p < limit; p += kGroup) {
        const float e = eval_candidate<FOUR>(sat, table[p], sum, mx, my, mz, &c);
        if (e < be) { be = e; bp = p; }
    }
    // Each lane saw p & 3 == lane & 3 in increasing order; combine lanes.
    // be is never NaN, so the order is total and any combining order gives
    // the same winner.
    static_for<0, 4>([&](auto k) {
        constexpr int r = 8 >> decltype(k)::value;
        const float oe = gror<r>(be);
        const int op = gror<r>(bp);
        if (key_less(oe, op, be, bp)) { be = oe; bp = op; }
    });
    if (!(be < FLT_MAX)) {
        // No lane improved on FLT_MAX: the SSE build returns lane 0's initial zeros.
        best->ax = best->ay = best->az = best->bx = best->by = best->bz = 0.0f;
        return;
    }
    eval_candidate<FOUR>(sat, table[bp], sum, mx, my, mz, best);
}

// icbc.h:1580-1595 vector3_to_color16.
__device__ __forceinline__ uint32_t to565(float x, float y, float z, const float *mid5, const float *mid6)
{
    float fr = x * 31.0f, fg = y * 63.0f, fb = z * 31.0f;
    fr = (0.0f < fr) ? fr : 0.0f; fr = (fr < 31.0f) ? fr : 31.0f;
    fg = (0.0f < fg) ? fg : 0.0f; fg = (fg < 63.0f) ? fg : 63.0f;
    fb = (0.0f < fb) ? fb : 0.0f; fb = (fb < 31.0f) ? fb : 31.0f;
    uint32_t r = uint32_t(fr), g = uint32_t(fg), b = uint32_t(fb);
    r += (x > mid5[r]) ? 1u : 0u;
    g += (y > mid6[g]) ? 1u : 0u;
    b += (z > mid5[b]) ? 1u : 0u;
    return (r << 11) | (g << 5) | b;
}

__device__ __forceinline__ void expand565(uint32_t c, int &r, int &g, int &b)
{
    const int r5 = (c >> 11) & 31, g6 = (c >> 5) & 63, b5 = c & 31;
    r = (r5 << 3) | (r5 >> 2);
    g = (g6 << 2) | (g6 >> 4);
    b = (b5 << 3) | (b5 >> 2);
}

// D3D10 palette (icbc.h:2557-2585) as floats in [0,1] (color_to_vector3,
// through the k / 255.0f table div255).
__device__ __forceinline__ void palette(uint32_t c0, uint32_t c1, float p[4][3], const float *div255)
{
    int q[4][3];
    expand565(c0, q[0][0], q[0][1], q[0][2]);
    expand565(c1, q[1][0], q[1][1], q[1][2]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (c0 > c1) {
            q[2][k] = (2 * q[0][k] + q[1][k]) / 3;
            q[3][k] = (2 * q[1][k] + q[0][k]) / 3;
        } else {
            q[2][k] = (q[0][k] + q[1][k]) / 2;
            q[3][k] = 0;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) p[i][k] = div255[q[i][k]];
}

// output_block4 / output_block3 (icbc.h:2944-2980) for the group: lane l
// classifies texel l; the weighted MSE is summed in texel order.
template <bool FOUR>
__device__ float output_block(const Cand &c, float tr, float tg, float tb, float tw, float cwx, float cwy, float cwz,
                              const float *mid5, const float *mid6, const float *div255, int lane, uint2 *blk)
{
    uint32_t c0 = to565(c.ax, c.ay, c.az, mid5, mid6);
    uint32_t c1 = to565(c.bx, c.by, c.bz, mid5, mid6);
    if (FOUR ? (c0 < c1) : (c0 > c1)) { const uint32_t t = c0; c0 = c1; c1 = t; }
    float p[4][3];
    palette(c0, c1, p, div255);
    const float vr = tr * cwx, vg = tg * cwy, vb = tb * cwz;
    float d[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float pr = p[k][0] * cwx, pg = p[k][1] * cwy, pb = p[k][2] * cwz;
        float dr, dg, db;
        if (FOUR) { dr = vr - pr; dg = vg - pg; db = vb - pb; }
        else { dr = pr - vr; dg = pg - vg; db = pb - vb; }
        d[k] = dr * dr + dg * dg + db * db;
    }
    uint32_t idx;
    if (FOUR) {
        // compute_indices4 (icbc.h:2803-2825).
        const bool x0 = ((d[1] > d[2]) && (d[0] > d[2])) || ((d[0] > d[3]) && (d[1] > d[3]));
        const bool x1 = (d[0] > d[3]) && (d[2] > d[3]);
        idx = (x1 ? 1u : 0u) | (x0 ? 2u : 0u);
    } else {
        // compute_indices3 with allow_transparent_black (icbc.h:2836-2856).
        const float d3 = vr * vr + vg * vg + vb * vb;
        const bool i1 = d[1] < d[2];
        const bool i2 = (d[2] <= d[0]) && (d[2] <= d[1]);
        const bool i3 = (d3 <= d[0]) && (d3 <= d[1]) && (d3 <= d[2]);
        idx = ((i1 || i3) ? 1u : 0u) | ((i2 || i3) ? 2u : 0u);
    }
    // evaluate_mse (icbc.h:2738-2749) term for this texel.
    const float er = ((p[idx][0] - tr) * cwx) * 255.0f;
    const float eg = ((p[idx][1] - tg) * cwy) * 255.0f;
    const float eb = ((p[idx][2] - tb) * cwz) * 255.0f;
    const float term = tw * (er * er + eg * eg + eb * eb);
    float err = 0.0f;
    static_for<0, 16>([&](auto i) { err += gbcast<decltype(i)::value>(term); });
    int bits = int(idx << (2 * lane));
    bits |= gror<8>(bits);
    bits |= gror<4>(bits);
    bits |= gror<2>(bits);
    bits |= gror<1>(bits);
    blk->x = c0 | (c1 << 16);
    blk->y = uint32_t(bits);
    return err;
}

// icbc.h:1710-1882: PCA (power method) of the n clusters held one per lane,
// stable sort by projection, and the weighted prefix-sum table in LDS.
// Returns false only for an empty set.
__device__ __forceinline__ void compute_sat(Bc1Shared &sh, int g, int lane, int n, float kr, float kg, float kb, float kw)
{
    const bool live = lane < n;
    // Centroid: sequential sums over clusters 0..n-1 (computeCentroid).
    const float pr = kr * kw, pg = kg * kw, pb = kb * kw;
    float total = 0.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    static_for<0, 16>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (j < n) {
            total += gbcast<j>(kw);
            cr += gbcast<j>(pr);
            cg += gbcast<j>(pg);
            cb += gbcast<j>(pb);
        }
    });
    const float inv = 1.0f / total;
    cr *= inv; cg *= inv; cb *= inv;
    // Covariance (computeCovariance).
    const float ar = kr - cr, ag = kg - cg, ab = kb - cb;
    const float br = ar * kw, bg = ag * kw, bb = ab * kw;
    const float q0 = ar * br, q1 = ar * bg, q2 = ar * bb, q3 = ag * bg, q4 = ag * bb, q5 = ab * bb;
    float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f, m4 = 0.0f, m5 = 0.0f;
    static_for<0, 16>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (j < n) {
            m0 += gbcast<j>(q0); m1 += gbcast<j>(q1); m2 += gbcast<j>(q2);
            m3 += gbcast<j>(q3); m4 += gbcast<j>(q4); m5 += gbcast<j>(q5);
        }
    });
    // firstEigenVector_PowerMethod (icbc.h:1767-1791).
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
    if (!(m0 == 0 && m3 == 0 && m5 == 0)) {
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
            float norm = (y < x) ? x : y;           // icbc max(a,b) = (b<a)?a:b
            norm = (z < norm) ? norm : z;
            const float iv = 1.0f / norm;
            vx = x * iv; vy = y * iv; vz = z * iv;
        }
    }
    const float dp = kr * vx + kg * vy + kb * vz;
    // Stable insertion sort by projection.  Without NaNs its result is the
    // stable order by key, i.e. rank = #smaller + #equal-with-lower-index.
    const bool has_nan = gballot(live && dp != dp) != 0;
    int rank = 0;
    if (!has_nan) {
        static_for<0, 16>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j < n) {
                const float o = gbcast<j>(dp);
                rank += (o < dp || (o == dp && j < lane)) ? 1 : 0;
            }
        });
        if (live) sh.order[g][rank] = lane;
    } else {
        if (live) { sh.dps[g][lane] = dp; sh.order[g][lane] = lane; }
        ntx_wave_sync();
        if (lane == 0) {
            for (int i = 0; i < n; i++)
                for (int j = i; j > 0 && sh.dps[g][j] < sh.dps[g][j - 1]; j--) {
                    const float t = sh.dps[g][j]; sh.dps[g][j] = sh.dps[g][j - 1]; sh.dps[g][j - 1] = t;
                    const int o = sh.order[g][j]; sh.order[g][j] = sh.order[g][j - 1]; sh.order[g][j - 1] = o;
                }
        }
    }
    ntx_wave_sync();
    // Lane i takes the cluster at sorted position i and its weighted colour.
    const int src = live ? sh.order[g][lane] : 0;
    const float sr = gshfl(kr, src), sg = gshfl(kg, src), sb = gshfl(kb, src), sw = gshfl(kw, src);
    if (live) sh.col[g][lane] = make_float4(sr * sw, sg * sw, sb * sw, sw);
    ntx_wave_sync();
    // Prefix sums in sorted order (compute_sat's sequential loop).
    float4 acc = sh.col[g][0];
    for (int i = 1; i <= lane && i < n; i++) {
        const float4 v = sh.col[g][i];
        acc.x = acc.x + v.x; acc.y = acc.y + v.y; acc.z = acc.z + v.z; acc.w = acc.w + v.w;
    }
    sh.sat[g][lane + 1] = live ? acc : make_float4(FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX);
    if (lane == 0) sh.sat[g][0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    ntx_wave_sync();
}

// One channel of entry `idx` of the D3D10 palette of (c0, c1) as an 8-bit
// integer: shift/width select the 565 field, e0/e1 its bit expansion.
__device__ __forceinline__ int palette_entry(int e0, int e1, int idx, bool four)
{
    if (idx == 0) return e0;
    if (idx == 1) return e1;
    if (four) return idx == 2 ? (2 * e0 + e1) / 3 : (2 * e1 + e0) / 3;
    return idx == 2 ? (e0 + e1) / 2 : 0;
}

// compute_indices (icbc.h:2885-2930, the SSE form) for this lane's texel
// against the palette of (c0, c1); the group's 32 index bits.
__device__ __forceinline__ int group_indices(uint32_t c0, uint32_t c1, float tr, float tg, float tb, float cwx, float cwy,
                                             float cwz, const float *div255, int lane)
{
    float p[4][3];
    palette(c0, c1, p, div255);
    const float cr = tr * cwx, cg = tg * cwy, cb = tb * cwz;
    float d[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float dr = cr - p[k][0] * cwx, dg = cg - p[k][1] * cwy, db = cb - p[k][2] * cwz;
        d[k] = dr * dr + dg * dg + db * db;
    }
    const bool i1 = (d[1] <= d[0]) && (d[1] < d[2]) && (d[1] < d[3]);
    const bool i2 = (d[2] <= d[0]) && (d[2] <= d[1]) && (d[2] < d[3]);
    const bool i3 = (d[3] <= d[0]) && (d[3] <= d[1]) && (d[3] <= d[2]);
    int bits = int(((i1 || i3) ? 1u : 0u) | ((i2 || i3) ? 2u : 0u)) << (2 * lane);
    bits |= gror<8>(bits);
    bits |= gror<4>(bits);
    bits |= gror<2>(bits);
    bits |= gror<1>(bits);
    return bits;
}

// refine_endpoints (icbc.h:3329-3406) with three_color_mode set, as
// compress_dxt1 calls it at Quality_Level9 (CompressorDXT1::compressBlock
// passes three_color_mode = true, BlockCompressor.cpp:211-222): up to 256
// single steps of one endpoint's 5:6:5 fields (the fields wrap, as the
// reference's bit-fields do), endpoint c1 for steps 0-15, c0 for 16-31, ...;
// each candidate takes the indices of the current block's palette and is
// kept when its evaluate_mse is lower; stops 33 steps after the last gain.
// Lane l holds texel l; the error is summed in texel order.
__device__ __forceinline__ uint2 refine_endpoints(uint2 blk, float best, float tr, float tg, float tb, float tw,
                                                  float cwx, float cwy, float cwz, const float *div255, int lane)
{
    // The 16 steps (dr, dg, db), each component + 1 in two bits.
    constexpr uint32_t kSteps[16] = { 0x16, 0x19, 0x25, 0x14, 0x11, 0x05, 0x1a, 0x26, 0x29,
                                      0x10, 0x04, 0x01, 0x18, 0x12, 0x21, 0x09 };
    uint32_t c0 = blk.x & 0xffffu, c1 = blk.x >> 16;
    int ind = int(blk.y);
    int idx = group_indices(c0, c1, tr, tg, tb, cwx, cwy, cwz, div255, lane);
    int mine = (idx >> (2 * lane)) & 3;
    int last = 0;
    for (int i = 0; i < 256; i++) {
        const uint32_t st = kSteps[i & 15];
        const int dr = int(st & 3) - 1, dg = int((st >> 2) & 3) - 1, db = int((st >> 4) & 3) - 1;
        uint32_t r0 = c0, r1 = c1;
        uint32_t &t = ((i >> 4) & 1) ? r0 : r1;
        t = (uint32_t((int(t >> 11) + dr) & 31) << 11) | (uint32_t((int((t >> 5) & 63) + dg) & 63) << 5) |
            uint32_t((int(t & 31) + db) & 31);
        // evaluate_mse (icbc.h:2738-2749) of (r0, r1, idx): this texel's term.
        int er0, eg0, eb0, er1, eg1, eb1;
        expand565(r0, er0, eg0, eb0);
        expand565(r1, er1, eg1, eb1);
        const bool four = r0 > r1;
        const float pr = div255[palette_entry(er0, er1, mine, four)];
        const float pg = div255[palette_entry(eg0, eg1, mine, four)];
        const float pb = div255[palette_entry(eb0, eb1, mine, four)];
        const float xr = ((pr - tr) * cwx) * 255.0f, xg = ((pg - tg) * cwy) * 255.0f, xb = ((pb - tb) * cwz) * 255.0f;
        const float term = tw * (xr * xr + xg * xg + xb * xb);
        float err = 0.0f;
        static_for<0, 16>([&](auto k) { err += gbcast<decltype(k)::value>(term); });
        if (err < best) {
            best = err;
            c0 = r0;
            c1 = r1;
            ind = idx;
            last = i;
            idx = group_indices(c0, c1, tr, tg, tb, cwx, cwy, cwz, div255, lane);
            mine = (idx >> (2 * lane)) & 3;
        }
        if (i - last > 32) break;
    }
    return make_uint2(c0 | (c1 << 16), uint32_t(ind));
}

// compress_dxt1_single_color_optimal (icbc.h:3274-3289) on the rounded
// colour (vector3_to_color32).
__device__ __forceinline__ uint2 single_color_block(const ntx::Tables *tab, float kr, float kg, float kb)
{
    const int r = int(uint8_t(sse_sat(kr) * 255 + 0.5f));
    const int gg = int(uint8_t(sse_sat(kg) * 255 + 0.5f));
    const int b = int(uint8_t(sse_sat(kb) * 255 + 0.5f));
    uint32_t c0 = uint32_t(tab->match5[r] & 0xff) << 11 | uint32_t(tab->match6[gg] & 0xff) << 5 | uint32_t(tab->match5[b] & 0xff);
    uint32_t c1 = uint32_t(tab->match5[r] >> 8) << 11 | uint32_t(tab->match6[gg] >> 8) << 5 | uint32_t(tab->match5[b] >> 8);
    uint32_t idx = 0xaaaaaaaau;
    if (c0 < c1) { const uint32_t t = c0; c0 = c1; c1 = t; idx ^= 0x55555555u; }
    return make_uint2(c0 | (c1 << 16), idx);
}

// The icbc scalar helpers (icbc.h:183-193): max(a, b) = (b < a) ? a : b,
// min(a, b) = (a < b) ? a : b, saturate = min(max(x, 0), 1).
__device__ __forceinline__ float icbc_max(float a, float b) { return (b < a) ? a : b; }
__device__ __forceinline__ float icbc_min(float a, float b) { return (a < b) ? a : b; }
__device__ __forceinline__ float icbc_saturate(float x) { return icbc_min(icbc_max(x, 0.0f), 1.0f); }

// Sums over the group's 16 lanes, each in lane order (the reference's
// sequential loops); the N sums advance together, one lane at a time, so only
// N broadcasts are live at once.
template <int N>
__device__ __forceinline__ void group_sums(const float (&v)[N], float (&s)[N])
{
    #pragma unroll
    for (int j = 0; j < N; j++) s[j] = 0.0f;
    static_for<0, 16>([&](auto k) {
        #pragma unroll
        for (int j = 0; j < N; j++) s[j] += gbcast<decltype(k)::value>(v[j]);
    });
}

// compress_dxt1 at Quality_Level1, nvtt's Quality_Fastest (icbc.h:3485-3540
// with box_fit and least_squares_fit and no cluster fit): all 16 texels
// (those outside the image as black), fit_colors_bbox + inset_bbox +
// select_diagonal (icbc.h:3096-3153), output_block4, then
// optimize_end_points4 on those indices (icbc.h:2984-3018) and output_block4
// again, keeping the lower error.
__device__ __forceinline__ uint2 encode_box_fit(float tr, float tg, float tb, float tw, float cwx, float cwy, float cwz,
                                                const float *mid5, const float *mid6, const float *div255, int lane)
{
    // fit_colors_bbox: c0 = running max from 0, c1 = running min from 1, in texel order.
    float x0 = 0.0f, y0 = 0.0f, z0 = 0.0f, x1 = 1.0f, y1 = 1.0f, z1 = 1.0f;
    static_for<0, 16>([&](auto k) {
        constexpr int i = decltype(k)::value;
        const float r = gbcast<i>(tr), g = gbcast<i>(tg), b = gbcast<i>(tb);
        x0 = icbc_max(x0, r); y0 = icbc_max(y0, g); z0 = icbc_max(z0, b);
        x1 = icbc_min(x1, r); y1 = icbc_min(y1, g); z1 = icbc_min(z1, b);
    });
    // inset_bbox.
    const float bias = (8.0f / 255.0f) / 16.0f;
    const float ix = (x0 - x1) / 16.0f - bias, iy = (y0 - y1) / 16.0f - bias, iz = (z0 - z1) / 16.0f - bias;
    x0 = icbc_saturate(x0 - ix); y0 = icbc_saturate(y0 - iy); z0 = icbc_saturate(z0 - iz);
    x1 = icbc_saturate(x1 + ix); y1 = icbc_saturate(y1 + iy); z1 = icbc_saturate(z1 + iz);
    // select_diagonal: signs of the x-z and y-z covariances about the box centre.
    const float cx = (x0 + x1) * 0.5f, cy = (y0 + y1) * 0.5f, cz = (z0 + z1) * 0.5f;
    const float tx = tr - cx, ty = tg - cy, tz = tb - cz;
    float cov[2];
    group_sums<2>({ tx * tz, ty * tz }, cov);
    const float cov_xz = cov[0], cov_yz = cov[1];
    if (cov_xz < 0.0f) { const float t = x0; x0 = x1; x1 = t; }
    if (cov_yz < 0.0f) { const float t = y0; y0 = y1; y1 = t; }
    Cand c = { x0, y0, z0, x1, y1, z1 };
    uint2 blk;
    float err = output_block<true>(c, tr, tg, tb, tw, cwx, cwy, cwz, mid5, mid6, div255, lane, &blk);
    // optimize_end_points4 (no weights): index bits -> beta in {0, 1, 1/3, 2/3}.
    const uint32_t bits = (blk.y >> (2 * lane)) & 3u;
    float beta = float(bits & 1u);
    if (bits & 2u) beta = (1.0f + beta) / 3.0f;
    const float alpha = 1.0f - beta;
    float m[9];
    group_sums<9>({ alpha * alpha, beta * beta, alpha * beta, alpha * tr, alpha * tg, alpha * tb, beta * tr, beta * tg,
                    beta * tb },
                  m);
    const float a2 = m[0], b2 = m[1], ab = m[2];
    const float axr = m[3], axg = m[4], axb = m[5];
    const float bxr = m[6], bxg = m[7], bxb = m[8];
    const float denom = a2 * b2 - ab * ab;
    // icbc returns the first block when equal(denom, 0).  The test is a
    // select at the end, not a branch: the lane broadcasts of the sums cannot
    // move into a branch that the four groups of a wave may take differently,
    // so a branch here keeps all of them live at once.  A degenerate denom
    // gives non-finite factors, which saturate() maps to 0: harmless.
    const bool degenerate = fabsf(denom - 0.0f) < 0.0001f;
    const float factor = 1.0f / denom;
    Cand o;
    o.ax = icbc_saturate((axr * b2 - bxr * ab) * factor);
    o.ay = icbc_saturate((axg * b2 - bxg * ab) * factor);
    o.az = icbc_saturate((axb * b2 - bxb * ab) * factor);
    o.bx = icbc_saturate((bxr * a2 - axr * ab) * factor);
    o.by = icbc_saturate((bxg * a2 - axg * ab) * factor);
    o.bz = icbc_saturate((bxb * a2 - axb * ab) * factor);
    uint2 blk2;
    const float err2 = output_block<true>(o, tr, tg, tb, tw, cwx, cwy, cwz, mid5, mid6, div255, lane, &blk2);
    return (!degenerate && err2 < err) ? blk2 : blk;
}

// Encodes block (bx, by) of the w x h planar RGBA image for this group;
// every lane of the group returns the 8-byte block, the colour count n
// after reduce_colors and the count m seen by the three-colour fit (-1 if
// it did not run).  Uses only wave-level synchronisation.  LEVEL: icbc
// quality level 1, 8 or 9, a template argument so that each instance holds
// only its own path's registers (the level-8 kernel keeps its occupancy).
// refine (LEVEL 9 only, uniform across the workgroup): run refine_endpoints,
// so that one instance serves levels 8 and 9 where both are needed.
template <int LEVEL = 8>
__device__ __forceinline__ uint2 encode_block(Bc1Shared &sh, const float *src, int w, int h, int bx, int by,
                                              int alpha_mode, float cwx, float cwy, float cwz,
                                              const ntx::Tables *tab, int g, int lane, int *n_out, int *m_out,
                                              bool refine = true)
{
    static_assert(LEVEL == 1 || LEVEL == 8 || LEVEL == 9, "icbc levels 1, 8 and 9");
    // Gather (BlockCompressor.cpp:116-166): texel l; outside the image -> colour 0, weight 0.
    const int x = bx * 4 + (lane & 3), y = by * 4 + (lane >> 2);
    const size_t plane = size_t(w) * size_t(h);
    float tr = 0.0f, tg = 0.0f, tb = 0.0f, tw = 0.0f;
    if (x < w && y < h) {
        const size_t o = size_t(y) * size_t(w) + size_t(x);
        tr = src[o];
        tg = src[o + plane];
        tb = src[o + 2 * plane];
        if (alpha_mode == NTX_ALPHA_TRANSPARENCY) {
            const float a = src[o + 3 * plane];
            float s = (0.0f < a) ? a : 0.0f;
            tw = (s < 1.0f) ? s : 1.0f;
        } else {
            tw = 1.0f;
        }
    }

    if constexpr (LEVEL == 1) {
        *n_out = 16;
        *m_out = -1;
        return encode_box_fit(tr, tg, tb, tw, cwx, cwy, cwz, tab->mid5, tab->mid6, sh.div255, lane);
    }

    // reduce_colors (icbc.h:1638-1683): lane j holds cluster j.
    const float thr = 1.0f / 256;
    float kr = 0.0f, kg = 0.0f, kb = 0.0f, kw = 0.0f;
    int n = 0;
    bool any_black = false;
    static_for<0, 16>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const float ir = gbcast<i>(tr), ig = gbcast<i>(tg), ib = gbcast<i>(tb), iw = gbcast<i>(tw);
        if (iw > 0) {
            const bool match = lane < n && fabsf(kr - ir) < thr && fabsf(kg - ig) < thr && fabsf(kb - ib) < thr;
            const uint32_t m = gballot(match);
            if (m) {
                if (lane == __builtin_ctz(m)) {
                    const float s = kw + iw;
                    const float nr = kr * kw + ir * iw, ng = kg * kw + ig * iw, nb = kb * kw + ib * iw;
                    kr = nr / s; kg = ng / s; kb = nb / s;
                    kw = s;
                }
            } else {
                if (lane == n) { kr = ir; kg = ig; kb = ib; kw = iw; }
                n++;
            }
            if (ir < 1.0f / 8 && ig < 1.0f / 8 && ib < 1.0f / 8) any_black = true;
        }
    });

    const float cwsq_x = cwx * cwx, cwsq_y = cwy * cwy, cwsq_z = cwz * cwz;
    const float *mid5 = tab->mid5;
    const float *mid6 = tab->mid6;
    uint2 result;
    int m_count = -1;
    if (n == 0) {
        result = make_uint2(0u, 0u);
    } else if (n == 1) {
        result = single_color_block(tab, kr, kg, kb);
    } else {
        // compress_dxt1_cluster_fit (icbc.h:3292-3326).
        compute_sat(sh, g, lane, n, kr, kg, kb, kw);
        float4 sum = sh.sat[g][n];
        Cand c;
        cluster_fit<true>(sh.sat[g], sh.four, tab->four_total[n - 1], lane, sum, cwsq_x, cwsq_y, cwsq_z, &c);
        uint2 b4;
        const float e4 = output_block<true>(c, tr, tg, tb, tw, cwx, cwy, cwz, mid5, mid6, sh.div255, lane, &b4);
        float best = e4;
        result = b4;
        bool run3 = true;
        int m = n;
        if (any_black) {
            // skip_blacks (icbc.h:1685-1703): compact the non-black clusters.
            const bool keep = lane < n && !(kr < 1.0f / 8 && kg < 1.0f / 8 && kb < 1.0f / 8);
            const uint32_t km = gballot(keep);
            m = __builtin_popcount(km);
            if (m == 0) {
                run3 = false;
            } else {
                // Lane d takes the d-th kept cluster, preserving order.
                int srcl = 0;
                if (lane < m) {
                    uint32_t mm = km;
                    for (int d = 0; d < lane; d++) mm &= mm - 1u;
                    srcl = __builtin_ctz(mm);
                }
                const float nr = gshfl(kr, srcl), ng = gshfl(kg, srcl), nb = gshfl(kb, srcl), nw = gshfl(kw, srcl);
                compute_sat(sh, g, lane, m, nr, ng, nb, nw);
                sum = sh.sat[g][m];
            }
        }
        if (run3) {
            m_count = m;
            cluster_fit<false>(sh.sat[g], sh.three, tab->three_total[m - 1], lane, sum, cwsq_x, cwsq_y, cwsq_z, &c);
            uint2 b3;
            const float e3 = output_block<false>(c, tr, tg, tb, tw, cwx, cwy, cwz, mid5, mid6, sh.div255, lane, &b3);
            if (e3 < best) { best = e3; result = b3; }
        }
        if (!(best < FLT_MAX)) result = make_uint2(0u, 0u);
        // Quality_Level9 (icbc.h:3551-3553): the cluster fit's block refined.
        else if constexpr (LEVEL == 9) {
            if (refine) result = refine_endpoints(result, best, tr, tg, tb, tw, cwx, cwy, cwz, sh.div255, lane);
        }
    }
    *n_out = n;
    *m_out = m_count;
    return result;
}

}  // namespace ntx_bc1
