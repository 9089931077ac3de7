// bc1_lane.hpp -- icbc v1.05 compress_dxt1 at Quality_Level8 (icbc.h:3485-3559,
// the SSE4.1 build: no FMA, exact division, round-half-even, 4-lane
// tie-break) with one 4x4 block per lane, for the two-pass BC1 encoder of
// large images (bc1.hip: bc1_classify_kernel, bc1_search_kernel).
//
// A block's clusters, and then its prefix-sum table, live in column `lane`
// of a [17][64] float4 LDS array (rows[r][lane]): row r is read by every
// lane at the same r, so the accesses are conflict-free.
//
// The cluster tables (icbc.h:1906-2012) list, for t = 1..16, the new
// prefix triples c0 <= c1 <= c2 = t in (c0, c1) order, and the pairs
// c0 <= c1 = t in c0 order; the table of a colour count n is the first
// C(n+3,3)-1 (C(n+2,2)-1) entries.  The searches below walk exactly that
// order as nested loops (c2 outer, c0, c1 inner), so the candidate index p
// advances by one per inner step, and evaluate the entries the SSE build's
// four-wide loop reads past the end (round-up to a multiple of 4) as well.
// The winner is the SSE build's: the least (error, p & 3, p >> 2).
#pragma once

#include "bc1_block.hpp"

namespace ntx_bc1 {

using LaneRows = float4 (*)[64];

__device__ __forceinline__ bool is_black(float r, float g, float b)
{
    // icbc.h:1628-1635.
    return r < 1.0f / 8 && g < 1.0f / 8 && b < 1.0f / 8;
}

// FloatColorCompressor's gather (BlockCompressor.cpp:116-166) of block
// (bx, by): texel i = (r, g, b, weight); outside the image colour 0, weight 0;
// weight = saturate(alpha) for Transparency, else 1.  rows16: every row of
// every plane starts 16-byte aligned, so an inner block reads each row of
// each plane with one 16-byte load.
__device__ __forceinline__ void lane_gather(const float *__restrict__ src, int w, int h, int bx, int by, int alpha_mode,
                                            bool rows16, float4 (&t)[16])
{
    const size_t plane = size_t(w) * size_t(h);
    const bool transparency = alpha_mode == NTX_ALPHA_TRANSPARENCY;
    if (rows16 && 4 * bx + 4 <= w && 4 * by + 4 <= h) {
        const float *p0 = src + size_t(4 * by) * size_t(w) + size_t(4 * bx);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float *p = p0 + size_t(r) * size_t(w);
            const float4 vr = *reinterpret_cast<const float4 *>(p);
            const float4 vg = *reinterpret_cast<const float4 *>(p + plane);
            const float4 vb = *reinterpret_cast<const float4 *>(p + 2 * plane);
            float4 wt = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            if (transparency) {
                const float4 va = *reinterpret_cast<const float4 *>(p + 3 * plane);
                wt = make_float4(icbc_saturate(va.x), icbc_saturate(va.y), icbc_saturate(va.z), icbc_saturate(va.w));
            }
            t[4 * r + 0] = make_float4(vr.x, vg.x, vb.x, wt.x);
            t[4 * r + 1] = make_float4(vr.y, vg.y, vb.y, wt.y);
            t[4 * r + 2] = make_float4(vr.z, vg.z, vb.z, wt.z);
            t[4 * r + 3] = make_float4(vr.w, vg.w, vb.w, wt.w);
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int x = 4 * bx + (i & 3), y = 4 * by + (i >> 2);
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (x < w && y < h) {
            const size_t o = size_t(y) * size_t(w) + size_t(x);
            v.x = src[o];
            v.y = src[o + plane];
            v.z = src[o + 2 * plane];
            if (transparency) {
                const float a = src[o + 3 * plane];
                float s = (0.0f < a) ? a : 0.0f;
                v.w = (s < 1.0f) ? s : 1.0f;
            } else {
                v.w = 1.0f;
            }
        }
        t[i] = v;
    }
}

// reduce_colors (icbc.h:1638-1683): texels of positive weight merge, in
// texel order, into the first earlier cluster within 1/256 in every channel
// (weighted running mean), or open a new one.  Leaves clusters 0..n-1 as
// (r, g, b, weight) in rows[0..n-1]; returns n.  any_black: some texel of
// positive weight is black (is_black).
__device__ __forceinline__ int lane_reduce(LaneRows rows, int lane, const float4 (&t)[16], bool &any_black)
{
    const float thr = 1.0f / 256;
    int n = 0;
    bool black = false;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float4 ci = t[i];
        if (ci.w > 0) {
            // First matching cluster: clusters j < n (n <= i here).
            int jm = -1;
#pragma unroll
            for (int j = i - 1; j >= 0; j--) {
                const float4 k = rows[j][lane];
                if (j < n && fabsf(k.x - ci.x) < thr && fabsf(k.y - ci.y) < thr && fabsf(k.z - ci.z) < thr) jm = j;
            }
            if (jm >= 0) {
                const float4 k = rows[jm][lane];
                const float s = k.w + ci.w;
                const float nr = k.x * k.w + ci.x * ci.w, ng = k.y * k.w + ci.y * ci.w, nb = k.z * k.w + ci.z * ci.w;
                rows[jm][lane] = make_float4(nr / s, ng / s, nb / s, s);
            } else {
                rows[n][lane] = ci;
                n++;
            }
            if (is_black(ci.x, ci.y, ci.z)) black = true;
        }
    }
    any_black = black;
    return n;
}

// skip_blacks (icbc.h:1685-1703) in place: keeps the clusters that are not
// black, in order; returns their count.
__device__ __forceinline__ int lane_skip_blacks(LaneRows rows, int lane, int n)
{
    int m = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        if (j < n) {
            const float4 k = rows[j][lane];
            if (!is_black(k.x, k.y, k.z)) {
                rows[m][lane] = k;
                m++;
            }
        }
    }
    return m;
}

// True if some cluster 0..n-1 is black (skip_blacks would drop it).
__device__ __forceinline__ bool lane_has_black_cluster(LaneRows rows, int lane, int n)
{
    bool any = false;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        if (j < n) {
            const float4 k = rows[j][lane];
            any = any || is_black(k.x, k.y, k.z);
        }
    }
    return any;
}

// compute_sat (icbc.h:1796-1882) of the n >= 1 clusters in rows[0..n-1]:
// weighted PCA by the power method, the stable insertion sort by projection,
// and the prefix sums in sorted order.  Leaves row 0 zero, rows 1..n the
// prefix sums (r*w, g*w, b*w, w) and rows n+1..16 FLT_MAX (the entries the
// SSE build reads past the count), and returns row n.
__device__ __forceinline__ float4 lane_sat(LaneRows rows, int lane, int n)
{
    float4 c[16];
#pragma unroll
    for (int j = 0; j < 16; j++) c[j] = rows[j][lane];
    // computeCentroid: sequential sums over clusters 0..n-1.
    float total = 0.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; j++)
        if (j < n) {
            total += c[j].w;
            cr += c[j].x * c[j].w;
            cg += c[j].y * c[j].w;
            cb += c[j].z * c[j].w;
        }
    const float inv = 1.0f / total;
    cr *= inv; cg *= inv; cb *= inv;
    // computeCovariance.
    float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f, m4 = 0.0f, m5 = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; j++)
        if (j < n) {
            const float ar = c[j].x - cr, ag = c[j].y - cg, ab = c[j].z - cb;
            const float br = ar * c[j].w, bg = ag * c[j].w, bb = ab * c[j].w;
            m0 += ar * br; m1 += ar * bg; m2 += ar * bb;
            m3 += ag * bg; m4 += ag * bb; m5 += ab * bb;
        }
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
            float norm = (y < x) ? x : y;  // icbc max(a,b) = (b<a)?a:b
            norm = (z < norm) ? norm : z;
            const float iv = 1.0f / norm;
            vx = x * iv; vy = y * iv; vz = z * iv;
        }
    }
    // The stable insertion sort as the network it executes: pass i moves
    // element i down while it is strictly below its left neighbour (a NaN
    // stops it, as in the reference loop).
    float dps[16];
    int ord[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        dps[j] = c[j].x * vx + c[j].y * vy + c[j].z * vz;
        ord[j] = j;
    }
#pragma unroll
    for (int i = 1; i < 16; i++) {
        bool go = i < n;
#pragma unroll
        for (int j = i; j > 0; j--) {
            go = go && dps[j] < dps[j - 1];
            if (go) {
                const float td = dps[j]; dps[j] = dps[j - 1]; dps[j - 1] = td;
                const int to = ord[j]; ord[j] = ord[j - 1]; ord[j - 1] = to;
            }
        }
    }
    // The clusters in sorted order (all read before any row is rewritten).
    float4 s[16];
#pragma unroll
    for (int r = 0; r < 16; r++) s[r] = rows[ord[r]][lane];
    const float4 fmax = make_float4(FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX);
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sum = acc;
    rows[0][lane] = acc;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const float wgt = s[r].w;
        const float4 v = make_float4(s[r].x * wgt, s[r].y * wgt, s[r].z * wgt, wgt);
        if (r == 0) acc = v;
        else acc = make_float4(acc.x + v.x, acc.y + v.y, acc.z + v.z, acc.w + v.w);
        rows[r + 1][lane] = r < n ? acc : fmax;
        if (r == n - 1) sum = acc;
    }
    return sum;
}

// Best-candidate bookkeeping: candidate p with error e replaces the current
// best when (e, p & 3, p >> 2) is smaller; the key packs (p & 3, p >> 2) so
// that a later p of the same residue never wins a tie.
struct Best {
    float e = FLT_MAX;
    int key = 0;
    __device__ __forceinline__ void offer(float err, int p)
    {
        const int k = ((p & 3) << 16) | (p >> 2);
        if (err < e || (err == e && k < key)) { e = err; key = k; }
    }
    __device__ __forceinline__ int index() const { return ((key & 0xffff) << 2) | (key >> 16); }
};

// cluster_fit_four (icbc.h:2250-2549) over the rows of a block with n
// clusters.  `units`: every texel of every live lane has weight 1, so the
// cluster weights are integer counts summing to 16 and the weight terms come
// from tab->bc1w4; when n is also 16 the counts are the loop indices and the
// terms are scalar loads shared by the wave.  n must be the same in every
// lane of the wave.
__device__ __forceinline__ Best lane_search4(LaneRows rows, int lane, int n, float4 sum, float mx, float my, float mz,
                                            const ntx::Tables *__restrict__ tab, bool units)
{
    Best best;
    int p = 0;
    const float4 *__restrict__ w4 = reinterpret_cast<const float4 *>(tab->bc1w4);
    if (units && n == 16) {
        for (int c = 1; c <= 16; c++) {
            const float4 x2 = rows[c][lane];
            for (int a = 0; a <= c; a++) {
                const float4 x0 = rows[a][lane];
                for (int b = a; b <= c; b++, p++) {
                    const float4 x1 = rows[b][lane];
                    const float4 t = w4[(a * 17 + b) * 17 + c];
                    best.offer(eval_terms<true>(x0, x1, x2, sum, t, mx, my, mz, nullptr), p);
                }
            }
        }
        return best;
    }
    if (units) {
        // Entries by 32-bit byte offsets from the table's (scalar) base, so
        // each load takes one vector offset: ((x0.w * 17 + x1.w) * 17 + x2.w) * 16.
        const char *w4b = reinterpret_cast<const char *>(w4);
        for (int c = 1; c <= n; c++) {
            const float4 x2 = rows[c][lane];
            for (int a = 0; a <= c; a++) {
                const float4 x0 = rows[a][lane];
                const uint32_t base = ((uint32_t(x0.w) * 17u) * 17u + uint32_t(x2.w)) * 16u;
                for (int b = a; b <= c; b++, p++) {
                    const float4 x1 = rows[b][lane];
                    const float4 t = *reinterpret_cast<const float4 *>(w4b + (base + uint32_t(x1.w) * 272u));
                    best.offer(eval_terms<true>(x0, x1, x2, sum, t, mx, my, mz, nullptr), p);
                }
            }
        }
    } else {
        for (int c = 1; c <= n; c++) {
            const float4 x2 = rows[c][lane];
            for (int a = 0; a <= c; a++) {
                const float4 x0 = rows[a][lane];
                for (int b = a; b <= c; b++, p++) {
                    const float4 x1 = rows[b][lane];
                    best.offer(eval_terms<true>(x0, x1, x2, sum, cluster_terms<true>(x0, x1, x2, sum.w), mx, my, mz,
                                                nullptr),
                               p);
                }
            }
        }
    }
    // The SSE loop's last group of four reads the next entries of the
    // table: (0, q, n + 1), whose row n + 1 is FLT_MAX.
    const int pad = (4 - (p & 3)) & 3;
    if (pad) {
        const float4 x0 = rows[0][lane], x2 = rows[n + 1][lane];
        for (int q = 0; q < pad; q++, p++) {
            const float4 x1 = rows[q][lane];
            best.offer(eval_terms<true>(x0, x1, x2, sum, cluster_terms<true>(x0, x1, x2, sum.w), mx, my, mz, nullptr),
                       p);
        }
    }
    return best;
}

// cluster_fit_three (icbc.h:2016-2247), the same way: pairs (c0, c1 = t) for
// t = 1..m, then the read-past entries (q, m + 1).  m need not be the same
// in every lane.  `units`: as for lane_search4, and the set's weights sum to
// 16 (no cluster was dropped by skip_blacks).
__device__ __forceinline__ Best lane_search3(LaneRows rows, int lane, int m, float4 sum, float mx, float my, float mz,
                                            const ntx::Tables *__restrict__ tab, bool units)
{
    Best best;
    int p = 0;
    const float4 *__restrict__ w3 = reinterpret_cast<const float4 *>(tab->bc1w3);
    if (units && m == 16) {
        for (int b = 1; b <= 16; b++) {
            const float4 x1 = rows[b][lane];
            for (int a = 0; a <= b; a++, p++) {
                const float4 x0 = rows[a][lane];
                best.offer(eval_terms<false>(x0, x1, x1, sum, w3[a * 17 + b], mx, my, mz, nullptr), p);
            }
        }
        return best;
    }
    const char *w3b = reinterpret_cast<const char *>(w3);
    for (int b = 1; b <= m; b++) {
        const float4 x1 = rows[b][lane];
        for (int a = 0; a <= b; a++, p++) {
            const float4 x0 = rows[a][lane];
            // As in lane_search4: a 32-bit byte offset (x0.w * 17 + x1.w) * 16.
            const float4 t = units ? *reinterpret_cast<const float4 *>(w3b + (uint32_t(x0.w) * 272u + uint32_t(x1.w) * 16u))
                                   : cluster_terms<false>(x0, x1, x1, sum.w);
            best.offer(eval_terms<false>(x0, x1, x1, sum, t, mx, my, mz, nullptr), p);
        }
    }
    const int pad = (4 - (p & 3)) & 3;
    if (pad) {
        const float4 x1 = rows[m + 1][lane];
        for (int q = 0; q < pad; q++, p++) {
            const float4 x0 = rows[q][lane];
            best.offer(eval_terms<false>(x0, x1, x1, sum, cluster_terms<false>(x0, x1, x1, sum.w), mx, my, mz, nullptr),
                       p);
        }
    }
    return best;
}

// The winning candidate's endpoints (the SSE build keeps lane 0's zeros when
// no candidate beat FLT_MAX).
template <bool FOUR>
__device__ __forceinline__ Cand lane_winner(LaneRows rows, int lane, const Best &best, float4 sum, float mx, float my,
                                           float mz, const ntx::Tables *__restrict__ tab)
{
    Cand c = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
    if (!(best.e < FLT_MAX)) return c;
    const uint32_t packed = FOUR ? tab->four[best.index()] : tab->three[best.index()];
    const float4 x0 = rows[packed & 0xff][lane];
    const float4 x1 = rows[(packed >> 8) & 0xff][lane];
    const float4 x2 = FOUR ? rows[(packed >> 16) & 0xff][lane] : x1;
    eval_terms<FOUR>(x0, x1, x2, sum, cluster_terms<FOUR>(x0, x1, x2, sum.w), mx, my, mz, &c);
    return c;
}

// output_block4 / output_block3 (icbc.h:2944-2980) for one lane's block:
// 5:6:5 endpoints (vector3_to_color16), order, the D3D10 palette, each
// texel's index (compute_indices4 / compute_indices3 with transparent
// black), and evaluate_mse summed in texel order.
template <bool FOUR>
__device__ __forceinline__ float lane_output(const Cand &c, const float4 (&t)[16], float cwx, float cwy, float cwz,
                                             const float *mid5, const float *mid6, const float *div255, uint2 *blk)
{
    uint32_t c0 = to565(c.ax, c.ay, c.az, mid5, mid6);
    uint32_t c1 = to565(c.bx, c.by, c.bz, mid5, mid6);
    if (FOUR ? (c0 < c1) : (c0 > c1)) { const uint32_t tmp = c0; c0 = c1; c1 = tmp; }
    float p[4][3];
    palette(c0, c1, p, div255);
    float pw[4][3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        pw[k][0] = p[k][0] * cwx;
        pw[k][1] = p[k][1] * cwy;
        pw[k][2] = p[k][2] * cwz;
    }
    float err = 0.0f;
    uint32_t bits = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float tr = t[i].x, tg = t[i].y, tb = t[i].z, tw = t[i].w;
        const float vr = tr * cwx, vg = tg * cwy, vb = tb * cwz;
        float d[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float dr, dg, db;
            if (FOUR) { dr = vr - pw[k][0]; dg = vg - pw[k][1]; db = vb - pw[k][2]; }
            else { dr = pw[k][0] - vr; dg = pw[k][1] - vg; db = pw[k][2] - vb; }
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
        // evaluate_mse (icbc.h:2738-2749) term of this texel.
        const float qr = idx == 0 ? p[0][0] : idx == 1 ? p[1][0] : idx == 2 ? p[2][0] : p[3][0];
        const float qg = idx == 0 ? p[0][1] : idx == 1 ? p[1][1] : idx == 2 ? p[2][1] : p[3][1];
        const float qb = idx == 0 ? p[0][2] : idx == 1 ? p[1][2] : idx == 2 ? p[2][2] : p[3][2];
        const float er = ((qr - tr) * cwx) * 255.0f;
        const float eg = ((qg - tg) * cwy) * 255.0f;
        const float eb = ((qb - tb) * cwz) * 255.0f;
        err += tw * (er * er + eg * eg + eb * eb);
        bits |= idx << (2 * i);
    }
    blk->x = c0 | (c1 << 16);
    blk->y = bits;
    return err;
}

}  // namespace ntx_bc1
