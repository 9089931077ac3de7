/*
 * ntx_oracle.c -- CPU restatement of the NVTT 2.1.2 hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Citations are to the reference tree (castano/nvidia-texture-tools 2.1.2).
 * Parity pins for each part are listed in the test files that exercise it
 * (tests/test_oracle_*.py); the functions below are written from the
 * algorithm descriptions, not copied.
 *
 * Build: gcc -O2 -std=c99 -ffp-contract=off (no -ffast-math, no -march=native).
 */
#include "ntx_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } v3;

/* =========================================================================
 * Tables
 * ========================================================================= */

typedef struct { uint8_t c0, c1, c2; } cl4;
typedef struct { uint8_t c0, c1; } cl3;

static cl4 g_four[968 + 8];
static int g_four_total[16];
static cl3 g_three[152 + 8];
static int g_three_total[16];
static uint8_t g_match5[256][2], g_match6[256][2];   /* icbc s_match5/6 (D3D10 decoder) */
static uint8_t g_omatch5[256][2], g_omatch6[256][2]; /* nvtt OMatch5/6 */
static uint8_t g_omatcha5[256][2], g_omatcha6[256][2]; /* nvtt OMatchAlpha5/6 */
static float g_mid5[32], g_mid6[64];
static int g_inited = 0;

/* icbc.h:1906-2012 init_cluster_tables.  Level t adds exactly the cumulative
 * partitions whose last boundary equals t (every partition with a smaller last
 * boundary was added at an earlier level), in (c0, c1) ascending order. */
static void build_cluster_tables(void)
{
    int i = 0;
    for (int t = 1; t <= 16; t++) {
        for (int c0 = 0; c0 <= t; c0++)
            for (int c1 = 0; c1 <= t - c0; c1++) {
                g_four[i].c0 = (uint8_t)c0;
                g_four[i].c1 = (uint8_t)(c0 + c1);
                g_four[i].c2 = (uint8_t)t;
                i++;
            }
        g_four_total[t - 1] = i;
    }
    for (int k = 0; k < 8; k++) g_four[968 + k] = g_four[967];
    i = 0;
    for (int t = 1; t <= 16; t++) {
        for (int c0 = 0; c0 <= t; c0++) {
            g_three[i].c0 = (uint8_t)c0;
            g_three[i].c1 = (uint8_t)t;
            i++;
        }
        g_three_total[t - 1] = i;
    }
    for (int k = 0; k < 8; k++) g_three[152 + k] = g_three[151];
}

/* icbc.h:3166-3262 PrepareOptTable5/6 with Decoder_D3D10: the error of an
 * endpoint pair is the worse of the AMD and NVIDIA interpolation errors. */
static void build_icbc_match(void)
{
    for (int i = 0; i < 256; i++) {
        int best = 256 * 100;
        for (int mn = 0; mn < 32; mn++)
            for (int mx = 0; mx < 32; mx++) {
                int mine = (mn << 3) | (mn >> 2), maxe = (mx << 3) | (mx >> 2);
                int amd = abs(((43 * maxe + 21 * mine + 32) >> 6) - i);
                int nv = abs((((2 * mx + mn) * 22) / 8) - i);
                int err = amd > nv ? amd : nv;
                if (err < best) { best = err; g_match5[i][0] = (uint8_t)mx; g_match5[i][1] = (uint8_t)mn; }
            }
    }
    for (int i = 0; i < 256; i++) {
        int best = 256 * 100;
        for (int mn = 0; mn < 64; mn++)
            for (int mx = 0; mx < 64; mx++) {
                int mine = (mn << 2) | (mn >> 4), maxe = (mx << 2) | (mx >> 4);
                int amd = abs(((43 * maxe + 21 * mine + 32) >> 6) - i);
                int d = maxe - mine; /* C division truncates toward zero, as in the reference */
                int nv = abs(((256 * mine + d / 4 + 128 + d * 80) / 256) - i);
                int err = amd > nv ? amd : nv;
                if (err < best) { best = err; g_match6[i][0] = (uint8_t)mx; g_match6[i][1] = (uint8_t)mn; }
            }
    }
}

/* SingleColorLookup.cpp:34-89 PrepareOptTable: err = 100*|lerp - i| + 3*|max-min|,
 * lerp = (2 max + min) / 3, or (max + min) / 2 for the alpha (three-colour) tables. */
static void build_omatch_mode(uint8_t (*tab)[2], int bits, int alpha_mode)
{
    int size = 1 << bits;
    for (int i = 0; i < 256; i++) {
        int best = 256 * 100;
        for (int mn = 0; mn < size; mn++)
            for (int mx = 0; mx < size; mx++) {
                int mine = bits == 5 ? ((mn << 3) | (mn >> 2)) : ((mn << 2) | (mn >> 4));
                int maxe = bits == 5 ? ((mx << 3) | (mx >> 2)) : ((mx << 2) | (mx >> 4));
                int lerp = alpha_mode ? (maxe + mine) / 2 : (maxe * 2 + mine) / 3;
                int err = abs(lerp - i) * 100 + abs(mx - mn) * 3;
                if (err < best) { tab[i][0] = (uint8_t)mx; tab[i][1] = (uint8_t)mn; best = err; }
            }
    }
}

/* icbc.h:1513-1542: 565 bit-expansion midpoints, as the reference tabulates
 * them (six decimals, then to float). */
static void build_midpoints(void)
{
    char buf[64];
    for (int i = 0; i < 31; i++) {
        float f0 = (float)(((i) << 3) | ((i) >> 2)) / 255.0f;
        float f1 = (float)(((i + 1) << 3) | ((i + 1) >> 2)) / 255.0f;
        snprintf(buf, sizeof buf, "%f", (double)((f0 + f1) * 0.5));
        g_mid5[i] = strtof(buf, NULL);
    }
    g_mid5[31] = FLT_MAX;
    for (int i = 0; i < 63; i++) {
        float f0 = (float)(((i) << 2) | ((i) >> 4)) / 255.0f;
        float f1 = (float)(((i + 1) << 2) | ((i + 1) >> 4)) / 255.0f;
        snprintf(buf, sizeof buf, "%f", (double)((f0 + f1) * 0.5));
        g_mid6[i] = strtof(buf, NULL);
    }
    g_mid6[63] = FLT_MAX;
}

void ntxo_init(void)
{
    if (g_inited) return;
    build_cluster_tables();
    build_icbc_match();
    build_omatch_mode(g_omatch5, 5, 0);
    build_omatch_mode(g_omatch6, 6, 0);
    build_omatch_mode(g_omatcha5, 5, 1);
    build_omatch_mode(g_omatcha6, 6, 1);
    build_midpoints();
    g_inited = 1;
}

int ntxo_cluster_table4(uint8_t *out, int *totals)
{
    ntxo_init();
    for (int i = 0; i < 976; i++) { out[3 * i] = g_four[i].c0; out[3 * i + 1] = g_four[i].c1; out[3 * i + 2] = g_four[i].c2; }
    memcpy(totals, g_four_total, sizeof g_four_total);
    return 976;
}
int ntxo_cluster_table3(uint8_t *out, int *totals)
{
    ntxo_init();
    for (int i = 0; i < 160; i++) { out[2 * i] = g_three[i].c0; out[2 * i + 1] = g_three[i].c1; }
    memcpy(totals, g_three_total, sizeof g_three_total);
    return 160;
}
void ntxo_icbc_match_tables(uint8_t *m5, uint8_t *m6) { ntxo_init(); memcpy(m5, g_match5, 512); memcpy(m6, g_match6, 512); }
void ntxo_omatch_tables(uint8_t *m5, uint8_t *m6) { ntxo_init(); memcpy(m5, g_omatch5, 512); memcpy(m6, g_omatch6, 512); }
void ntxo_omatch_alpha_tables(uint8_t *m5, uint8_t *m6) { ntxo_init(); memcpy(m5, g_omatcha5, 512); memcpy(m6, g_omatcha6, 512); }
void ntxo_midpoints(float *m5, float *m6) { ntxo_init(); memcpy(m5, g_mid5, sizeof g_mid5); memcpy(m6, g_mid6, sizeof g_mid6); }

/* =========================================================================
 * BC1: icbc v1.05 Level8, SSE4.1 arithmetic
 * ========================================================================= */

/* icbc::max/min (icbc.h:178-188): max(a,b) = (b<a)?a:b, min(a,b) = (a<b)?a:b. */
static inline float imax(float a, float b) { return (b < a) ? a : b; }
static inline float imin(float a, float b) { return (a < b) ? a : b; }

/* SSE vsaturate (icbc.h:474-478): minps(maxps(a, 0), 1). maxps returns the
 * second operand unless the first is greater; NaN therefore becomes 0. */
static inline float sse_saturate(float a)
{
    float t = (a > 0.0f) ? a : 0.0f;
    return (t < 1.0f) ? t : 1.0f;
}
/* vround5/vround6 without ICBC_PERFECT_ROUND (icbc.h:1544-1578): roundps to nearest even. */
static inline float round5(float x) { return rintf(x * 31.0f) * (1.0f / 31.0f); }
static inline float round6(float x) { return rintf(x * 63.0f) * (1.0f / 63.0f); }

typedef struct { float r[16], g[16], b[16], w[16]; } sat_t;

/* icbc.h:1628-1683 reduce_colors: merge texels whose channels are all within
 * the threshold of an existing colour (first match, running weighted mean). */
static int reduce_colors(const float *in, const float *inw, v3 *col, float *w, int *any_black)
{
    const float thr = 1.0f / 256;
    int n = 0;
    *any_black = 0;
    for (int i = 0; i < 16; i++) {
        v3 ci = { in[4 * i + 0], in[4 * i + 1], in[4 * i + 2] };
        float wi = inw[i];
        if (wi > 0) {
            int j;
            for (j = 0; j < n; j++) {
                if (fabsf(col[j].x - ci.x) < thr && fabsf(col[j].y - ci.y) < thr && fabsf(col[j].z - ci.z) < thr) {
                    float s = w[j] + wi;
                    v3 m = { col[j].x * w[j] + ci.x * wi, col[j].y * w[j] + ci.y * wi, col[j].z * w[j] + ci.z * wi };
                    col[j].x = m.x / s; col[j].y = m.y / s; col[j].z = m.z / s;
                    w[j] += wi;
                    break;
                }
            }
            if (j == n) { col[n] = ci; w[n] = wi; n++; }
            if (ci.x < 1.0f / 8 && ci.y < 1.0f / 8 && ci.z < 1.0f / 8) *any_black = 1;
        }
    }
    return n;
}

static inline int is_black(v3 c) { return c.x < 1.0f / 8 && c.y < 1.0f / 8 && c.z < 1.0f / 8; }

/* icbc.h:1685-1703 skip_blacks. */
static int skip_blacks(const v3 *col, const float *w, int count, v3 *oc, float *ow)
{
    int n = 0;
    for (int i = 0; i < count; i++) {
        if (is_black(col[i])) continue;
        oc[n] = col[i]; ow[n] = w[i]; n++;
    }
    return n;
}

/* icbc.h:1710-1806: weighted centroid, covariance, power iteration (8 steps,
 * normalised by the largest component via reciprocal multiply). */
static v3 principal_component(int n, const v3 *p, const float *w)
{
    v3 c = { 0, 0, 0 };
    float total = 0.0f;
    for (int i = 0; i < n; i++) {
        total += w[i];
        c.x += p[i].x * w[i]; c.y += p[i].y * w[i]; c.z += p[i].z * w[i];
    }
    float inv = 1.0f / total;
    c.x *= inv; c.y *= inv; c.z *= inv;
    float m[6] = { 0, 0, 0, 0, 0, 0 };
    for (int i = 0; i < n; i++) {
        v3 a = { p[i].x - c.x, p[i].y - c.y, p[i].z - c.z };
        v3 b = { a.x * w[i], a.y * w[i], a.z * w[i] };
        m[0] += a.x * b.x; m[1] += a.x * b.y; m[2] += a.x * b.z;
        m[3] += a.y * b.y; m[4] += a.y * b.z; m[5] += a.z * b.z;
    }
    v3 v = { 0, 0, 0 };
    if (m[0] == 0 && m[3] == 0 && m[5] == 0) return v;
    float r0 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    float r1 = m[1] * m[1] + m[3] * m[3] + m[4] * m[4];
    float r2 = m[2] * m[2] + m[4] * m[4] + m[5] * m[5];
    if (r0 > r1 && r0 > r2) { v.x = m[0]; v.y = m[1]; v.z = m[2]; }
    else if (r1 > r2)       { v.x = m[1]; v.y = m[3]; v.z = m[4]; }
    else                    { v.x = m[2]; v.y = m[4]; v.z = m[5]; }
    for (int it = 0; it < 8; it++) {
        float x = v.x * m[0] + v.y * m[1] + v.z * m[2];
        float y = v.x * m[1] + v.y * m[3] + v.z * m[4];
        float z = v.x * m[2] + v.y * m[4] + v.z * m[5];
        float norm = imax(imax(x, y), z);
        float iv = 1.0f / norm;
        v.x = x * iv; v.y = y * iv; v.z = z * iv;
    }
    return v;
}

/* icbc.h:1812-1882 compute_sat: sort by projection (stable insertion sort),
 * weighted prefix sums, FLT_MAX past the end. */
static int compute_sat(const v3 *col, const float *w, int count, sat_t *s)
{
    v3 pc = principal_component(count, col, w);
    int order[16];
    float dps[16];
    for (int i = 0; i < count; i++) {
        order[i] = i;
        dps[i] = col[i].x * pc.x + col[i].y * pc.y + col[i].z * pc.z;
    }
    for (int i = 0; i < count; i++)
        for (int j = i; j > 0 && dps[j] < dps[j - 1]; j--) {
            float t = dps[j]; dps[j] = dps[j - 1]; dps[j - 1] = t;
            int o = order[j]; order[j] = order[j - 1]; order[j - 1] = o;
        }
    float wt = w[order[0]];
    s->r[0] = col[order[0]].x * wt; s->g[0] = col[order[0]].y * wt; s->b[0] = col[order[0]].z * wt; s->w[0] = wt;
    for (int i = 1; i < count; i++) {
        wt = w[order[i]];
        s->r[i] = s->r[i - 1] + col[order[i]].x * wt;
        s->g[i] = s->g[i - 1] + col[order[i]].y * wt;
        s->b[i] = s->b[i - 1] + col[order[i]].z * wt;
        s->w[i] = s->w[i - 1] + wt;
    }
    for (int i = count; i < 16; i++) s->r[i] = s->g[i] = s->b[i] = s->w[i] = FLT_MAX;
    return count;
}

static inline void sat_fetch(const sat_t *s, int c, v3 *x, float *w)
{
    if (c >= 0) { x->x = s->r[c]; x->y = s->g[c]; x->z = s->b[c]; *w = s->w[c]; }
    else { x->x = x->y = x->z = 0.0f; *w = 0.0f; }
}

/* Shared tail of icbc.h:2201-2236 / 2501-2538: least-squares endpoints, snap to
 * the 565 grid, error of the snapped endpoints (constant term dropped). */
static inline float ls_error(v3 ax, v3 sum, float alpha2, float beta2, float ab, float factor, v3 msq, v3 *pa, v3 *pb)
{
    v3 bx = { sum.x - ax.x, sum.y - ax.y, sum.z - ax.z };
    v3 a = { (ax.x * beta2 - bx.x * ab) * factor, (ax.y * beta2 - bx.y * ab) * factor, (ax.z * beta2 - bx.z * ab) * factor };
    v3 b = { (bx.x * alpha2 - ax.x * ab) * factor, (bx.y * alpha2 - ax.y * ab) * factor, (bx.z * alpha2 - ax.z * ab) * factor };
    a.x = round5(sse_saturate(a.x)); a.y = round6(sse_saturate(a.y)); a.z = round5(sse_saturate(a.z));
    b.x = round5(sse_saturate(b.x)); b.y = round6(sse_saturate(b.y)); b.z = round5(sse_saturate(b.z));
    v3 e2 = { (a.x * (b.x * ab - ax.x) - b.x * bx.x) * 2.0f,
              (a.y * (b.y * ab - ax.y) - b.y * bx.y) * 2.0f,
              (a.z * (b.z * ab - ax.z) - b.z * bx.z) * 2.0f };
    v3 e1 = { (a.x * a.x) * alpha2 + ((b.x * b.x) * beta2 + e2.x),
              (a.y * a.y) * alpha2 + ((b.y * b.y) * beta2 + e2.y),
              (a.z * a.z) * alpha2 + ((b.z * b.z) * beta2 + e2.z) };
    *pa = a; *pb = b;
    return e1.x * msq.x + e1.y * msq.y + e1.z * msq.z;
}

/* icbc.h:2250-2549 cluster_fit_four (generic lane-load path of the SSE build). */
static float fit4_candidate(const sat_t *s, int count, int p, v3 msq, v3 *pa, v3 *pb)
{
    v3 sum = { s->r[count - 1], s->g[count - 1], s->b[count - 1] };
    float wsum = s->w[count - 1];
    v3 x0, x1, x2; float w0, w1, w2;
    sat_fetch(s, g_four[p].c0 - 1, &x0, &w0);
    sat_fetch(s, g_four[p].c1 - 1, &x1, &w1);
    sat_fetch(s, g_four[p].c2 - 1, &x2, &w2);
    float w3 = wsum - w2;
    x2.x = x2.x - x1.x; x2.y = x2.y - x1.y; x2.z = x2.z - x1.z;
    x1.x = x1.x - x0.x; x1.y = x1.y - x0.y; x1.z = x1.z - x0.z;
    w2 = w2 - w1;
    w1 = w1 - w0;
    float alpha2 = w2 * (1.0f / 9.0f) + (w1 * (4.0f / 9.0f) + w0);
    float beta2 = w1 * (1.0f / 9.0f) + (w2 * (4.0f / 9.0f) + w3);
    float ab = (w1 + w2) * (2.0f / 9.0f);
    float factor = 1.0f / (alpha2 * beta2 - ab * ab);
    v3 ax = { x2.x * (1.0f / 3.0f) + (x1.x * (2.0f / 3.0f) + x0.x),
              x2.y * (1.0f / 3.0f) + (x1.y * (2.0f / 3.0f) + x0.y),
              x2.z * (1.0f / 3.0f) + (x1.z * (2.0f / 3.0f) + x0.z) };
    return ls_error(ax, sum, alpha2, beta2, ab, factor, msq, pa, pb);
}

/* icbc.h:2016-2247 cluster_fit_three. */
static float fit3_candidate(const sat_t *s, int count, int p, v3 msq, v3 *pa, v3 *pb)
{
    v3 sum = { s->r[count - 1], s->g[count - 1], s->b[count - 1] };
    float wsum = s->w[count - 1];
    v3 x0, x1; float w0, w1;
    sat_fetch(s, g_three[p].c0 - 1, &x0, &w0);
    sat_fetch(s, g_three[p].c1 - 1, &x1, &w1);
    float w2 = wsum - w1;
    x1.x = x1.x - x0.x; x1.y = x1.y - x0.y; x1.z = x1.z - x0.z;
    w1 = w1 - w0;
    float ab = w1 * 0.25f;
    float alpha2 = w0 + ab;
    float beta2 = w2 + ab;
    float factor = 1.0f / (alpha2 * beta2 - ab * ab);
    v3 ax = { x0.x + x1.x * 0.5f, x0.y + x1.y * 0.5f, x0.z + x1.z * 0.5f };
    return ls_error(ax, sum, alpha2, beta2, ab, factor, msq, pa, pb);
}

/* The SSE build evaluates partitions four at a time; each lane keeps its own
 * best with a strict '<' and reduce_min_index (icbc.h:533-545) picks the lowest
 * lane holding the minimum.  Partitions past the level total are evaluated too
 * (they read FLT_MAX entries of the SAT). */
static void cluster_fit(int four, const sat_t *s, int count, v3 msq, v3 *start, v3 *end)
{
    const int total = four ? g_four_total[count - 1] : g_three_total[count - 1];
    float best[4] = { FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX };
    v3 bs[4], be[4];
    memset(bs, 0, sizeof bs);
    memset(be, 0, sizeof be);
    for (int i = 0; i < total; i += 4)
        for (int l = 0; l < 4; l++) {
            v3 a, b;
            float e = four ? fit4_candidate(s, count, i + l, msq, &a, &b) : fit3_candidate(s, count, i + l, msq, &a, &b);
            if (e < best[l]) { best[l] = e; bs[l] = a; be[l] = b; }
        }
    float m01 = (best[0] < best[1]) ? best[0] : best[1];
    float m23 = (best[2] < best[3]) ? best[2] : best[3];
    float mn = (m01 < m23) ? m01 : m23;
    int li = 0;
    while (li < 3 && !(best[li] <= mn)) li++;
    *start = bs[li];
    *end = be[li];
}

/* icbc.h:1580-1595 vector3_to_color16: truncate, then round exactly on the
 * 565 bit-expansion midpoints. */
static uint16_t v3_to_565(v3 v)
{
    unsigned r = (unsigned)imin(imax(v.x * 31.0f, 0.0f), 31.0f);
    unsigned g = (unsigned)imin(imax(v.y * 63.0f, 0.0f), 63.0f);
    unsigned b = (unsigned)imin(imax(v.z * 31.0f, 0.0f), 31.0f);
    r += (v.x > g_mid5[r]);
    g += (v.y > g_mid6[g]);
    b += (v.z > g_mid5[b]);
    return (uint16_t)((r << 11) | (g << 5) | b);
}

static void expand565(uint16_t c, int *r, int *g, int *b)
{
    int r5 = (c >> 11) & 31, g6 = (c >> 5) & 63, b5 = c & 31;
    *r = (r5 << 3) | (r5 >> 2);
    *g = (g6 << 2) | (g6 >> 4);
    *b = (b5 << 3) | (b5 >> 2);
}

/* icbc.h:2557-2585 evaluate_palette_d3d10, as 8-bit integers (rgb per entry). */
static void palette_d3d10(uint16_t c0, uint16_t c1, int pal[4][3])
{
    expand565(c0, &pal[0][0], &pal[0][1], &pal[0][2]);
    expand565(c1, &pal[1][0], &pal[1][1], &pal[1][2]);
    for (int k = 0; k < 3; k++) {
        if (c0 > c1) {
            pal[2][k] = (2 * pal[0][k] + pal[1][k]) / 3;
            pal[3][k] = (2 * pal[1][k] + pal[0][k]) / 3;
        } else {
            pal[2][k] = (pal[0][k] + pal[1][k]) / 2;
            pal[3][k] = 0;
        }
    }
}

static void palette_float(uint16_t c0, uint16_t c1, v3 pf[4])
{
    int pal[4][3];
    palette_d3d10(c0, c1, pal);
    for (int i = 0; i < 4; i++) { pf[i].x = pal[i][0] / 255.0f; pf[i].y = pal[i][1] / 255.0f; pf[i].z = pal[i][2] / 255.0f; }
}

static inline float dist2(v3 a, v3 b) { float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z; return dx * dx + dy * dy + dz * dz; }

/* icbc.h:2803-2825 compute_indices4 (index = x1 | x0 << 1). */
static uint32_t indices4(const float *in, v3 cw, const v3 pal[4])
{
    v3 p[4];
    for (int k = 0; k < 4; k++) { p[k].x = pal[k].x * cw.x; p[k].y = pal[k].y * cw.y; p[k].z = pal[k].z * cw.z; }
    uint32_t idx = 0;
    for (int i = 0; i < 16; i++) {
        v3 c = { in[4 * i] * cw.x, in[4 * i + 1] * cw.y, in[4 * i + 2] * cw.z };
        float d0 = dist2(c, p[0]), d1 = dist2(c, p[1]), d2 = dist2(c, p[2]), d3 = dist2(c, p[3]);
        int x0 = ((d1 > d2) && (d0 > d2)) || ((d0 > d3) && (d1 > d3));
        int x1 = (d0 > d3) && (d2 > d3);
        idx |= (uint32_t)(x1 | (x0 << 1)) << (2 * i);
    }
    return idx;
}

/* icbc.h:2827-2882 compute_indices3 with allow_transparent_black. */
static uint32_t indices3(const float *in, v3 cw, const v3 pal[4])
{
    v3 p[3];
    for (int k = 0; k < 3; k++) { p[k].x = pal[k].x * cw.x; p[k].y = pal[k].y * cw.y; p[k].z = pal[k].z * cw.z; }
    uint32_t idx = 0;
    for (int i = 0; i < 16; i++) {
        v3 c = { in[4 * i] * cw.x, in[4 * i + 1] * cw.y, in[4 * i + 2] * cw.z };
        float d0 = dist2(p[0], c), d1 = dist2(p[1], c), d2 = dist2(p[2], c);
        float d3 = c.x * c.x + c.y * c.y + c.z * c.z;
        int i1 = d1 < d2;
        int i2 = (d2 <= d0) && (d2 <= d1);
        int i3 = (d3 <= d0) && (d3 <= d1) && (d3 <= d2);
        idx |= (uint32_t)((i1 | i3) | ((i2 | i3) << 1)) << (2 * i);
    }
    return idx;
}

/* icbc.h:2687-2736 evaluate_mse over the 16 input texels (0..255 units). */
static float block_mse(const float *in, const float *w, v3 cw, const v3 pal[4], uint32_t idx)
{
    float err = 0.0f;
    for (int i = 0; i < 16; i++) {
        v3 p = pal[(idx >> (2 * i)) & 3];
        v3 d = { ((p.x - in[4 * i]) * cw.x) * 255.0f, ((p.y - in[4 * i + 1]) * cw.y) * 255.0f, ((p.z - in[4 * i + 2]) * cw.z) * 255.0f };
        err += w[i] * (d.x * d.x + d.y * d.y + d.z * d.z);
    }
    return err;
}

static void put_block(uint8_t *out, uint16_t c0, uint16_t c1, uint32_t idx)
{
    out[0] = (uint8_t)c0; out[1] = (uint8_t)(c0 >> 8);
    out[2] = (uint8_t)c1; out[3] = (uint8_t)(c1 >> 8);
    out[4] = (uint8_t)idx; out[5] = (uint8_t)(idx >> 8); out[6] = (uint8_t)(idx >> 16); out[7] = (uint8_t)(idx >> 24);
}

/* icbc.h:2963-2980 output_block4 / 2944-2961 output_block3. */
static float output_block(int four, const float *in, const float *w, v3 cw, v3 s, v3 e, uint8_t *out)
{
    uint16_t c0 = v3_to_565(s), c1 = v3_to_565(e);
    if (four ? (c0 < c1) : (c0 > c1)) { uint16_t t = c0; c0 = c1; c1 = t; }
    v3 pal[4];
    palette_float(c0, c1, pal);
    uint32_t idx = four ? indices4(in, cw, pal) : indices3(in, cw, pal);
    put_block(out, c0, c1, idx);
    return block_mse(in, w, cw, pal, idx);
}

/* icbc.h:3274-3289 compress_dxt1_single_color_optimal. */
static void single_color(v3 c, uint8_t *out)
{
    int r = (int)(uint8_t)(sse_saturate(c.x) * 255 + 0.5f);
    int g = (int)(uint8_t)(sse_saturate(c.y) * 255 + 0.5f);
    int b = (int)(uint8_t)(sse_saturate(c.z) * 255 + 0.5f);
    uint16_t c0 = (uint16_t)((g_match5[r][0] << 11) | (g_match6[g][0] << 5) | g_match5[b][0]);
    uint16_t c1 = (uint16_t)((g_match5[r][1] << 11) | (g_match6[g][1] << 5) | g_match5[b][1]);
    uint32_t idx = 0xaaaaaaaau;
    if (c0 < c1) { uint16_t t = c0; c0 = c1; c1 = t; idx ^= 0x55555555u; }
    put_block(out, c0, c1, idx);
}

/* icbc.h:2885-2930 compute_indices (SSE form): nearest of the four palette
 * entries with the reference's comparison order. */
static uint32_t indices_any(const float *in, v3 cw, const v3 pal[4])
{
    v3 p[4];
    for (int k = 0; k < 4; k++) { p[k].x = pal[k].x * cw.x; p[k].y = pal[k].y * cw.y; p[k].z = pal[k].z * cw.z; }
    uint32_t idx = 0;
    for (int i = 0; i < 16; i++) {
        v3 c = { in[4 * i] * cw.x, in[4 * i + 1] * cw.y, in[4 * i + 2] * cw.z };
        float d0 = dist2(c, p[0]), d1 = dist2(c, p[1]), d2 = dist2(c, p[2]), d3 = dist2(c, p[3]);
        int i1 = (d1 <= d0) && (d1 < d2) && (d1 < d3);
        int i2 = (d2 <= d0) && (d2 <= d1) && (d2 < d3);
        int i3 = (d3 <= d0) && (d3 <= d1) && (d3 <= d2);
        idx |= (uint32_t)((i1 | i3) | ((i2 | i3) << 1)) << (2 * i);
    }
    return idx;
}

/* icbc.h:3329-3406 refine_endpoints with three_color_mode = true: step one
 * endpoint's 5:6:5 bit-fields (wrapping) by deltas[i % 16], endpoint 1 for
 * i / 16 even and endpoint 0 for odd; the candidate's indices are those of
 * the current block's palette; keep it when its evaluate_mse is lower; stop
 * 33 steps after the last improvement. */
static float refine_endpoints(const float *in, const float *w, v3 cw, float input_error, uint8_t *out)
{
    static const int deltas[16][3] = {
        { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 }, { -1, 0, 0 }, { 0, -1, 0 }, { 0, 0, -1 },
        { 1, 1, 0 }, { 1, 0, 1 }, { 0, 1, 1 }, { -1, -1, 0 }, { -1, 0, -1 }, { 0, -1, -1 },
        { -1, 1, 0 }, { 1, -1, 0 }, { 0, -1, 1 }, { 0, 1, -1 },
    };
    uint16_t c0 = (uint16_t)(out[0] | (out[1] << 8)), c1 = (uint16_t)(out[2] | (out[3] << 8));
    uint32_t idx = (uint32_t)out[4] | ((uint32_t)out[5] << 8) | ((uint32_t)out[6] << 16) | ((uint32_t)out[7] << 24);
    float best = input_error;
    int last = 0;
    for (int i = 0; i < 256; i++) {
        const int *d = deltas[i % 16];
        uint16_t r0 = c0, r1 = c1;
        uint16_t *t = ((i / 16) & 1) ? &r0 : &r1;
        int r = ((*t >> 11) + d[0]) & 31, g = (((*t >> 5) & 63) + d[1]) & 63, b = ((*t & 31) + d[2]) & 31;
        *t = (uint16_t)((r << 11) | (g << 5) | b);
        v3 pal[4], rpal[4];
        palette_float(c0, c1, pal);
        uint32_t ridx = indices_any(in, cw, pal);
        palette_float(r0, r1, rpal);
        float e = block_mse(in, w, cw, rpal, ridx);
        if (e < best) { best = e; c0 = r0; c1 = r1; idx = ridx; last = i; }
        if (i - last > 32) break;
    }
    put_block(out, c0, c1, idx);
    return best;
}

/* BlockCompressor.cpp:211-217 qualityLevel: Fastest -> Level1, Production ->
 * Level9, otherwise Level8. */
static int bc1_level(int quality)
{
    return quality == NTXO_QUALITY_FASTEST ? 1 : quality == NTXO_QUALITY_PRODUCTION ? 9 : 8;
}

/* icbc.h:191-193, 291-293 clamp / saturate on the scalar path. */
static inline float isat(float x) { return imin(imax(x, 0.0f), 1.0f); }

/* compress_dxt1 at Quality_Level1 (icbc.h:3485-3540, box_fit and
 * least_squares_fit): fit_colors_bbox, inset_bbox, select_diagonal
 * (icbc.h:3096-3153), output_block4, optimize_end_points4 (icbc.h:2984-3018)
 * and output_block4 again, keeping the lower error.  All 16 input colours. */
static float bc1_box_fit(const float *in, const float *inw, v3 cw, uint8_t *out)
{
    v3 c0 = { 0, 0, 0 }, c1 = { 1, 1, 1 };
    for (int i = 0; i < 16; i++) {
        c0.x = imax(c0.x, in[4 * i]); c0.y = imax(c0.y, in[4 * i + 1]); c0.z = imax(c0.z, in[4 * i + 2]);
        c1.x = imin(c1.x, in[4 * i]); c1.y = imin(c1.y, in[4 * i + 1]); c1.z = imin(c1.z, in[4 * i + 2]);
    }
    const float bias = (8.0f / 255.0f) / 16.0f;
    v3 ins = { (c0.x - c1.x) / 16.0f - bias, (c0.y - c1.y) / 16.0f - bias, (c0.z - c1.z) / 16.0f - bias };
    c0.x = isat(c0.x - ins.x); c0.y = isat(c0.y - ins.y); c0.z = isat(c0.z - ins.z);
    c1.x = isat(c1.x + ins.x); c1.y = isat(c1.y + ins.y); c1.z = isat(c1.z + ins.z);
    v3 ctr = { (c0.x + c1.x) * 0.5f, (c0.y + c1.y) * 0.5f, (c0.z + c1.z) * 0.5f };
    float cov_xz = 0.0f, cov_yz = 0.0f;
    for (int i = 0; i < 16; i++) {
        v3 t = { in[4 * i] - ctr.x, in[4 * i + 1] - ctr.y, in[4 * i + 2] - ctr.z };
        cov_xz += t.x * t.z;
        cov_yz += t.y * t.z;
    }
    if (cov_xz < 0) { float t = c0.x; c0.x = c1.x; c1.x = t; }
    if (cov_yz < 0) { float t = c0.y; c0.y = c1.y; c1.y = t; }
    float err = output_block(1, in, inw, cw, c0, c1, out);
    uint32_t idx = (uint32_t)out[4] | ((uint32_t)out[5] << 8) | ((uint32_t)out[6] << 16) | ((uint32_t)out[7] << 24);
    float a2 = 0.0f, b2 = 0.0f, ab = 0.0f;
    v3 ax = { 0, 0, 0 }, bx = { 0, 0, 0 };
    for (int i = 0; i < 16; i++) {
        const uint32_t bits = idx >> (2 * i);
        float beta = (float)(bits & 1);
        if (bits & 2) beta = (1 + beta) / 3.0f;
        float alpha = 1.0f - beta;
        a2 += alpha * alpha; b2 += beta * beta; ab += alpha * beta;
        ax.x += alpha * in[4 * i]; ax.y += alpha * in[4 * i + 1]; ax.z += alpha * in[4 * i + 2];
        bx.x += beta * in[4 * i]; bx.y += beta * in[4 * i + 1]; bx.z += beta * in[4 * i + 2];
    }
    float denom = a2 * b2 - ab * ab;
    if (fabsf(denom - 0.0f) < 0.0001f) return err; /* icbc equal(denom, 0) */
    float factor = 1.0f / denom;
    v3 a = { isat((ax.x * b2 - bx.x * ab) * factor), isat((ax.y * b2 - bx.y * ab) * factor),
             isat((ax.z * b2 - bx.z * ab) * factor) };
    v3 b = { isat((bx.x * a2 - ax.x * ab) * factor), isat((bx.y * a2 - ax.y * ab) * factor),
             isat((bx.z * a2 - ax.z * ab) * factor) };
    uint8_t blk2[8];
    float err2 = output_block(1, in, inw, cw, a, b, blk2);
    if (err2 < err) { err = err2; memcpy(out, blk2, 8); }
    return err;
}

/* icbc.h:3485-3559 compress_dxt1 with three_color_mode and three_color_black
 * enabled (BlockCompressor.cpp:211-222) at Quality_Level8 (level 8, nvtt's
 * Normal and Highest), Quality_Level9 (level 9, Production) or
 * Quality_Level1 (level 1, Fastest). */
static float bc1_block_level(const float *in, const float *inw, const float *cwp, int level, uint8_t *out, int *n_out,
                             int *m_out)
{
    ntxo_init();
    v3 cw = { cwp[0], cwp[1], cwp[2] };
    if (level == 1) {
        if (n_out) *n_out = 16;
        if (m_out) *m_out = -1;
        return bc1_box_fit(in, inw, cw, out);
    }
    v3 col[16];
    float w[16];
    int any_black = 0;
    int count = reduce_colors(in, inw, col, w, &any_black);
    if (n_out) *n_out = count;
    if (m_out) *m_out = -1;
    if (count == 0) { memset(out, 0, 8); return 0.0f; }
    if (count == 1) {
        single_color(col[0], out);
        v3 pal[4];
        uint16_t c0 = (uint16_t)(out[0] | (out[1] << 8)), c1 = (uint16_t)(out[2] | (out[3] << 8));
        uint32_t idx = (uint32_t)out[4] | ((uint32_t)out[5] << 8) | ((uint32_t)out[6] << 16) | ((uint32_t)out[7] << 24);
        palette_float(c0, c1, pal);
        return block_mse(in, inw, cw, pal, idx);
    }
    v3 msq = { cw.x * cw.x, cw.y * cw.y, cw.z * cw.z };
    sat_t sat;
    int sat_count = compute_sat(col, w, count, &sat);
    v3 s, e;
    cluster_fit(1, &sat, sat_count, msq, &s, &e);
    uint8_t best_blk[8], blk3[8];
    float best = output_block(1, in, inw, cw, s, e, best_blk);
    int run3 = 1;
    if (any_black) {
        v3 tc[16];
        float tw[16];
        int tn = skip_blacks(col, w, count, tc, tw);
        if (!tn) run3 = 0;
        else sat_count = compute_sat(tc, tw, tn, &sat);
    }
    if (run3) {
        if (m_out) *m_out = sat_count;
        cluster_fit(0, &sat, sat_count, msq, &s, &e);
        float e3 = output_block(0, in, inw, cw, s, e, blk3);
        if (e3 < best) { best = e3; memcpy(best_blk, blk3, 8); }
    }
    if (best < FLT_MAX) memcpy(out, best_blk, 8);
    else { memset(out, 0, 8); return best; } /* the reference leaves the block unwritten here */
    if (level == 9) best = refine_endpoints(in, inw, cw, best, out);
    return best;
}

float ntxo_bc1_block(const float *in, const float *inw, const float *cwp, uint8_t *out, int *n_out, int *m_out)
{
    return bc1_block_level(in, inw, cwp, 8, out, n_out, m_out);
}

float ntxo_bc1_block_q(const float *in, const float *inw, const float *cwp, int quality, uint8_t *out)
{
    return bc1_block_level(in, inw, cwp, bc1_level(quality), out, NULL, NULL);
}

void ntxo_bc1_decode(const uint8_t *blk, uint8_t *rgba)
{
    uint16_t c0 = (uint16_t)(blk[0] | (blk[1] << 8)), c1 = (uint16_t)(blk[2] | (blk[3] << 8));
    uint32_t idx = (uint32_t)blk[4] | ((uint32_t)blk[5] << 8) | ((uint32_t)blk[6] << 16) | ((uint32_t)blk[7] << 24);
    int pal[4][3];
    palette_d3d10(c0, c1, pal);
    for (int i = 0; i < 16; i++) {
        int k = (idx >> (2 * i)) & 3;
        rgba[4 * i] = (uint8_t)pal[k][0]; rgba[4 * i + 1] = (uint8_t)pal[k][1]; rgba[4 * i + 2] = (uint8_t)pal[k][2];
        rgba[4 * i + 3] = (c0 <= c1 && k == 3) ? 0 : 255;
    }
}

/* =========================================================================
 * BC3 colour: nvsquish WeightedClusterFit, scalar build (SQUISH_USE_SSE=0)
 * ========================================================================= */

/* std::min / std::max semantics used by squish (maths.h Min/Max). */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }

/* colourblock.cpp:30-54 FloatToInt / FloatTo565. */
static int float_to_int(float a, int limit)
{
    int i = (int)(a + 0.5f);
    if (i < 0) i = 0; else if (i > limit) i = limit;
    return i;
}
static int float_to_565(v3 c)
{
    return (float_to_int(31.0f * c.x, 31) << 11) | (float_to_int(63.0f * c.y, 63) << 5) | float_to_int(31.0f * c.z, 31);
}

static void squish_fit(const uint8_t *bgra, int weight_by_alpha, const float *cwp, int dxt1, uint8_t *out);

void ntxo_bc3_color_block(const uint8_t *bgra, int weight_by_alpha, const float *cwp, uint8_t *out)
{
    ntxo_init();
    /* CompressorDX9.cpp:160-163: single RGB colour -> OptimalCompress::compressDXT1. */
    int single = 1;
    for (int i = 1; i < 16 && single; i++)
        single = bgra[4 * i] == bgra[0] && bgra[4 * i + 1] == bgra[1] && bgra[4 * i + 2] == bgra[2];
    if (single) {
        /* OptimalCompressDXT.cpp:254-269 */
        int r = bgra[2], g = bgra[1], b = bgra[0];
        uint16_t c0 = (uint16_t)((g_omatch5[r][0] << 11) | (g_omatch6[g][0] << 5) | g_omatch5[b][0]);
        uint16_t c1 = (uint16_t)((g_omatch5[r][1] << 11) | (g_omatch6[g][1] << 5) | g_omatch5[b][1]);
        uint32_t idx = 0xaaaaaaaau;
        if (c0 < c1) { uint16_t t = c0; c0 = c1; c1 = t; idx ^= 0x55555555u; }
        put_block(out, c0, c1, idx);
        return;
    }
    squish_fit(bgra, weight_by_alpha, cwp, 0, out);
}

/* nvsquish WeightedClusterFit over ColourSet (colourfit.cpp:41-57): dxt1 set
 * (kDxt1): texels of alpha 0 leave the set and take index 3, Compress3 runs,
 * then Compress4 unless a texel was transparent; a later fit replaces the
 * block only with a strictly lower error. */
static void squish_fit(const uint8_t *bgra, int weight_by_alpha, const float *cwp, int dxt1, uint8_t *out)
{
    v3 metric = { cwp[0], cwp[1], cwp[2] };
    v3 msq = { metric.x * metric.x, metric.y * metric.y, metric.z * metric.z };
    int transparent = 0;
    /* colourset.cpp:35-101 with the header default createMinimalSet=true
     * (colourset.h:40): texels with identical RGB bytes share one point whose
     * weight accumulates in texel order; remap[] sends texels to points. */
    v3 pts[16];
    float wts[16];
    int remap[16];
    int count = 0;
    for (int i = 0; i < 16; i++) {
        float wa = (float)(bgra[4 * i + 3] + 1) / 256.0f;
        if (dxt1 && bgra[4 * i + 3] == 0) { remap[i] = -1; transparent = 1; continue; }
        int j;
        for (j = 0; j < i; j++)
            if (bgra[4 * i] == bgra[4 * j] && bgra[4 * i + 1] == bgra[4 * j + 1] && bgra[4 * i + 2] == bgra[4 * j + 2] &&
                (bgra[4 * j + 3] != 0 || !dxt1)) break;
        if (j == i) {
            pts[count].x = (float)bgra[4 * i + 2] / 255.0f;
            pts[count].y = (float)bgra[4 * i + 1] / 255.0f;
            pts[count].z = (float)bgra[4 * i + 0] / 255.0f;
            wts[count] = weight_by_alpha ? wa : 1.0f;
            remap[i] = count++;
        } else {
            wts[remap[j]] += weight_by_alpha ? wa : 1.0f;
            remap[i] = remap[j];
        }
    }
    /* maths.cpp:32-59 ComputeWeightedCovariance (metric applied to deltas). */
    float total = 0.0f;
    v3 c = { 0, 0, 0 };
    for (int i = 0; i < count; i++) {
        total += wts[i];
        c.x += pts[i].x * wts[i]; c.y += pts[i].y * wts[i]; c.z += pts[i].z * wts[i];
    }
    float t = 1.0f / total;
    c.x *= t; c.y *= t; c.z *= t;
    float m[6] = { 0, 0, 0, 0, 0, 0 };
    for (int i = 0; i < count; i++) {
        v3 a = { (pts[i].x - c.x) * metric.x, (pts[i].y - c.y) * metric.y, (pts[i].z - c.z) * metric.z };
        v3 b = { a.x * wts[i], a.y * wts[i], a.z * wts[i] };
        m[0] += a.x * b.x; m[1] += a.x * b.y; m[2] += a.x * b.z;
        m[3] += a.y * b.y; m[4] += a.y * b.z; m[5] += a.z * b.z;
    }
    /* maths.cpp:62-78, 109-131 principal component, scalar power method. */
    v3 v;
    {
        float r0 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
        float r1 = m[1] * m[1] + m[3] * m[3] + m[4] * m[4];
        float r2 = m[2] * m[2] + m[4] * m[4] + m[5] * m[5];
        if (r0 > r1 && r0 > r2) { v.x = m[0]; v.y = m[1]; v.z = m[2]; }
        else if (r1 > r2)       { v.x = m[1]; v.y = m[3]; v.z = m[4]; }
        else                    { v.x = m[2]; v.y = m[4]; v.z = m[5]; }
        for (int it = 0; it < 8; it++) {
            float x = v.x * m[0] + v.y * m[1] + v.z * m[2];
            float y = v.x * m[1] + v.y * m[3] + v.z * m[4];
            float z = v.x * m[2] + v.y * m[4] + v.z * m[5];
            float norm = smax(smax(x, y), z);
            float iv = 1.0f / norm;
            if (norm == 0.0f) { v.x = v.y = v.z = 0.0f; break; }
            v.x = x * iv; v.y = y * iv; v.z = z * iv;
        }
    }
    /* weightedclusterfit.cpp:56-104: order, weighted points, sums. */
    int order[16];
    float dps[16];
    for (int i = 0; i < count; i++) { dps[i] = pts[i].x * v.x + pts[i].y * v.y + pts[i].z * v.z; order[i] = i; }
    for (int i = 0; i < count; i++)
        for (int j = i; j > 0 && dps[j] < dps[j - 1]; j--) {
            float tt = dps[j]; dps[j] = dps[j - 1]; dps[j - 1] = tt;
            int o = order[j]; order[j] = order[j - 1]; order[j - 1] = o;
        }
    v3 wp[17];
    float ww[17];
    v3 xsum = { 0, 0, 0 };
    float wsum = 0.0f;
    for (int i = 0; i < count; i++) {
        int p = order[i];
        wp[i].x = wts[p] * pts[p].x; wp[i].y = wts[p] * pts[p].y; wp[i].z = wts[p] * pts[p].z;
        xsum.x += wp[i].x; xsum.y += wp[i].y; xsum.z += wp[i].z;
        ww[i] = wts[p];
        wsum += ww[i];
    }
    for (int i = count; i < 17; i++) { wp[i].x = wp[i].y = wp[i].z = 0.0f; ww[i] = 0.0f; } /* read past the end by the reference loops; never used */
    float m_best = FLT_MAX; /* ColourFit::m_besterror */
    if (dxt1) {
        /* weightedclusterfit.cpp:372-474 scalar Compress3: palette {a, b, (a+b)/2}. */
        v3 s3 = { 0, 0, 0 }, e3 = { 0, 0, 0 };
        float best3 = FLT_MAX;
        int d0 = 0, d1 = 0;
        v3 x0 = { 0, 0, 0 };
        float w0 = 0.0f;
        for (int c0 = 0; c0 <= count; c0++) {
            v3 x1 = { 0, 0, 0 };
            float w1 = 0.0f;
            for (int c1 = 0; c1 <= count - c0; c1++) {
                float w2 = (wsum - w0) - w1;
                float alpha2 = w0 + w1 * 0.25f;
                float beta2 = w2 + w1 * 0.25f;
                float ab = w1 * 0.25f;
                float factor = 1.0f / (alpha2 * beta2 - ab * ab);
                v3 ax = { x0.x + x1.x * 0.5f, x0.y + x1.y * 0.5f, x0.z + x1.z * 0.5f };
                v3 bx = { xsum.x - ax.x, xsum.y - ax.y, xsum.z - ax.z };
                v3 a = { (ax.x * beta2 - bx.x * ab) * factor, (ax.y * beta2 - bx.y * ab) * factor, (ax.z * beta2 - bx.z * ab) * factor };
                v3 b = { (bx.x * alpha2 - ax.x * ab) * factor, (bx.y * alpha2 - ax.y * ab) * factor, (bx.z * alpha2 - ax.z * ab) * factor };
                a.x = smin(1.0f, smax(0.0f, a.x)); a.y = smin(1.0f, smax(0.0f, a.y)); a.z = smin(1.0f, smax(0.0f, a.z));
                b.x = smin(1.0f, smax(0.0f, b.x)); b.y = smin(1.0f, smax(0.0f, b.y)); b.z = smin(1.0f, smax(0.0f, b.z));
                a.x = floorf(31.0f * a.x + 0.5f) * (1.0f / 31.0f);
                a.y = floorf(63.0f * a.y + 0.5f) * (1.0f / 63.0f);
                a.z = floorf(31.0f * a.z + 0.5f) * (1.0f / 31.0f);
                b.x = floorf(31.0f * b.x + 0.5f) * (1.0f / 31.0f);
                b.y = floorf(63.0f * b.y + 0.5f) * (1.0f / 63.0f);
                b.z = floorf(31.0f * b.z + 0.5f) * (1.0f / 31.0f);
                v3 e1 = {
                    ((a.x * a.x) * alpha2 + (b.x * b.x) * beta2) + 2.0f * ((((a.x * b.x) * ab) - a.x * ax.x) - b.x * bx.x),
                    ((a.y * a.y) * alpha2 + (b.y * b.y) * beta2) + 2.0f * ((((a.y * b.y) * ab) - a.y * ax.y) - b.y * bx.y),
                    ((a.z * a.z) * alpha2 + (b.z * b.z) * beta2) + 2.0f * ((((a.z * b.z) * ab) - a.z * ax.z) - b.z * bx.z) };
                float error = e1.x * msq.x + e1.y * msq.y + e1.z * msq.z;
                if (error < best3) { best3 = error; s3 = a; e3 = b; d0 = c0; d1 = c1; }
                int k = c0 + c1;
                x1.x += wp[k].x; x1.y += wp[k].y; x1.z += wp[k].z; w1 += ww[k];
            }
            x0.x += wp[c0].x; x0.y += wp[c0].y; x0.z += wp[c0].z; w0 += ww[c0];
        }
        if (best3 < m_best) {
            uint8_t sorted_idx[16], ordered[16], idx[16];
            for (int i = 0; i < count; i++) sorted_idx[i] = (uint8_t)(i < d0 ? 0 : i < d0 + d1 ? 2 : 1);
            for (int i = 0; i < count; i++) ordered[order[i]] = sorted_idx[i];
            for (int i = 0; i < 16; i++) idx[i] = remap[i] < 0 ? 3 : ordered[remap[i]]; /* colourset.cpp:140-150 */
            /* colourblock.cpp:75-106 WriteColourBlock3 (byte indices). */
            int ca = float_to_565(s3), cb = float_to_565(e3);
            if (ca > cb) {
                int tt = ca; ca = cb; cb = tt;
                for (int i = 0; i < 16; i++) idx[i] = (uint8_t)(idx[i] == 0 ? 1 : idx[i] == 1 ? 0 : idx[i]);
            }
            uint32_t packed = 0;
            for (int i = 0; i < 16; i++) packed |= (uint32_t)idx[i] << (2 * i);
            put_block(out, (uint16_t)ca, (uint16_t)cb, packed);
            m_best = best3;
        }
        if (transparent) return;
    }
    /* weightedclusterfit.cpp:476-589 scalar Compress4. */
    v3 beststart = { 0, 0, 0 }, bestend = { 0, 0, 0 };
    float besterror = FLT_MAX;
    int b0 = 0, b1 = 0, b2 = 0;
    v3 x0 = { 0, 0, 0 };
    float w0 = 0.0f;
    for (int c0 = 0; c0 <= count; c0++) {
        v3 x1 = { 0, 0, 0 };
        float w1 = 0.0f;
        for (int c1 = 0; c1 <= count - c0; c1++) {
            v3 x2 = { 0, 0, 0 };
            float w2 = 0.0f;
            for (int c2 = 0; c2 <= count - c0 - c1; c2++) {
                float w3 = ((wsum - w0) - w1) - w2;
                float alpha2 = (w0 + w1 * (4.0f / 9.0f)) + w2 * (1.0f / 9.0f);
                float beta2 = (w3 + w2 * (4.0f / 9.0f)) + w1 * (1.0f / 9.0f);
                float ab = (w1 + w2) * (2.0f / 9.0f);
                float factor = 1.0f / (alpha2 * beta2 - ab * ab);
                v3 ax = { (x0.x + x1.x * (2.0f / 3.0f)) + x2.x * (1.0f / 3.0f),
                          (x0.y + x1.y * (2.0f / 3.0f)) + x2.y * (1.0f / 3.0f),
                          (x0.z + x1.z * (2.0f / 3.0f)) + x2.z * (1.0f / 3.0f) };
                v3 bx = { xsum.x - ax.x, xsum.y - ax.y, xsum.z - ax.z };
                v3 a = { (ax.x * beta2 - bx.x * ab) * factor, (ax.y * beta2 - bx.y * ab) * factor, (ax.z * beta2 - bx.z * ab) * factor };
                v3 b = { (bx.x * alpha2 - ax.x * ab) * factor, (bx.y * alpha2 - ax.y * ab) * factor, (bx.z * alpha2 - ax.z * ab) * factor };
                a.x = smin(1.0f, smax(0.0f, a.x)); a.y = smin(1.0f, smax(0.0f, a.y)); a.z = smin(1.0f, smax(0.0f, a.z));
                b.x = smin(1.0f, smax(0.0f, b.x)); b.y = smin(1.0f, smax(0.0f, b.y)); b.z = smin(1.0f, smax(0.0f, b.z));
                a.x = floorf(31.0f * a.x + 0.5f) * (1.0f / 31.0f);
                a.y = floorf(63.0f * a.y + 0.5f) * (1.0f / 63.0f);
                a.z = floorf(31.0f * a.z + 0.5f) * (1.0f / 31.0f);
                b.x = floorf(31.0f * b.x + 0.5f) * (1.0f / 31.0f);
                b.y = floorf(63.0f * b.y + 0.5f) * (1.0f / 63.0f);
                b.z = floorf(31.0f * b.z + 0.5f) * (1.0f / 31.0f);
                v3 e1 = {
                    ((a.x * a.x) * alpha2 + (b.x * b.x) * beta2) + 2.0f * ((((a.x * b.x) * ab) - a.x * ax.x) - b.x * bx.x),
                    ((a.y * a.y) * alpha2 + (b.y * b.y) * beta2) + 2.0f * ((((a.y * b.y) * ab) - a.y * ax.y) - b.y * bx.y),
                    ((a.z * a.z) * alpha2 + (b.z * b.z) * beta2) + 2.0f * ((((a.z * b.z) * ab) - a.z * ax.z) - b.z * bx.z) };
                float error = e1.x * msq.x + e1.y * msq.y + e1.z * msq.z;
                if (error < besterror) {
                    besterror = error; beststart = a; bestend = b; b0 = c0; b1 = c1; b2 = c2;
                }
                int k = c0 + c1 + c2;
                x2.x += wp[k].x; x2.y += wp[k].y; x2.z += wp[k].z; w2 += ww[k];
            }
            int k = c0 + c1;
            x1.x += wp[k].x; x1.y += wp[k].y; x1.z += wp[k].z; w1 += ww[k];
        }
        x0.x += wp[c0].x; x0.y += wp[c0].y; x0.z += wp[c0].z; w0 += ww[c0];
    }
    if (!(besterror < m_best)) { /* no improvement: the block stays as written, or unwritten (zeros) */
        if (!(m_best < FLT_MAX)) memset(out, 0, 8);
        return;
    }
    uint8_t sorted_idx[16], ordered[16], idx[16];
    for (int i = 0; i < count; i++) sorted_idx[i] = (uint8_t)(i < b0 ? 0 : i < b0 + b1 ? 2 : i < b0 + b1 + b2 ? 3 : 1);
    for (int i = 0; i < count; i++) ordered[order[i]] = sorted_idx[i];
    for (int i = 0; i < 16; i++) idx[i] = remap[i] < 0 ? 3 : ordered[remap[i]]; /* colourset.cpp:140-150 RemapIndices */
    /* colourblock.cpp:108-138 WriteColourBlock4. */
    int ca = float_to_565(beststart), cb = float_to_565(bestend);
    if (ca < cb) { int tt = ca; ca = cb; cb = tt; for (int i = 0; i < 16; i++) idx[i] = (uint8_t)((idx[i] ^ 1) & 3); }
    else if (ca == cb) { for (int i = 0; i < 16; i++) idx[i] = 0; }
    uint32_t packed = 0;
    for (int i = 0; i < 16; i++) packed |= (uint32_t)idx[i] << (2 * i);
    put_block(out, (uint16_t)ca, (uint16_t)cb, packed);
}

/* =========================================================================
 * BC4 channel: QuickCompress::compressDXT5A (QuickCompressDXT.cpp:505-592, 643-647, 772-821)
 * ========================================================================= */

typedef struct { uint64_t u; } ablk;

static inline unsigned ab_a0(ablk b) { return (unsigned)(b.u & 0xff); }
static inline unsigned ab_a1(ablk b) { return (unsigned)((b.u >> 8) & 0xff); }
static inline unsigned ab_index(ablk b, int i) { return (unsigned)((b.u >> (16 + 3 * i)) & 7); }
static inline void ab_set_index(ablk *b, int i, unsigned v)
{
    uint64_t m = (uint64_t)7 << (16 + 3 * i);
    b->u = (b->u & ~m) | ((uint64_t)v << (16 + 3 * i));
}
static inline void ab_set_ends(ablk *b, unsigned a0, unsigned a1)
{
    b->u = (b->u & ~(uint64_t)0xffff) | (uint64_t)(a0 & 0xff) | ((uint64_t)(a1 & 0xff) << 8);
}

/* BlockDXT.cpp:336-378 evaluatePalette8/6 (d3d9=false, no bias). */
static void alpha_palette(unsigned a0, unsigned a1, uint8_t pal[8])
{
    pal[0] = (uint8_t)a0; pal[1] = (uint8_t)a1;
    if (a0 > a1) {
        for (int k = 1; k <= 6; k++) pal[1 + k] = (uint8_t)(((7 - k) * a0 + k * a1) / 7);
    } else {
        for (int k = 1; k <= 4; k++) pal[1 + k] = (uint8_t)(((5 - k) * a0 + k * a1) / 5);
        pal[6] = 0; pal[7] = 255;
    }
}

static unsigned alpha_indices(const uint8_t *v, ablk *b)
{
    uint8_t pal[8];
    alpha_palette(ab_a0(*b), ab_a1(*b), pal);
    unsigned total = 0;
    for (int i = 0; i < 16; i++) {
        unsigned besterr = 256 * 256, best = 8;
        for (unsigned p = 0; p < 8; p++) {
            int d = pal[p] - v[i];
            unsigned e = (unsigned)(d * d);
            if (e < besterr) { besterr = e; best = p; }
        }
        total += besterr;
        ab_set_index(b, i, best);
    }
    return total;
}

/* nv::max / nv::min (nvcore/Utils.h:160-188) on floats. */
static inline float nvmax(float a, float b) { return (b < a) ? a : b; }
static inline float nvmin(float a, float b) { return (a < b) ? a : b; }

static void optimize_alpha8(const uint8_t *v, ablk *b)
{
    float alpha2 = 0, beta2 = 0, ab = 0, ax = 0, bx = 0;
    for (int i = 0; i < 16; i++) {
        unsigned idx = ab_index(*b, i);
        float alpha = (idx < 2) ? 1.0f - (float)idx : (8.0f - (float)idx) / 7.0f;
        float beta = 1 - alpha;
        alpha2 += alpha * alpha;
        beta2 += beta * beta;
        ab += alpha * beta;
        ax += alpha * v[i];
        bx += beta * v[i];
    }
    const float factor = 1.0f / (alpha2 * beta2 - ab * ab);
    float a = (ax * beta2 - bx * ab) * factor;
    float bb = (bx * alpha2 - ax * ab) * factor;
    unsigned a0 = (unsigned)nvmin(nvmax(a, 0.0f), 255.0f);
    unsigned a1 = (unsigned)nvmin(nvmax(bb, 0.0f), 255.0f);
    if (a0 < a1) {
        unsigned t = a0; a0 = a1; a1 = t;
        for (int i = 0; i < 16; i++) {
            unsigned idx = ab_index(*b, i);
            ab_set_index(b, i, idx < 2 ? 1 - idx : 9 - idx);
        }
    } else if (a0 == a1) {
        for (int i = 0; i < 16; i++) ab_set_index(b, i, 0);
    }
    ab_set_ends(b, a0, a1);
}

void ntxo_bc4_channel(const uint8_t *v, uint8_t *out)
{
    unsigned mx = 0, mn = 255;
    for (int i = 0; i < 16; i++) { if (mx < v[i]) mx = v[i]; if (v[i] < mn) mn = v[i]; }
    ablk block = { 0 };
    ab_set_ends(&block, (mx - (mx - mn) / 34) & 0xff, (mn + (mx - mn) / 34) & 0xff);
    unsigned besterr = alpha_indices(v, &block);
    ablk best = block;
    for (int it = 0; it < 8; it++) {
        optimize_alpha8(v, &block);
        unsigned err = alpha_indices(v, &block);
        if (err >= besterr) break;
        /* sameIndices (QuickCompressDXT.cpp:643-647) ORs ~0xFFFF into both words,
         * so it compares only the two endpoint bytes. */
        if ((block.u & 0xffff) == (best.u & 0xffff)) { best = block; break; }
        besterr = err;
        best = block;
    }
    for (int i = 0; i < 8; i++) out[i] = (uint8_t)(best.u >> (8 * i));
}

void ntxo_bc4_channels(const uint8_t *v, size_t n, uint8_t *out)
{
    for (size_t i = 0; i < n; i++) ntxo_bc4_channel(v + 16 * i, out + 8 * i);
}

void ntxo_bc4_decode(const uint8_t *blk, uint8_t *v)
{
    uint64_t u = 0;
    for (int i = 0; i < 8; i++) u |= (uint64_t)blk[i] << (8 * i);
    uint8_t pal[8];
    alpha_palette(blk[0], blk[1], pal);
    for (int i = 0; i < 16; i++) v[i] = pal[(u >> (16 + 3 * i)) & 7];
}

void ntxo_bc3_block(const uint8_t *bgra, int weight_by_alpha, const float *cw, uint8_t *out)
{
    uint8_t a[16];
    for (int i = 0; i < 16; i++) a[i] = bgra[4 * i + 3];
    ntxo_bc4_channel(a, out);
    ntxo_bc3_color_block(bgra, weight_by_alpha, cw, out + 8);
}

/* The squish branch of CompressorDXT1a::compressBlock (CompressorDX9.cpp:98-109):
 * kDxt1 colour set, plus kWeightColourByAlpha when weight_by_alpha. */
void ntxo_bc1a_squish(const uint8_t *bgra, int weight_by_alpha, const float *cw, uint8_t *out)
{
    ntxo_init();
    squish_fit(bgra, weight_by_alpha, cw, 1, out);
}

/* CompressorDXT1a::compressBlock (CompressorDX9.cpp:83-111): the alpha-0
 * mask; one RGB over all 16 texels -> OptimalCompress::compressDXT1a
 * (OptimalCompressDXT.cpp:271-292: compressDXT1 without transparent texels,
 * else the OMatchAlpha endpoints, col0 <= col1 by swapping, indices
 * 0xaaaaaaaa | mask); otherwise the kDxt1 squish fit. */
void ntxo_bc1a_block(const uint8_t *bgra, int weight_by_alpha, const float *cw, uint8_t *out)
{
    ntxo_init();
    uint32_t mask = 0;
    int single = 1;
    for (int i = 0; i < 16; i++) {
        if (bgra[4 * i + 3] == 0) mask |= 3u << (2 * i);
        single = single && bgra[4 * i] == bgra[0] && bgra[4 * i + 1] == bgra[1] && bgra[4 * i + 2] == bgra[2];
    }
    if (!single) { squish_fit(bgra, weight_by_alpha, cw, 1, out); return; }
    const int r = bgra[2], g = bgra[1], b = bgra[0];
    if (mask == 0) {
        uint16_t c0 = (uint16_t)((g_omatch5[r][0] << 11) | (g_omatch6[g][0] << 5) | g_omatch5[b][0]);
        uint16_t c1 = (uint16_t)((g_omatch5[r][1] << 11) | (g_omatch6[g][1] << 5) | g_omatch5[b][1]);
        uint32_t idx = 0xaaaaaaaau;
        if (c0 < c1) { uint16_t t = c0; c0 = c1; c1 = t; idx ^= 0x55555555u; }
        put_block(out, c0, c1, idx);
        return;
    }
    uint16_t c0 = (uint16_t)((g_omatcha5[r][0] << 11) | (g_omatcha6[g][0] << 5) | g_omatcha5[b][0]);
    uint16_t c1 = (uint16_t)((g_omatcha5[r][1] << 11) | (g_omatcha6[g][1] << 5) | g_omatcha5[b][1]);
    if (c0 > c1) { uint16_t t = c0; c0 = c1; c1 = t; }
    put_block(out, c0, c1, 0xaaaaaaaau | mask);
}

/* CompressorDXT5n::compressBlock at Normal / Production
 * (CompressorDX9.cpp:179-222).  Colour: a block whose texels share one green
 * takes OptimalCompress::compressDXT1G(g) (OptimalCompressDXT.cpp:294-311:
 * red 31, blue 0, green from OMatch6); any other goes through the weighted
 * cluster fit on the swizzled texels (red 0xFF, G, blue 0, A) with metric
 * (0, 1, 0).  Alpha: QuickCompress::compressDXT5A of the red channel
 * (swizzle(4, 1, 5, 0) moves R into alpha). */
void ntxo_bc3n_block(const uint8_t *bgra, int weight_by_alpha, uint8_t *out)
{
    ntxo_init();
    uint8_t tile[64], x[16];
    int single = 1;
    for (int i = 0; i < 16; i++) {
        tile[4 * i] = 0;
        tile[4 * i + 1] = bgra[4 * i + 1];
        tile[4 * i + 2] = 0xFF;
        tile[4 * i + 3] = bgra[4 * i + 3];
        x[i] = bgra[4 * i + 2];
        single = single && bgra[4 * i + 1] == bgra[1];
    }
    if (single) {
        const int g = bgra[1];
        uint16_t c0 = (uint16_t)((31 << 11) | (g_omatch6[g][0] << 5));
        uint16_t c1 = (uint16_t)((31 << 11) | (g_omatch6[g][1] << 5));
        uint32_t idx = 0xaaaaaaaau;
        if (c0 < c1) { uint16_t t = c0; c0 = c1; c1 = t; idx ^= 0x55555555u; }
        put_block(out + 8, c0, c1, idx);
    } else {
        const float green[4] = { 0.0f, 1.0f, 0.0f, 1.0f };
        ntxo_bc3_color_block(tile, weight_by_alpha, green, out + 8);
    }
    ntxo_bc4_channel(x, out);
}

/* =========================================================================
 * Highest / Production channel blocks: OptimalCompress::compressDXT5A
 * (OptimalCompressDXT.cpp:512-607, computeAlphaError :189-216,
 * computeAlphaIndices :219-246, alphaDistance :52-57)
 * ========================================================================= */

/* computeAlphaError with unit weights: sum over texels of the squared
 * distance to the nearest palette entry (integers, so the reference's float
 * sum is exact).  Like the reference it stops once the sum exceeds `bound`:
 * such a pair cannot win, so the outcome is the full search's. */
static unsigned alpha_error(const uint8_t *v, unsigned a0, unsigned a1, unsigned bound)
{
    uint8_t pal[8];
    alpha_palette(a0, a1, pal);
    unsigned total = 0;
    for (int i = 0; i < 16; i++) {
        unsigned best = ~0u;
        for (int p = 0; p < 8; p++) {
            const int d = (int)pal[p] - (int)v[i];
            if ((unsigned)(d * d) < best) best = (unsigned)(d * d);
        }
        total += best;
        if (total > bound) break;
    }
    return total;
}

/* OptimalCompress::compressDXT5A.  The reference seeds its running best
 * error from whatever endpoints the output block held before the call
 * (line 546); here the seed is the error of (max, min), the endpoints the
 * search keeps when no pair beats the seed. */
void ntxo_bc4_channel_optimal(const uint8_t *v, uint8_t *out)
{
    unsigned mina = 255, maxa = 0, mina01 = 255, maxa01 = 0;
    for (int i = 0; i < 16; i++) {
        const unsigned a = v[i];
        if (a < mina) mina = a;
        if (a > maxa) maxa = a;
        if (a != 0 && a != 255) {
            if (a < mina01) mina01 = a;
            if (a > maxa01) maxa01 = a;
        }
    }
    unsigned e0, e1;
    if ((int)maxa - (int)mina < 8) {
        e0 = maxa; e1 = mina;
    } else if ((int)maxa01 - (int)mina01 < 6) {
        e0 = mina01; e1 = maxa01;
    } else {
        unsigned best = alpha_error(v, maxa, mina, ~0u);
        e0 = maxa; e1 = mina;
        /* Eight-value palettes: a0 - a1 > 8 within the range widened by 8. */
        const int lo = mina <= 8 ? 0 : (int)mina - 8, hi = maxa >= 247 ? 255 : (int)maxa + 8;
        for (int a0 = lo + 9; a0 < hi; a0++)
            for (int a1 = lo; a1 < a0 - 8; a1++) {
                const unsigned e = alpha_error(v, (unsigned)a0, (unsigned)a1, best);
                if (e < best) { best = e; e0 = (unsigned)a0; e1 = (unsigned)a1; }
            }
        /* Six-value palettes (0 and 255 implicit): endpoints swapped, the
         * range of the texels other than 0 and 255 widened by 6. */
        const int lo6 = mina01 <= 6 ? 0 : (int)mina01 - 6, hi6 = maxa01 >= 249 ? 255 : (int)maxa01 + 6;
        for (int a0 = lo6 + 9; a0 < hi6; a0++)
            for (int a1 = lo6; a1 < a0 - 8; a1++) {
                const unsigned e = alpha_error(v, (unsigned)a1, (unsigned)a0, best);
                if (e < best) { best = e; e0 = (unsigned)a1; e1 = (unsigned)a0; }
            }
    }
    ablk block = { 0 };
    ab_set_ends(&block, e0, e1);
    alpha_indices(v, &block);
    for (int i = 0; i < 8; i++) out[i] = (uint8_t)(block.u >> (8 * i));
}

void ntxo_bc4_channels_optimal(const uint8_t *v, size_t n, uint8_t *out)
{
    for (size_t i = 0; i < n; i++) ntxo_bc4_channel_optimal(v + 16 * i, out + 8 * i);
}

/* OptimalCompress::compressDXT1G(const ColorBlock &) (OptimalCompressDXT.cpp:315-390):
 * green-only BC1 (red 31, blue 0) by exhaustive search of 6-bit green
 * pairs, with computeGreenError (:71-110: four-colour palette of the 8-bit
 * expansions, squared distance) and computeGreenIndices (:112-145).  The
 * quantised green of 255 is (255 + 1) >> 2 = 64, which the 6-bit endpoint
 * field stores as 0; the seed endpoints keep that wrap as the reference
 * does. */
static unsigned expand6(unsigned g) { g &= 63; return (g << 2) | (g >> 4); }
static unsigned expand5(unsigned r) { r &= 31; return (r << 3) | (r >> 2); }

static unsigned green_error(const uint8_t *bgra, unsigned g0, unsigned g1)
{
    int p[4];
    p[0] = (int)expand6(g0); p[1] = (int)expand6(g1);
    p[2] = (2 * p[0] + p[1]) / 3; p[3] = (2 * p[1] + p[0]) / 3;
    unsigned total = 0;
    for (int i = 0; i < 16; i++) {
        const int g = bgra[4 * i + 1];
        unsigned best = ~0u;
        for (int k = 0; k < 4; k++) {
            const unsigned e = (unsigned)((g - p[k]) * (g - p[k]));
            if (e < best) best = e;
        }
        total += best;
    }
    return total;
}

void ntxo_bc1_green_optimal(const uint8_t *bgra, uint8_t *out)
{
    ntxo_init();
    unsigned ming = 63, maxg = 0;
    int single = 1;
    for (int i = 0; i < 16; i++) {
        const unsigned g = ((unsigned)bgra[4 * i + 1] + 1) >> 2;
        if (g < ming) ming = g;
        if (g > maxg) maxg = g;
        single = single && bgra[4 * i + 1] == bgra[1];
    }
    if (single) {
        const int g = bgra[1];
        uint16_t c0 = (uint16_t)((31 << 11) | (g_omatch6[g][0] << 5));
        uint16_t c1 = (uint16_t)((31 << 11) | (g_omatch6[g][1] << 5));
        uint32_t idx = 0xaaaaaaaau;
        if (c0 < c1) { uint16_t t = c0; c0 = c1; c1 = t; idx ^= 0x55555555u; }
        put_block(out, c0, c1, idx);
        return;
    }
    unsigned best = green_error(bgra, maxg, ming), b0 = maxg, b1 = ming;
    const unsigned lo = ming <= 4 ? 0 : ming - 4, hi = maxg >= 59 ? 63 : maxg + 4;
    for (unsigned g0 = lo + 1; g0 <= hi; g0++)
        for (unsigned g1 = lo; g1 < g0; g1++) {
            const unsigned e = green_error(bgra, g0, g1);
            if (e < best) { best = e; b0 = g0; b1 = g1; }
        }
    /* BlockDXT1::evaluatePalette (BlockDXT.cpp:43-99), green only. */
    const uint16_t c0 = (uint16_t)((31 << 11) | ((b0 & 63) << 5)), c1 = (uint16_t)((31 << 11) | ((b1 & 63) << 5));
    int p[4];
    p[0] = (int)expand6(b0); p[1] = (int)expand6(b1);
    if (c0 > c1) { p[2] = (2 * p[0] + p[1]) / 3; p[3] = (2 * p[1] + p[0]) / 3; }
    else { p[2] = (p[0] + p[1]) / 2; p[3] = 0; }
    uint32_t idx = 0;
    for (int i = 0; i < 16; i++) {
        const int g = bgra[4 * i + 1];
        const unsigned d0 = (unsigned)((p[0] - g) * (p[0] - g)), d1 = (unsigned)((p[1] - g) * (p[1] - g));
        const unsigned d2 = (unsigned)((p[2] - g) * (p[2] - g)), d3 = (unsigned)((p[3] - g) * (p[3] - g));
        const unsigned x0 = (d1 > d2) & (d0 > d2), x1 = (d0 > d3) & (d1 > d3), x2 = (d0 > d3) & (d2 > d3);
        idx |= (x2 | ((x0 | x1) << 1)) << (2 * i);
    }
    put_block(out, c0, c1, idx);
}

/* =========================================================================
 * Fastest: QuickCompress::compressDXT1 / compressDXT5
 * (QuickCompressDXT.cpp:45-452, 653-713, 823-827), in the reference's float
 * operation order (nvmath Vector3: lerp a*s + t*b, dot x+y+z, division by a
 * scalar per component).
 * ========================================================================= */

static v3 q_v3(float x, float y, float z) { v3 r = { x, y, z }; return r; }
static float q_dist(v3 a, v3 b)
{
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return dx * dx + dy * dy + dz * dz;
}
static v3 q_lerp(v3 a, v3 b, float t)
{
    const float s = 1.0f - t;
    return q_v3(a.x * s + t * b.x, a.y * s + t * b.y, a.z * s + t * b.z);
}
static float q_clamp(float x, float a, float b) { return nvmin(nvmax(x, a), b); }

/* findMinMaxColorsBox (:72-82), selectDiagonal (:85-110), insetBBox (:112-117). */
static void q_box(const v3 *blk, int num, v3 *mx, v3 *mn)
{
    *mx = q_v3(0, 0, 0);
    *mn = q_v3(255, 255, 255);
    for (int i = 0; i < num; i++) {
        mx->x = nvmax(mx->x, blk[i].x); mx->y = nvmax(mx->y, blk[i].y); mx->z = nvmax(mx->z, blk[i].z);
        mn->x = nvmin(mn->x, blk[i].x); mn->y = nvmin(mn->y, blk[i].y); mn->z = nvmin(mn->z, blk[i].z);
    }
    const v3 c = q_v3((mx->x + mn->x) * 0.5f, (mx->y + mn->y) * 0.5f, (mx->z + mn->z) * 0.5f);
    float cx = 0.0f, cy = 0.0f;
    for (int i = 0; i < num; i++) {
        const float tx = blk[i].x - c.x, ty = blk[i].y - c.y, tz = blk[i].z - c.z;
        cx += tx * tz;
        cy += ty * tz;
    }
    float x0 = mx->x, y0 = mx->y, x1 = mn->x, y1 = mn->y;
    if (cx < 0) { float t = x0; x0 = x1; x1 = t; }
    if (cy < 0) { float t = y0; y0 = y1; y1 = t; }
    mx->x = x0; mx->y = y0;
    mn->x = x1; mn->y = y1;
    const float k = (8.0f / 255.0f) / 16.0f;
    const v3 in = q_v3((mx->x - mn->x) / 16.0f - k, (mx->y - mn->y) / 16.0f - k, (mx->z - mn->z) / 16.0f - k);
    *mx = q_v3(q_clamp(mx->x - in.x, 0.0f, 255.0f), q_clamp(mx->y - in.y, 0.0f, 255.0f), q_clamp(mx->z - in.z, 0.0f, 255.0f));
    *mn = q_v3(q_clamp(mn->x + in.x, 0.0f, 255.0f), q_clamp(mn->y + in.y, 0.0f, 255.0f), q_clamp(mn->z + in.z, 0.0f, 255.0f));
}

/* roundAndExpand (:122-150): floor to 5:6:5, step up when the next code's
 * expansion is strictly nearer; v becomes the expanded colour. */
static unsigned q_round_expand(v3 *v)
{
    unsigned r = (unsigned)floorf(q_clamp(v->x * (31.0f / 255.0f), 0.0f, 31.0f));
    unsigned g = (unsigned)floorf(q_clamp(v->y * (63.0f / 255.0f), 0.0f, 63.0f));
    unsigned b = (unsigned)floorf(q_clamp(v->z * (31.0f / 255.0f), 0.0f, 31.0f));
    if (fabsf(v->x - (float)((((r + 1) << 3) | ((r + 1) >> 2)))) < fabsf(v->x - (float)(((r << 3) | (r >> 2)))))
        r = r + 1 < 31 ? r + 1 : 31;
    if (fabsf(v->y - (float)((((g + 1) << 2) | ((g + 1) >> 4)))) < fabsf(v->y - (float)(((g << 2) | (g >> 4)))))
        g = g + 1 < 63 ? g + 1 : 63;
    if (fabsf(v->z - (float)((((b + 1) << 3) | ((b + 1) >> 2)))) < fabsf(v->z - (float)(((b << 3) | (b >> 2)))))
        b = b + 1 < 31 ? b + 1 : 31;
    *v = q_v3((float)expand5(r), (float)expand6(g), (float)expand5(b));
    return (r << 11) | (g << 5) | b;
}

/* computeIndices4 (:195-229). */
static uint32_t q_indices4(const v3 *blk, v3 mx, v3 mn)
{
    const v3 p2 = q_lerp(mx, mn, 1.0f / 3.0f), p3 = q_lerp(mx, mn, 2.0f / 3.0f);
    uint32_t idx = 0;
    for (int i = 0; i < 16; i++) {
        const float d0 = q_dist(mx, blk[i]), d1 = q_dist(mn, blk[i]), d2 = q_dist(p2, blk[i]), d3 = q_dist(p3, blk[i]);
        const unsigned b0 = d0 > d3, b1 = d1 > d2, b2 = d0 > d2, b3 = d1 > d3, b4 = d2 > d3;
        const unsigned x0 = b1 & b2, x1 = b0 & b3, x2 = b0 & b4;
        idx |= (x2 | ((x0 | x1) << 1)) << (2 * i);
    }
    return idx;
}

/* optimizeEndPoints4 (:406-452): least squares on the current indices. */
static void q_optimize4(const v3 *blk, uint16_t *c0, uint16_t *c1, uint32_t *idx)
{
    float a2 = 0.0f, b2 = 0.0f, ab = 0.0f;
    v3 ax = q_v3(0, 0, 0), bx = q_v3(0, 0, 0);
    for (int i = 0; i < 16; i++) {
        const unsigned bits = *idx >> (2 * i);
        float beta = (float)(bits & 1);
        if (bits & 2) beta = (1 + beta) / 3.0f;
        const float alpha = 1.0f - beta;
        a2 += alpha * alpha;
        b2 += beta * beta;
        ab += alpha * beta;
        ax.x += blk[i].x * alpha; ax.y += blk[i].y * alpha; ax.z += blk[i].z * alpha;
        bx.x += blk[i].x * beta; bx.y += blk[i].y * beta; bx.z += blk[i].z * beta;
    }
    const float denom = a2 * b2 - ab * ab;
    /* nv::equal(denom, 0) (nvmath.h:139-143): |denom| <= 1e-4 * max(1, |denom|). */
    if (fabsf(denom) <= 0.0001f * nvmax(1.0f, nvmax(fabsf(denom), 0.0f))) return;
    const float f = 1.0f / denom;
    v3 a = q_v3((ax.x * b2 - bx.x * ab) * f, (ax.y * b2 - bx.y * ab) * f, (ax.z * b2 - bx.z * ab) * f);
    v3 b = q_v3((bx.x * a2 - ax.x * ab) * f, (bx.y * a2 - ax.y * ab) * f, (bx.z * a2 - ax.z * ab) * f);
    a = q_v3(q_clamp(a.x, 0.0f, 255.0f), q_clamp(a.y, 0.0f, 255.0f), q_clamp(a.z, 0.0f, 255.0f));
    b = q_v3(q_clamp(b.x, 0.0f, 255.0f), q_clamp(b.y, 0.0f, 255.0f), q_clamp(b.z, 0.0f, 255.0f));
    unsigned k0 = q_round_expand(&a), k1 = q_round_expand(&b);
    if (k0 < k1) { v3 t = a; a = b; b = t; unsigned u = k0; k0 = k1; k1 = u; }
    *c0 = (uint16_t)k0;
    *c1 = (uint16_t)k1;
    *idx = q_indices4(blk, a, b);
}

/* QuickCompress::compressDXT1 (:653-713) on BGRA texels. */
static void q_dxt1(const uint8_t *bgra, uint8_t *out)
{
    int single = 1;
    for (int i = 1; i < 16; i++)
        single = single && bgra[4 * i] == bgra[0] && bgra[4 * i + 1] == bgra[1] && bgra[4 * i + 2] == bgra[2];
    if (single) {
        const int r = bgra[2], g = bgra[1], b = bgra[0];
        uint16_t c0 = (uint16_t)((g_omatch5[r][0] << 11) | (g_omatch6[g][0] << 5) | g_omatch5[b][0]);
        uint16_t c1 = (uint16_t)((g_omatch5[r][1] << 11) | (g_omatch6[g][1] << 5) | g_omatch5[b][1]);
        uint32_t idx = 0xaaaaaaaau;
        if (c0 < c1) { uint16_t t = c0; c0 = c1; c1 = t; idx ^= 0x55555555u; }
        put_block(out, c0, c1, idx);
        return;
    }
    v3 blk[16];
    for (int i = 0; i < 16; i++) blk[i] = q_v3(bgra[4 * i + 2], bgra[4 * i + 1], bgra[4 * i]);
    v3 mx, mn;
    q_box(blk, 16, &mx, &mn);
    unsigned k0 = q_round_expand(&mx), k1 = q_round_expand(&mn);
    if (k0 < k1) { v3 t = mx; mx = mn; mn = t; unsigned u = k0; k0 = k1; k1 = u; }
    uint16_t c0 = (uint16_t)k0, c1 = (uint16_t)k1;
    uint32_t idx = q_indices4(blk, mx, mn);
    q_optimize4(blk, &c0, &c1, &idx);
    put_block(out, c0, c1, idx);
}

/* CompressorDXT5 / FastCompressorDXT5 / CompressorDXT5n / FastCompressorDXT5n
 * at every quality (CompressorDX9.cpp:67-80, 145-222; chosen by
 * Context.cpp:1080-1097).  Highest: alpha (BC3n: red) by
 * OptimalCompress::compressDXT5A; BC3n colour by compressDXT1G.  Fastest:
 * QuickCompress::compressDXT5 (:823-827), BC3n after swizzle(4, 1, 5, 0). */
void ntxo_bc3_block_q(const uint8_t *bgra, int quality, int weight_by_alpha, const float *cw, uint8_t *out)
{
    ntxo_init();
    uint8_t a[16];
    for (int i = 0; i < 16; i++) a[i] = bgra[4 * i + 3];
    if (quality == NTXO_QUALITY_FASTEST) {
        q_dxt1(bgra, out + 8);
        ntxo_bc4_channel(a, out);
        return;
    }
    if (quality == NTXO_QUALITY_HIGHEST) ntxo_bc4_channel_optimal(a, out);
    else ntxo_bc4_channel(a, out);
    ntxo_bc3_color_block(bgra, weight_by_alpha, cw, out + 8);
}

void ntxo_bc3n_block_q(const uint8_t *bgra, int quality, int weight_by_alpha, uint8_t *out)
{
    ntxo_init();
    uint8_t x[16];
    for (int i = 0; i < 16; i++) x[i] = bgra[4 * i + 2];
    if (quality == NTXO_QUALITY_FASTEST) {
        uint8_t sw[64];
        for (int i = 0; i < 16; i++) { sw[4 * i] = 0; sw[4 * i + 1] = bgra[4 * i + 1]; sw[4 * i + 2] = 0xFF; sw[4 * i + 3] = x[i]; }
        q_dxt1(sw, out + 8);
        ntxo_bc4_channel(x, out);
        return;
    }
    if (quality == NTXO_QUALITY_HIGHEST) {
        ntxo_bc1_green_optimal(bgra, out + 8);
        ntxo_bc4_channel_optimal(x, out);
        return;
    }
    ntxo_bc3n_block(bgra, weight_by_alpha, out);
}

/* =========================================================================
 * Whole images
 * ========================================================================= */

size_t ntxo_output_size(int format, int w, int h)
{
    size_t bs = (format == NTXO_FMT_BC1 || format == NTXO_FMT_BC1A || format == NTXO_FMT_BC4) ? 8 : 16;
    return (size_t)((w + 3) / 4) * (size_t)((h + 3) / 4) * bs;
}

/* ColorBlock.cpp:80-110: truncating float->u8, edge blocks repeat pixels. */
void ntxo_gather_u8(int w, int h, const float *d, int bx, int by, uint8_t *bgra)
{
    int x = 4 * bx, y = 4 * by;
    int bw = w - x < 4 ? w - x : 4, bh = h - y < 4 ? h - y : 4;
    size_t plane = (size_t)w * h;
    for (int i = 0; i < 4; i++)
        for (int e = 0; e < 4; e++) {
            size_t idx = (size_t)(y + i % bh) * w + x + e % bw;
            uint8_t *c = bgra + 4 * (4 * i + e);
            c[2] = (uint8_t)(255 * nvmin(nvmax(d[idx], 0.0f), 1.0f));
            c[1] = (uint8_t)(255 * nvmin(nvmax(d[idx + plane], 0.0f), 1.0f));
            c[0] = (uint8_t)(255 * nvmin(nvmax(d[idx + 2 * plane], 0.0f), 1.0f));
            c[3] = (uint8_t)(255 * nvmin(nvmax(d[idx + 3 * plane], 0.0f), 1.0f));
        }
}

/* BlockCompressor.cpp:116-166: float gather for BC1; texels outside the image
 * get colour 0 and weight 0. */
static void gather_f32(int w, int h, const float *d, int bx, int by, int alpha_mode, float *col, float *wt)
{
    int x0 = 4 * bx, y0 = 4 * by;
    int bw = w - x0 < 4 ? w - x0 : 4, bh = h - y0 < 4 ? h - y0 : 4;
    size_t plane = (size_t)w * h;
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) {
            int k = 4 * y + x;
            if (x < bw && y < bh) {
                size_t s = (size_t)(y + y0) * w + x + x0;
                col[4 * k] = d[s]; col[4 * k + 1] = d[s + plane]; col[4 * k + 2] = d[s + 2 * plane]; col[4 * k + 3] = d[s + 3 * plane];
                wt[k] = (alpha_mode == NTXO_ALPHA_TRANSPARENCY) ? nvmin(nvmax(d[s + 3 * plane], 0.0f), 1.0f) : 1.0f;
            } else {
                col[4 * k] = col[4 * k + 1] = col[4 * k + 2] = col[4 * k + 3] = 0.0f;
                wt[k] = 0.0f;
            }
        }
}

typedef struct {
    int format, quality, alpha_mode, w, h, bw;
    const float *cw;
    const float *data;
    uint8_t *out;
    int16_t *nm;
    atomic_int next;
    int count;
} job_t;

static void do_block(job_t *j, int i)
{
    int bx = i % j->bw, by = i / j->bw;
    if (j->format == NTXO_FMT_BC1) {
        float col[64], wt[16];
        int n, m;
        gather_f32(j->w, j->h, j->data, bx, by, j->alpha_mode, col, wt);
        bc1_block_level(col, wt, j->cw, bc1_level(j->quality), j->out + (size_t)i * 8, &n, &m);
        if (j->nm) { j->nm[2 * i] = (int16_t)n; j->nm[2 * i + 1] = (int16_t)m; }
        return;
    }
    uint8_t bgra[64];
    ntxo_gather_u8(j->w, j->h, j->data, bx, by, bgra);
    const int wba = j->alpha_mode == NTXO_ALPHA_TRANSPARENCY;
    /* Context.cpp:1098-1116: BC4/BC5 Production and Highest take
     * ProductionCompressorBC4/BC5 (CompressorDX10.cpp:65-87). */
    void (*channel)(const uint8_t *, uint8_t *) =
        j->quality >= NTXO_QUALITY_PRODUCTION ? ntxo_bc4_channel_optimal : ntxo_bc4_channel;
    if (j->format == NTXO_FMT_BC1A) {
        ntxo_bc1a_block(bgra, wba, j->cw, j->out + (size_t)i * 8);
    } else if (j->format == NTXO_FMT_BC3) {
        ntxo_bc3_block_q(bgra, j->quality, wba, j->cw, j->out + (size_t)i * 16);
    } else if (j->format == NTXO_FMT_BC3N) {
        ntxo_bc3n_block_q(bgra, j->quality, wba, j->out + (size_t)i * 16);
    } else if (j->format == NTXO_FMT_BC4) {
        uint8_t v[16];
        for (int k = 0; k < 16; k++) v[k] = bgra[4 * k + 2];
        channel(v, j->out + (size_t)i * 8);
    } else {
        uint8_t v[16];
        for (int k = 0; k < 16; k++) v[k] = bgra[4 * k + 2];
        channel(v, j->out + (size_t)i * 16);
        for (int k = 0; k < 16; k++) v[k] = bgra[4 * k + 1];
        channel(v, j->out + (size_t)i * 16 + 8);
    }
}

/* ParallelFor.cpp:14-60: workers pull one block at a time from an atomic counter. */
static void *worker(void *arg)
{
    job_t *j = (job_t *)arg;
    for (;;) {
        int i = atomic_fetch_add(&j->next, 1);
        if (i >= j->count) break;
        do_block(j, i);
    }
    return NULL;
}

static void run_job(job_t *j, int threads)
{
    ntxo_init();
    atomic_init(&j->next, 0);
    if (threads <= 1) { worker(j); return; }
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    for (int t = 0; t < threads; t++) pthread_create(&tid[t], NULL, worker, j);
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}

void ntxo_compress_image_q(int format, int quality, int alpha_mode, const float *cw4, int w, int h, const float *planar,
                           uint8_t *out, int threads)
{
    job_t j;
    j.format = format; j.quality = quality; j.alpha_mode = alpha_mode; j.w = w; j.h = h; j.bw = (w + 3) / 4;
    j.cw = cw4; j.data = planar; j.out = out; j.nm = NULL;
    j.count = j.bw * ((h + 3) / 4);
    run_job(&j, threads);
}

void ntxo_compress_image(int format, int alpha_mode, const float *cw4, int w, int h, const float *planar, uint8_t *out, int threads)
{
    job_t j;
    j.format = format; j.quality = NTXO_QUALITY_NORMAL; j.alpha_mode = alpha_mode; j.w = w; j.h = h; j.bw = (w + 3) / 4;
    j.cw = cw4; j.data = planar; j.out = out; j.nm = NULL;
    j.count = j.bw * ((h + 3) / 4);
    run_job(&j, threads);
}

void ntxo_compress_bc1_counts(int alpha_mode, const float *cw4, int w, int h, const float *planar, uint8_t *out, int16_t *nm, int threads)
{
    job_t j;
    j.format = NTXO_FMT_BC1; j.quality = NTXO_QUALITY_NORMAL; j.alpha_mode = alpha_mode; j.w = w; j.h = h;
    j.bw = (w + 3) / 4;
    j.cw = cw4; j.data = planar; j.out = out; j.nm = nm;
    j.count = j.bw * ((h + 3) / 4);
    run_job(&j, threads);
}

/* =========================================================================
 * Mipmaps
 * ========================================================================= */

/* FloatImage.cpp:559-737 fastDownSample (2x2 box; 3-tap polyphase box on odd sizes). */
void ntxo_fast_downsample(int W, int H, const float *src_all, float *dst_all)
{
    const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
    if (W == 1 && H == 1) {
        /* No next level exists (canMakeNextMipmap, Surface.cpp:155-167); the
         * 3-tap 1-D rule below would read past the single texel. */
        for (int c = 0; c < 4; c++) dst_all[c] = src_all[c];
        return;
    }
    for (int c = 0; c < 4; c++) {
        const float *src = src_all + (size_t)c * W * H;
        float *dst = dst_all + (size_t)c * w * h;
        if (W == 1 || H == 1) {
            const int n = w * h;
            if ((W * H) & 1) {
                const float scale = 1.0f / (2 * n + 1);
                for (int x = 0; x < n; x++) {
                    const float w0 = (float)(n - x), w1 = (float)n, w2 = (float)(1 + x);
                    const float *s = src + 2 * x;
                    dst[x] = scale * ((w0 * s[0] + w1 * s[1]) + w2 * s[2]);
                }
            } else {
                for (int x = 0; x < n; x++) dst[x] = 0.5f * (src[2 * x] + src[2 * x + 1]);
            }
        } else if ((W & 1) == 0 && (H & 1) == 0) {
            for (int y = 0; y < h; y++)
                for (int x = 0; x < w; x++) {
                    const float *s = src + (size_t)(2 * y) * W + 2 * x;
                    dst[(size_t)y * w + x] = 0.25f * (((s[0] + s[1]) + s[W]) + s[W + 1]);
                }
        } else if ((W & 1) && (H & 1)) {
            const float scale = 1.0f / (float)(W * H);
            for (int y = 0; y < h; y++) {
                const float v0 = (float)(h - y), v1 = (float)h, v2 = (float)(1 + y);
                const float *s = src + (size_t)(2 * y) * W;
                for (int x = 0; x < w; x++) {
                    const float w0 = (float)(w - x), w1 = (float)w, w2 = (float)(1 + x);
                    float f = 0.0f;
                    f += v0 * ((w0 * s[2 * x] + w1 * s[2 * x + 1]) + w2 * s[2 * x + 2]);
                    f += v1 * ((w0 * s[W + 2 * x] + w1 * s[W + 2 * x + 1]) + w2 * s[W + 2 * x + 2]);
                    f += v2 * ((w0 * s[2 * W + 2 * x] + w1 * s[2 * W + 2 * x + 1]) + w2 * s[2 * W + 2 * x + 2]);
                    dst[(size_t)y * w + x] = f * scale;
                }
            }
        } else if (W & 1) {
            const float scale = 1.0f / (float)(2 * W);
            for (int y = 0; y < h; y++) {
                const float *s = src + (size_t)(2 * y) * W;
                for (int x = 0; x < w; x++) {
                    const float w0 = (float)(w - x), w1 = (float)w, w2 = (float)(1 + x);
                    float f = 0.0f;
                    f += w0 * (s[2 * x] + s[W + 2 * x]);
                    f += w1 * (s[2 * x + 1] + s[W + 2 * x + 1]);
                    f += w2 * (s[2 * x + 2] + s[W + 2 * x + 2]);
                    dst[(size_t)y * w + x] = f * scale;
                }
            }
        } else {
            const float scale = 1.0f / (float)(2 * H);
            for (int y = 0; y < h; y++) {
                const float v0 = (float)(h - y), v1 = (float)h, v2 = (float)(1 + y);
                const float *s = src + (size_t)(2 * y) * W;
                for (int x = 0; x < w; x++) {
                    float f = 0.0f;
                    f += v0 * (s[2 * x] + s[2 * x + 1]);
                    f += v1 * (s[W + 2 * x] + s[W + 2 * x + 1]);
                    f += v2 * (s[2 * W + 2 * x] + s[2 * W + 2 * x + 1]);
                    dst[(size_t)y * w + x] = f * scale;
                }
            }
        }
    }
}

#define NTX_PI ((float)3.1415926535897932384626433833)

/* Filter.cpp:50-83 sincf / bessel0. */
static float sincf_ref(float x)
{
    if (fabs(x) < 0.0001f) return 1.0f + x * x * (-1.0f / 6.0f + x * x * 1.0f / 120.0f);
    return sinf(x) / x;
}
static float bessel0(float x)
{
    const float eps = 1e-6f;
    float xh = 0.5f * x, sum = 1.0f, pw = 1.0f, ds = 1.0f;
    int k = 0;
    while (ds > sum * eps) {
        ++k;
        pw = pw * (xh / k);
        ds = pw * pw;
        sum = sum + ds;
    }
    return sum;
}

typedef struct { int type; float width, alpha, stretch; } filt_t;

/* Filter.cpp:157-175, 257-271 Box/Triangle/Kaiser evaluate. */
static float filt_eval(const filt_t *f, float x)
{
    if (f->type == NTXO_FILTER_BOX) return fabsf(x) <= f->width ? 1.0f : 0.0f;
    if (f->type == NTXO_FILTER_TRIANGLE) { x = fabsf(x); return x < f->width ? f->width - x : 0.0f; }
    const float sv = sincf_ref(NTX_PI * x * f->stretch);
    const float t = x / f->width;
    if ((1 - t * t) >= 0) return sv * bessel0(f->alpha * sqrtf(1 - t * t)) / bessel0(f->alpha);
    return 0;
}

/* Filter.cpp:115-131 sampleBox: double accumulation of 'samples' point samples. */
static float sample_box(const filt_t *f, float x, float scale, int samples)
{
    double sum = 0;
    float is = 1.0f / (float)samples;
    for (int s = 0; s < samples; s++) {
        float p = (x + ((float)s + 0.5f) * is) * scale;
        sum += filt_eval(f, p);
    }
    return (float)(sum * is);
}

int ntxo_kernel_window(int filter, float width, int src_len, int dst_len, float *kwidth)
{
    float scale = (float)dst_len / (float)src_len;
    const float iscale = 1.0f / scale;
    (void)filter;
    float kw = width * iscale;
    if (kwidth) *kwidth = kw;
    return (int)ceilf(kw * 2) + 1;
}

/* Filter.cpp:563-608 PolyphaseKernel. */
void ntxo_polyphase_kernel(int filter, float width, const float *params, int src_len, int dst_len, float *wts)
{
    filt_t f = { filter, width, 4.0f, 1.0f };
    if (filter == NTXO_FILTER_KAISER && params) { f.alpha = params[0]; f.stretch = params[1]; }
    float scale = (float)dst_len / (float)src_len;
    const float iscale = 1.0f / scale;
    int samples = 32;
    if (scale > 1) { samples = 1; scale = 1; }
    const float kw = width * iscale;
    const int ws = (int)ceilf(kw * 2) + 1;
    for (int i = 0; i < dst_len; i++) {
        const float center = (0.5f + i) * iscale;
        const int left = (int)floorf(center - kw);
        float total = 0.0f;
        for (int j = 0; j < ws; j++) {
            const float s = sample_box(&f, left + j - center, scale, samples);
            wts[i * ws + j] = s;
            total += s;
        }
        for (int j = 0; j < ws; j++) wts[i * ws + j] /= total;
    }
}

/* FloatImage.h:298-351 wrap helpers. */
static int wrap_index(int x, int w, int wm)
{
    if (wm == NTXO_WRAP_CLAMP) return x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
    if (wm == NTXO_WRAP_REPEAT) return x >= 0 ? x % w : (x + 1) % w + w - 1;
    if (w == 1) x = 0;
    x = abs(x);
    while (x >= w) x = abs(w + w - x - 2);
    return x;
}

/* FloatImage.cpp:761-808 resize + 1115-1176 applyKernelX/Y: X pass into a
 * w x H temporary, then Y pass; taps summed in order without fusion. */
void ntxo_downsample(int filter, float width, const float *params, int wrap, int W, int H, const float *src_all, float *dst_all)
{
    const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
    const int wsx = ntxo_kernel_window(filter, width, W, w, NULL);
    const int wsy = ntxo_kernel_window(filter, width, H, h, NULL);
    float *kx = (float *)malloc(sizeof(float) * (size_t)wsx * w);
    float *ky = (float *)malloc(sizeof(float) * (size_t)wsy * h);
    float *tmp = (float *)malloc(sizeof(float) * (size_t)w * H);
    ntxo_polyphase_kernel(filter, width, params, W, w, kx);
    ntxo_polyphase_kernel(filter, width, params, H, h, ky);
    float kwx, kwy;
    ntxo_kernel_window(filter, width, W, w, &kwx);
    ntxo_kernel_window(filter, width, H, h, &kwy);
    const float isx = 1.0f / ((float)w / (float)W);
    const float isy = 1.0f / ((float)h / (float)H);
    for (int c = 0; c < 4; c++) {
        const float *src = src_all + (size_t)c * W * H;
        float *dst = dst_all + (size_t)c * w * h;
        for (int y = 0; y < H; y++)
            for (int i = 0; i < w; i++) {
                const float center = (0.5f + i) * isx;
                const int left = (int)floorf(center - kwx);
                float sum = 0;
                for (int j = 0; j < wsx; j++) sum += kx[i * wsx + j] * src[(size_t)y * W + wrap_index(left + j, W, wrap)];
                tmp[(size_t)y * w + i] = sum;
            }
        for (int x = 0; x < w; x++)
            for (int i = 0; i < h; i++) {
                const float center = (0.5f + i) * isy;
                const int left = (int)floorf(center - kwy);
                float sum = 0;
                for (int j = 0; j < wsy; j++) sum += ky[i * wsy + j] * tmp[(size_t)wrap_index(left + j, H, wrap) * w + x];
                dst[(size_t)i * w + x] = sum;
            }
    }
    free(kx); free(ky); free(tmp);
}

/* Surface.cpp:1121-1141, 1344-1406 (AlphaMode_None, depth 1). */
void ntxo_build_next_mipmap(int filter, float width, const float *params, int wrap, int W, int H, const float *src, float *dst)
{
    if (filter == NTXO_FILTER_BOX && width == 0.5f) ntxo_fast_downsample(W, H, src, dst);
    else ntxo_downsample(filter, width, params, wrap, W, H, src, dst);
}

/* Surface.cpp:1562-1592 fromSrgb on the R,G,B planes. */
void ntxo_srgb_to_linear(size_t n, float *d)
{
    for (size_t i = 0; i < n; i++) {
        float f = d[i];
        if (f < 0.0f) f = 0.0f;
        else if (f < 0.04045f) f = f / 12.92f;
        else if (f <= 1.0f) f = powf((f + 0.055f) / 1.055f, 2.4f);
        else f = 1.0f;
        d[i] = f;
    }
}

/* ---------------------------------------------------------------- gamma --
 * Gamma.cpp:30-300: exponent tables.  Entry k is 2^((k - 127) p) for the
 * sign-and-exponent field k of the input, rounded to float; 0 for zero and
 * subnormal inputs (k = 0) and for negative ones (k >= 256); +inf for inf and
 * NaN (k = 255).  Rebuilt here from that closed form. */
static float g_pow_5_11[512], g_pow_11_5[512];
static int g_gamma_ready;

static void gamma_tables(void)
{
    if (g_gamma_ready) return;
    for (int k = 0; k < 512; k++) {
        if (k == 0 || k >= 256) { g_pow_5_11[k] = 0.0f; g_pow_11_5[k] = 0.0f; }
        else if (k == 255) { g_pow_5_11[k] = INFINITY; g_pow_11_5[k] = INFINITY; }
        else {
            g_pow_5_11[k] = (float)pow(2.0, (double)(k - 127) * 5.0 / 11.0);
            g_pow_11_5[k] = (float)pow(2.0, (double)(k - 127) * 11.0 / 5.0);
        }
    }
    g_gamma_ready = 1;
}

void ntxo_gamma_table(int which, float *out512)
{
    gamma_tables();
    memcpy(out512, which ? g_pow_11_5 : g_pow_5_11, sizeof g_pow_5_11);
}

/* Mantissa of x with the exponent of 1.0: a float in [1, 2). */
static float unit_mantissa(float x, unsigned *k)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    *k = u >> 23;
    u = (u & ((1u << 23) - 1u)) | (127u << 23);
    float m;
    memcpy(&m, &u, 4);
    return m;
}

/* Gamma.cpp:311-330 _powf_5_11: table(exponent) * quartic(mantissa). */
float ntxo_powf_5_11(float x)
{
    gamma_tables();
    unsigned k;
    const float m = unit_mantissa(x, &k);
    const float pe = g_pow_5_11[k];
    const float pm = (((-0.0110083047f * m + 0.0905038750f) * m - 0.324697506f) * m + 0.876040946f) * m + 0.369160989f;
    return pe * pm;
}

/* Gamma.cpp:338-357 _powf_11_5. */
float ntxo_powf_11_5(float x)
{
    gamma_tables();
    unsigned k;
    const float m = unit_mantissa(x, &k);
    const float pe = g_pow_11_5[k];
    const float pm = (((-0.00916587552f * m + 0.119315466f) * m + 1.01847068f) * m - 0.158338739f) * m + 0.0297184721f;
    return pe * pm;
}

/* FloatImage.cpp:286-297 exponentiate: powf(nv::max(0, x), power). */
static void exponentiate(size_t n, float *d, float power)
{
    for (size_t i = 0; i < n; i++) {
        const float x = d[i] < 0.0f ? 0.0f : d[i];
        d[i] = powf(x, power);
    }
}

/* FloatImage.cpp:258-270 toLinear. */
void ntxo_to_linear(size_t n, float *d, float gamma)
{
    if (gamma == 2.2f) { for (size_t i = 0; i < n; i++) d[i] = ntxo_powf_11_5(d[i]); }
    else exponentiate(n, d, gamma);
}

/* FloatImage.cpp:272-284 toGamma. */
void ntxo_to_gamma(size_t n, float *d, float gamma)
{
    if (gamma == 2.2f) { for (size_t i = 0; i < n; i++) d[i] = ntxo_powf_5_11(d[i]); }
    else exponentiate(n, d, 1.0f / gamma);
}

/* Context.cpp:329-336: scaleBias(2, -2*127/255) (FloatImage.cpp:231-242),
 * normalizeSafe(v, 0, 0) (Vector.inl:549-595), scaleBias(0.5, 0.5). */
void ntxo_normalize_normals(size_t n, float *p)
{
    const float bias = -2.0f * 127.0f / 255.0f;
    for (size_t i = 0; i < n; i++) {
        float x = 2.0f * p[i] + bias, y = 2.0f * p[i + n] + bias, z = 2.0f * p[i + 2 * n] + bias;
        const float l = sqrtf(x * x + y * y + z * z);
        if (fabsf(l) <= 0.0f) { x = y = z = 0.0f; }
        else { const float s = 1.0f / l; x = x * s; y = y * s; z = z * s; }
        p[i] = 0.5f * x + 0.5f; p[i + n] = 0.5f * y + 0.5f; p[i + 2 * n] = 0.5f * z + 0.5f;
    }
}

/* ------------------------------------------------------ decode and errors --
 * Surface::setImage2D (Surface.cpp:1022-1107) for BC1/BC3/BC4/BC5: the nvimage
 * block decoders (BlockDXT.cpp:43-98, 200-217, 336-432, 475-566) to 8-bit
 * RGBA, then float(c) * 1.0f / 255.0f per channel.  d3d9 only affects BC4/BC5
 * (the BC1/BC3 colour decoder is called with d3d9 = false for both D3D
 * decoders). */
static void decode_alpha_palette(const uint8_t *b, int d3d9, int pal[8])
{
    const int a0 = b[0], a1 = b[1];
    pal[0] = a0; pal[1] = a1;
    if (a0 > a1) {
        const int bias = d3d9 ? 3 : 0;
        for (int i = 2; i < 8; i++) pal[i] = ((8 - i) * a0 + (i - 1) * a1 + bias) / 7;
    } else {
        const int bias = d3d9 ? 2 : 0;
        for (int i = 2; i < 6; i++) pal[i] = ((6 - i) * a0 + (i - 1) * a1 + bias) / 5;
        pal[6] = 0; pal[7] = 255;
    }
}

static int alpha_index(const uint8_t *b, int t)
{
    uint64_t u = 0;
    for (int k = 0; k < 8; k++) u |= (uint64_t)b[k] << (8 * k);
    return (int)((u >> (16 + 3 * t)) & 7);
}

void ntxo_decode_image(int format, int d3d9, const uint8_t *blocks, int w, int h, float *planar)
{
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const size_t n = (size_t)w * h;
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            const uint8_t *b = blocks + (size_t)(by * bw + bx) * ((format == NTXO_FMT_BC1 || format == NTXO_FMT_BC4) ? 8 : 16);
            uint8_t rgba[64];
            if (format == NTXO_FMT_BC1) {
                ntxo_bc1_decode(b, rgba);
            } else if (format == NTXO_FMT_BC3) {
                int pal[8];
                ntxo_bc1_decode(b + 8, rgba);
                decode_alpha_palette(b, 0, pal);
                for (int t = 0; t < 16; t++) rgba[4 * t + 3] = (uint8_t)pal[alpha_index(b, t)];
            } else if (format == NTXO_FMT_BC4) {
                int pal[8];
                decode_alpha_palette(b, d3d9, pal);
                for (int t = 0; t < 16; t++) {
                    const uint8_t v = (uint8_t)pal[alpha_index(b, t)];
                    rgba[4 * t] = rgba[4 * t + 1] = rgba[4 * t + 2] = v;
                    rgba[4 * t + 3] = 255;
                }
            } else {
                int px[8], py[8];
                decode_alpha_palette(b, d3d9, px);
                decode_alpha_palette(b + 8, d3d9, py);
                for (int t = 0; t < 16; t++) {
                    rgba[4 * t] = (uint8_t)px[alpha_index(b, t)];
                    rgba[4 * t + 1] = (uint8_t)py[alpha_index(b + 8, t)];
                    rgba[4 * t + 2] = 0;
                    rgba[4 * t + 3] = 255;
                }
            }
            for (int t = 0; t < 16; t++) {
                const int x = 4 * bx + (t & 3), y = 4 * by + (t >> 2);
                if (x >= w || y >= h) continue;
                const size_t o = (size_t)y * w + x;
                for (int c = 0; c < 4; c++) planar[o + c * n] = (float)rgba[4 * t + c] * 1.0f / 255.0f;
            }
        }
}

/* ErrorMetric.cpp:13-46 rmsColorError and :50-69 rmsAlphaError, summed one
 * texel after another in double.  metric 0: colour, 1: colour weighted by the
 * reference's alpha squared, 2: alpha. */
float ntxo_rms_error(int metric, const float *ref, const float *img, size_t n)
{
    double mse = 0;
    for (size_t i = 0; i < n; i++) {
        if (metric == 2) {
            float a = img[i + 3 * n] - ref[i + 3 * n];
            mse += a * a;
        } else {
            float r = ref[i] - img[i], g = ref[i + n] - img[i + n], b = ref[i + 2 * n] - img[i + 2 * n];
            float a = 1;
            if (metric == 1) a = ref[i + 3 * n] * ref[i + 3 * n];
            mse += (r * r) * a;
            mse += (g * g) * a;
            mse += (b * b) * a;
        }
    }
    return (float)sqrt(mse / (unsigned)n);
}

/* FloatImage.cpp:201-219 normalize(0) alone (the KTX branch of process). */
void ntxo_normalize_vectors(size_t n, float *p)
{
    for (size_t i = 0; i < n; i++) {
        float x = p[i], y = p[i + n], z = p[i + 2 * n];
        const float l = sqrtf(x * x + y * y + z * z);
        if (fabsf(l) <= 0.0f) { x = y = z = 0.0f; }
        else { const float s = 1.0f / l; x = x * s; y = y * s; z = z * s; }
        p[i] = x; p[i + n] = y; p[i + 2 * n] = z;
    }
}
