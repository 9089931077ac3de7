// ntx_tables.cpp -- host-built lookup tables uploaded to the device once per
// context, and the polyphase filter weights built per mip level.
//
// Built with -ffp-contract=off like the kernels: the polyphase weights feed
// the mip kernels and must round exactly as Filter.cpp's do.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gamma.hpp"
#include "ntx_common.h"

namespace ntx {

static int expand5(int v) { return (v << 3) | (v >> 2); }
static int expand6(int v) { return (v << 2) | (v >> 4); }

void build_tables(Tables *t)
{
    std::memset(t, 0, sizeof(*t));

    // icbc.h:1906-2012.  Level k of the table holds the cumulative partitions
    // (c0, c0+c1, k): every partition of fewer points already appears at an
    // earlier level.  The tail replicates the last entry 8 times.
    int i = 0;
    for (int k = 1; k <= 16; k++) {
        for (int a = 0; a <= k; a++)
            for (int b = a; b <= k; b++) t->four[i++] = uint32_t(a) | uint32_t(b) << 8 | uint32_t(k) << 16;
        t->four_total[k - 1] = i;
    }
    for (int r = 0; r < 8; r++) t->four[968 + r] = t->four[967];
    i = 0;
    for (int k = 1; k <= 16; k++) {
        for (int a = 0; a <= k; a++) t->three[i++] = uint32_t(a) | uint32_t(k) << 8;
        t->three_total[k - 1] = i;
    }
    for (int r = 0; r < 8; r++) t->three[152 + r] = t->three[151];

    // icbc.h:3166-3262, Decoder_D3D10: worst of the AMD and NVIDIA interpolators.
    for (int v = 0; v < 256; v++) {
        int best5 = 25600, best6 = 25600;
        for (int lo = 0; lo < 32; lo++)
            for (int hi = 0; hi < 32; hi++) {
                int el = expand5(lo), eh = expand5(hi);
                int amd = std::abs(((43 * eh + 21 * el + 32) >> 6) - v);
                int nv = std::abs(((2 * hi + lo) * 22) / 8 - v);
                int e = amd > nv ? amd : nv;
                if (e < best5) { best5 = e; t->match5[v] = uint16_t(hi | lo << 8); }
            }
        for (int lo = 0; lo < 64; lo++)
            for (int hi = 0; hi < 64; hi++) {
                int el = expand6(lo), eh = expand6(hi), d = eh - el;
                int amd = std::abs(((43 * eh + 21 * el + 32) >> 6) - v);
                int nv = std::abs((256 * el + d / 4 + 128 + d * 80) / 256 - v);
                int e = amd > nv ? amd : nv;
                if (e < best6) { best6 = e; t->match6[v] = uint16_t(hi | lo << 8); }
            }
    }

    // SingleColorLookup.cpp:34-89 (colour tables): 100*|lerp13 - v| + 3*|hi-lo|.
    for (int v = 0; v < 256; v++) {
        int b5 = 25600, b6 = 25600;
        for (int lo = 0; lo < 32; lo++)
            for (int hi = 0; hi < 32; hi++) {
                int e = std::abs((expand5(hi) * 2 + expand5(lo)) / 3 - v) * 100 + std::abs(hi - lo) * 3;
                if (e < b5) { b5 = e; t->omatch5[v] = uint16_t(hi | lo << 8); }
            }
        for (int lo = 0; lo < 64; lo++)
            for (int hi = 0; hi < 64; hi++) {
                int e = std::abs((expand6(hi) * 2 + expand6(lo)) / 3 - v) * 100 + std::abs(hi - lo) * 3;
                if (e < b6) { b6 = e; t->omatch6[v] = uint16_t(hi | lo << 8); }
            }
    }
    // The same search with the three-colour midpoint (OMatchAlpha5/6).
    for (int v = 0; v < 256; v++) {
        int b5 = 25600, b6 = 25600;
        for (int lo = 0; lo < 32; lo++)
            for (int hi = 0; hi < 32; hi++) {
                int e = std::abs((expand5(hi) + expand5(lo)) / 2 - v) * 100 + std::abs(hi - lo) * 3;
                if (e < b5) { b5 = e; t->omatcha5[v] = uint16_t(hi | lo << 8); }
            }
        for (int lo = 0; lo < 64; lo++)
            for (int hi = 0; hi < 64; hi++) {
                int e = std::abs((expand6(hi) + expand6(lo)) / 2 - v) * 100 + std::abs(hi - lo) * 3;
                if (e < b6) { b6 = e; t->omatcha6[v] = uint16_t(hi | lo << 8); }
            }
    }

    // icbc.h:1516-1526: midpoints between consecutive bit-expanded levels,
    // printed to six decimals in the reference table.
    char buf[64];
    for (int k = 0; k < 31; k++) {
        float m = (float(expand5(k)) / 255.0f + float(expand5(k + 1)) / 255.0f) * 0.5f;
        std::snprintf(buf, sizeof buf, "%f", double(m));
        t->mid5[k] = std::strtof(buf, nullptr);
    }
    t->mid5[31] = FLT_MAX;
    for (int k = 0; k < 63; k++) {
        float m = (float(expand6(k)) / 255.0f + float(expand6(k + 1)) / 255.0f) * 0.5f;
        std::snprintf(buf, sizeof buf, "%f", double(m));
        t->mid6[k] = std::strtof(buf, nullptr);
    }
    t->mid6[63] = FLT_MAX;

    // weightedclusterfit.cpp:495-553: (c0, c1, c2) with c0+c1+c2 <= n, c0
    // outermost, in loop order, for each point count n.
    i = 0;
    for (int n = 0; n <= 16; n++) {
        t->sq_off[n] = i;
        for (int a = 0; a <= n; a++)
            for (int b = 0; b <= n - a; b++)
                for (int c = 0; c <= n - a - b; c++) t->sq[i++] = uint32_t(a) | uint32_t(b) << 8 | uint32_t(c) << 16;
    }
    t->sq_off[17] = i;

    // Cluster-weight terms for integer weights (a, b, c) of 16 in all: the
    // operations of the kernel's sq_weights on exact integer values (float
    // arithmetic, no contraction, IEEE division), hence the same values.
    // sq16 lists them in the n = 16 loop order (unit weights: the weights
    // are the counts), sqw by (a, b, c).
    std::memset(t->sqw, 0, sizeof t->sqw);
    i = 0;
    for (int a = 0; a <= 16; a++)
        for (int b = 0; b <= 16 - a; b++)
            for (int c = 0; c <= 16 - a - b; c++, i++) {
                const float x0 = float(a), x1 = float(b), x2 = float(c), sum = 16.0f;
                const float w3 = ((sum - x0) - x1) - x2;
                const float alpha2 = (x0 + x1 * (4.0f / 9.0f)) + x2 * (1.0f / 9.0f);
                const float beta2 = (w3 + x2 * (4.0f / 9.0f)) + x1 * (1.0f / 9.0f);
                const float ab = (x1 + x2) * (2.0f / 9.0f);
                const float factor = 1.0f / (alpha2 * beta2 - ab * ab);
                const float v[4] = { alpha2, beta2, ab, factor };
                for (int j = 0; j < 4; j++) {
                    t->sq16[i][j] = v[j];
                    t->sqw[(a * 17 + b) * 17 + c][j] = v[j];
                }
            }

    // icbc cluster-weight terms for integer weights summing to 16, by the
    // prefix weights (W0, W1, W2) the kernel reads from its prefix-sum rows:
    // the operations of ntx_bc1::cluster_terms (icbc.h:2501-2511, 2201-2208).
    std::memset(t->bc1w4, 0, sizeof t->bc1w4);
    std::memset(t->bc1w3, 0, sizeof t->bc1w3);
    for (int a = 0; a <= 16; a++)
        for (int b = a; b <= 16; b++) {
            {
                const float x0w = float(a), x1w = float(b), sumw = 16.0f;
                const float w2 = sumw - x1w;
                const float w1 = x1w - x0w;
                const float ab = w1 * 0.25f;
                const float alpha2 = x0w + ab;
                const float beta2 = w2 + ab;
                const float v[4] = { alpha2, beta2, ab, 1.0f / (alpha2 * beta2 - ab * ab) };
                for (int j = 0; j < 4; j++) t->bc1w3[a * 17 + b][j] = v[j];
            }
            for (int c = b; c <= 16; c++) {
                const float x0w = float(a), x1w = float(b), x2w = float(c), sumw = 16.0f;
                const float w3 = sumw - x2w;
                const float w2 = x2w - x1w;
                const float w1 = x1w - x0w;
                const float w0 = x0w;
                const float alpha2 = w2 * (1.0f / 9.0f) + (w1 * (4.0f / 9.0f) + w0);
                const float beta2 = w1 * (1.0f / 9.0f) + (w2 * (4.0f / 9.0f) + w3);
                const float ab = (w1 + w2) * (2.0f / 9.0f);
                const float v[4] = { alpha2, beta2, ab, 1.0f / (alpha2 * beta2 - ab * ab) };
                for (int j = 0; j < 4; j++) t->bc1w4[(a * 17 + b) * 17 + c][j] = v[j];
            }
        }

    for (int k = 0; k < 256; k++) t->div255[k] = float(k) / 255.0f;
    for (int k = 0; k < 256; k++) {
        const float c = t->div255[k];
        float m = c, w = 1.0f;
        for (int i = 1; i < 16; i++) {
            const float s1 = w + 1.0f;
            m = (m * w + c) / s1;
            w = s1;
        }
        t->mean16[k] = m;
    }

    // Gamma-2.2 exponent tables (Gamma.cpp:30-300) from their closed form.
    for (int k = 0; k < 512; k++) {
        t->pow_5_11[k] = ntx_gamma::table_entry(k, 5.0, 11.0);
        t->pow_11_5[k] = ntx_gamma::table_entry(k, 11.0, 5.0);
    }
}

// ---------------------------------------------------------------- filters --

static const float kPi = float(3.1415926535897932384626433833);

// Filter.cpp:50-58 sincf.
static float sinc(float x)
{
    if (std::fabs(x) < 0.0001f) return 1.0f + x * x * (-1.0f / 6.0f + x * x * 1.0f / 120.0f);
    return std::sin(x) / x;
}

// Filter.cpp:63-82 bessel0 (series until the term drops below 1e-6 of the sum).
static float bessel_i0(float x)
{
    float half = 0.5f * x, sum = 1.0f, term = 1.0f, sq = 1.0f;
    int k = 0;
    while (sq > sum * 1e-6f) {
        ++k;
        term = term * (half / k);
        sq = term * term;
        sum = sum + sq;
    }
    return sum;
}

struct FilterFn {
    int type;
    float width, alpha, stretch;
    // Filter.cpp:157-175, 257-263.
    float operator()(float x) const
    {
        if (type == NTX_FILTER_BOX) return std::fabs(x) <= width ? 1.0f : 0.0f;
        if (type == NTX_FILTER_TRIANGLE) {
            x = std::fabs(x);
            return x < width ? width - x : 0.0f;
        }
        const float s = sinc(kPi * x * stretch);
        const float t = x / width;
        if ((1 - t * t) >= 0) return s * bessel_i0(alpha * std::sqrt(1 - t * t)) / bessel_i0(alpha);
        return 0.0f;
    }
};

int polyphase_window(float width, int src_len, int dst_len)
{
    const float scale = float(dst_len) / float(src_len);
    const float iscale = 1.0f / scale;
    return int(std::ceil(width * iscale * 2)) + 1;
}

// Filter.cpp:563-608 PolyphaseKernel: per output sample, box-integrate the
// filter over each source texel with 32 point samples (double accumulator,
// Filter.cpp:115-131), then normalise the row.
void build_polyphase(int filter, float width, const float *params, int src_len, int dst_len, Kernel1D *k)
{
    FilterFn f{ filter, width, 4.0f, 1.0f };
    if (filter == NTX_FILTER_KAISER && params) { f.alpha = params[0]; f.stretch = params[1]; }
    float scale = float(dst_len) / float(src_len);
    const float iscale = 1.0f / scale;
    int samples = 32;
    if (scale > 1) { samples = 1; scale = 1; }
    k->length = dst_len;
    k->iscale = iscale;
    k->width = width * iscale;
    k->window = int(std::ceil(k->width * 2)) + 1;
    k->weights = new float[size_t(k->window) * dst_len];
    const float inv_samples = 1.0f / float(samples);
    for (int i = 0; i < dst_len; i++) {
        const float center = (0.5f + i) * iscale;
        const int left = int(std::floor(center - k->width));
        float total = 0.0f;
        float *row = k->weights + size_t(i) * k->window;
        for (int j = 0; j < k->window; j++) {
            const float x = left + j - center;
            double acc = 0;
            for (int s = 0; s < samples; s++) acc += f((x + (float(s) + 0.5f) * inv_samples) * scale);
            row[j] = float(acc * inv_samples);
            total += row[j];
        }
        for (int j = 0; j < k->window; j++) row[j] /= total;
    }
}

void free_polyphase(Kernel1D *k)
{
    delete[] k->weights;
    k->weights = nullptr;
}

}  // namespace ntx
