// ref_harness.cpp -- exposes the reference's own BC1 (icbc.h) and BC3-colour
// (nvsquish WeightedClusterFit) encoders through a C ABI so the tests can pin
// the oracle against them.  TEST INFRASTRUCTURE ONLY.
//
// Nothing from the reference is copied here: the Makefile compiles the
// reference sources where they lie (REF/src/nvtt/icbc.h and
// REF/src/nvtt/squish/*.cpp) with the pinned settings ICBC_SIMD=2 (SSE4.1)
// and SQUISH_USE_SSE=0 (scalar).  Output goes to oracle/_ref/ only.
#include <stdint.h>

#define ICBC_IMPLEMENTATION
#include "icbc.h"

#include "squish/colourset.h"
#include "squish/weightedclusterfit.h"

extern "C" {

// Context.cpp:72 -- nvtt initialises icbc with the D3D10 decoder.
void ntxref_init(void) { icbc::init_dxt1(icbc::Decoder_D3D10); }

// BlockCompressor.cpp:219-222 -- CompressorDXT1::compressBlock at Quality_Normal.
float ntxref_bc1_block(const float *colors, const float *weights, const float *cw, uint8_t *out)
{
    return icbc::compress_dxt1(icbc::Quality_Level8, colors, weights, cw, true, true, out);
}

// BlockCompressor.cpp:211-222 -- CompressorDXT1::compressBlock at any nvtt
// quality (qualityLevel: Fastest -> Quality_Fast, Production ->
// Quality_Max, otherwise Quality_Default).
float ntxref_bc1_block_q(const float *colors, const float *weights, const float *cw, int quality, uint8_t *out)
{
    const icbc::Quality q = quality == 0 ? icbc::Quality_Fast : quality == 2 ? icbc::Quality_Max : icbc::Quality_Default;
    return icbc::compress_dxt1(q, colors, weights, cw, true, true, out);
}

void ntxref_bc1_decode(const uint8_t *blk, uint8_t *rgba)
{
    icbc::decode_dxt1(blk, rgba, icbc::Decoder_D3D10);
}

// CompressorDX9.cpp:163-174 -- the squish branch of CompressorDXT5::compressBlock.
void ntxref_bc3_color_squish(const uint8_t *bgra, int weight_by_alpha, const float *cw, uint8_t *out8)
{
    nvsquish::WeightedClusterFit fit;
    fit.SetMetric(cw[0], cw[1], cw[2]);
    int flags = weight_by_alpha ? nvsquish::kWeightColourByAlpha : 0;
    nvsquish::ColourSet colours(bgra, flags);
    fit.SetColourSet(&colours, 0);
    for (int i = 0; i < 8; i++) out8[i] = 0;
    fit.Compress(out8);
}

// CompressorDX9.cpp:98-109 -- the squish branch of CompressorDXT1a::compressBlock.
void ntxref_bc1a_squish(const uint8_t *bgra, int weight_by_alpha, const float *cw, uint8_t *out8)
{
    nvsquish::WeightedClusterFit fit;
    fit.SetMetric(cw[0], cw[1], cw[2]);
    int flags = nvsquish::kDxt1 | (weight_by_alpha ? nvsquish::kWeightColourByAlpha : 0);
    nvsquish::ColourSet colours(bgra, flags);
    fit.SetColourSet(&colours, nvsquish::kDxt1);
    for (int i = 0; i < 8; i++) out8[i] = 0;
    fit.Compress(out8);
}

}

// Whole-image BC1 with the reference encoder, for the CPU baseline.  The
// gather restates FloatColorCompressorTask (BlockCompressor.cpp:116-166,
// AlphaMode_None); workers pull one block at a time from an atomic counter
// like nv::ParallelFor (ParallelFor.cpp:14-60).
#include <atomic>
#include <thread>
#include <vector>

extern "C" void ntxref_compress_bc1_image(const float *planar, int w, int h, uint8_t *out, int threads)
{
    const int bw = (w + 3) / 4, bh = (h + 3) / 4, count = bw * bh;
    const size_t plane = size_t(w) * size_t(h);
    const float cw[3] = { 1.0f, 1.0f, 1.0f };
    std::atomic<int> next(0);
    auto work = [&]() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= count) return;
            const int bx = i % bw, by = i / bw;
            float colors[64], weights[16];
            for (int y = 0; y < 4; y++)
                for (int x = 0; x < 4; x++) {
                    const int k = 4 * y + x, X = 4 * bx + x, Y = 4 * by + y;
                    if (X < w && Y < h) {
                        const size_t s = size_t(Y) * w + X;
                        colors[4 * k] = planar[s];
                        colors[4 * k + 1] = planar[s + plane];
                        colors[4 * k + 2] = planar[s + 2 * plane];
                        colors[4 * k + 3] = planar[s + 3 * plane];
                        weights[k] = 1.0f;
                    } else {
                        colors[4 * k] = colors[4 * k + 1] = colors[4 * k + 2] = colors[4 * k + 3] = 0.0f;
                        weights[k] = 0.0f;
                    }
                }
            icbc::compress_dxt1(icbc::Quality_Level8, colors, weights, cw, true, true, out + size_t(i) * 8);
        }
    };
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
}
