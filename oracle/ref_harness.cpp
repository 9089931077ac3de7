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

}
