// ntx_common.h -- shared definitions for the gfx950 BCn / mipmap kernels.
//
// All kernels are compiled with -ffp-contract=off: the reference CPU path has
// no fused multiply-add, and every BCn decision (partition winner, 565
// rounding, index choice) must see the same float roundings it does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ntx_hip.h"

namespace ntx {

// The last levels of a mip chain built by one launch (launch_mip_tail): the
// first source level (W x H, at most kTailSrc texels per plane), then per
// produced level its destination and, for the polyphase filters, its X and
// Y tap tables (as PolyDev holds them); box: FloatImage::fastDownSample.
// One image of the multi-image two-pass BC1 encoder (Level8): its planar
// FP32 RGBA texels, its blocks (8 bytes each), and how they are weighted.
struct Bc1Image {
    const float *src;
    void *out;
    int w, h;
    int alpha_mode;
    int rows16;  // src 16-byte aligned and w % 4 == 0: rows read as float4
    float cw[3];
    int pad;
};

constexpr int kTailMaxLevels = 16;
constexpr int kTailSrc = 8192;
// Packed taps of the tail's levels: (weight, wrapped source offset as int
// bits) pairs, for each level in order whose X and Y taps still fit in
// kTailTaps entries: the w * wx X taps (offset within a source row), then
// the h * wy Y taps (offset of a temporary row, already times w).
constexpr int kTailTaps = 4096;
struct MipTail {
    int W, H, levels, box;
    float *dst[kTailMaxLevels];
    const float *kx[kTailMaxLevels], *ky[kTailMaxLevels];
    const int *lx[kTailMaxLevels], *ly[kTailMaxLevels];
    int wx[kTailMaxLevels], wy[kTailMaxLevels];
    const float *packed;  // 2 * npacked floats, built on the host
    int npacked;
};

// Tables shared by host set-up and kernels.  Layout is fixed here so the host
// builders and device code agree.
struct Tables {
    // icbc cluster tables (icbc.h:1906-2012): packed c0 | c1<<8 | c2<<16.
    uint32_t four[976];
    int four_total[16];
    uint32_t three[160];
    int three_total[16];
    // icbc single-colour tables for the D3D10 decoder (icbc.h:3166-3262): mx | mn<<8.
    uint16_t match5[256];
    uint16_t match6[256];
    // nvtt OMatch5/6 (SingleColorLookup.cpp:34-89): mx | mn<<8.
    uint16_t omatch5[256];
    uint16_t omatch6[256];
    // 565 bit-expansion midpoints as tabulated by icbc (icbc.h:1513-1526).
    float mid5[32];
    float mid6[64];
    // squish Compress4 loop order (weightedclusterfit.cpp:495-553) for a point
    // count n: entries [sq_off[n], sq_off[n+1]) hold c0 | c1<<8 | c2<<16.
    int sq_off[18];
    uint32_t sq[4845];
    // For 16 points of weight 1 (an unweighted block with 16 distinct
    // colours) the cluster weights of candidate (c0, c1, c2) are the counts,
    // so the fit's alpha2, beta2, alphabeta and 1 / (alpha2 beta2 - ab^2)
    // (weightedclusterfit.cpp:511-518) depend on the candidate only; here
    // they are, in loop order, computed with the kernel's own operations.
    float sq16[969][4];
    // The same terms for any integer cluster weights (an unweighted block:
    // weights are point multiplicities, 16 in all): entry
    // (w0 * 17 + w1) * 17 + w2 for w0 + w1 + w2 <= 16, others zero.
    float sqw[17 * 17 * 17][4];
    // icbc's cluster-weight terms (alpha2, beta2, alphabeta, 1 / (alpha2
    // beta2 - alphabeta^2), icbc.h:2501-2511 and :2201-2208) of a block whose
    // cluster weights are integers summing to 16 (every texel of weight 1):
    // bc1w4 by the prefix counts (W0 * 17 + W1) * 17 + W2 at the candidate's
    // c0 <= c1 <= c2, bc1w3 by W0 * 17 + W1; entries with W0 > W1 or W1 > W2
    // are zero.  Computed with the kernel's own operations on exact values.
    float bc1w4[17 * 17 * 17][4];
    float bc1w3[17 * 17][4];
    // Exponent tables of the gamma-2.2 transforms (Gamma.cpp:33-300): entry k
    // (a float's sign-and-exponent field) is 2^((k - 127) * 5/11), resp.
    // 2^((k - 127) * 11/5), as a float; 0 for k == 0 and k >= 256, +inf for 255.
    float pow_5_11[512];
    float pow_11_5[512];
    // nvtt OMatchAlpha5/6 (SingleColorLookup.cpp:34-89 with the three-colour
    // midpoint (hi + lo) / 2): mx | mn<<8, for BC1a blocks of one colour with
    // transparent texels (OptimalCompress::compressDXT1a).
    uint16_t omatcha5[256];
    uint16_t omatcha6[256];
    // k / 255.0f for k = 0..255 (icbc's color_to_vector3, icbc.h:1611-1613).
    float div255[256];
    // reduce_colors' running weighted mean (icbc.h:1638-1683) of 16 equal
    // colour channels of value div255[k], weights 1, in texel order: what a
    // uniform 4x4 block of an 8-bit source averages to.
    float mean16[256];
};

void build_tables(Tables *t);

// Polyphase weights (Filter.cpp:563-608) built on the host.
struct Kernel1D {
    int length = 0;      // output samples
    int window = 0;      // taps per output
    float width = 0.0f;  // kernel half-width in source samples
    float iscale = 0.0f; // source samples per output sample
    float *weights = nullptr; // length * window
};
void build_polyphase(int filter, float width, const float *params, int src_len, int dst_len, Kernel1D *k);
void free_polyphase(Kernel1D *k);
int polyphase_window(float width, int src_len, int dst_len);

}  // namespace ntx

#ifdef __HIP_DEVICE_COMPILE__
#define NTX_DEVICE 1
#endif

// Synchronisation of the lanes of one wavefront: orders LDS traffic between
// lanes of the same wave (a compiler fence; the hardware keeps one wave's LDS
// operations in order).
__device__ __forceinline__ void ntx_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
