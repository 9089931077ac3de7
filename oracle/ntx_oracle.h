/*
 * ntx_oracle.h -- CPU restatement of NVTT 2.1.2's BC1/BC3/BC4/BC5 block
 * compressors and mipmap filters.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product (libntx_hip.so / libnvtt_amd.so) never links or
 * calls it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * pinned to the reference build the project treats as authoritative:
 *   - BC1: icbc v1.05 at Quality_Level8 with ICBC_SIMD=2 (SSE4.1): no FMA,
 *     round-half-even, exact division, 4-lane tie-break.
 *   - BC3 colour: nvsquish WeightedClusterFit with SQUISH_USE_SSE=0 (scalar).
 *   - BC4/BC5 and BC3 alpha: QuickCompress::compressDXT5A (Fastest, Normal)
 *     and OptimalCompress::compressDXT5A with its first error seed taken
 *     from the (max, min) endpoints (Production, Highest).
 *   - BC3/BC3n Fastest: QuickCompress::compressDXT5; BC3n Highest colour:
 *     OptimalCompress::compressDXT1G; BC1a: CompressorDXT1a.
 *   - Mips: FloatImage::fastDownSample and the separable polyphase resize.
 * Compile with -ffp-contract=off and without -ffast-math.
 */
#ifndef NTX_ORACLE_H
#define NTX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* nvtt::Format values (nvtt.h:77-97). */
enum { NTXO_FMT_BC1 = 1, NTXO_FMT_BC1A = 2, NTXO_FMT_BC3 = 4, NTXO_FMT_BC3N = 5, NTXO_FMT_BC4 = 6, NTXO_FMT_BC5 = 7 };
/* nvtt::MipmapFilter / nvtt::WrapMode / nvtt::AlphaMode values (nvtt.h). */
enum { NTXO_FILTER_BOX = 0, NTXO_FILTER_TRIANGLE = 1, NTXO_FILTER_KAISER = 2 };
enum { NTXO_WRAP_CLAMP = 0, NTXO_WRAP_REPEAT = 1, NTXO_WRAP_MIRROR = 2 };
enum { NTXO_ALPHA_NONE = 0, NTXO_ALPHA_TRANSPARENCY = 1 };
/* nvtt::Quality values (nvtt.h:148-154). */
enum { NTXO_QUALITY_FASTEST = 0, NTXO_QUALITY_NORMAL = 1, NTXO_QUALITY_PRODUCTION = 2, NTXO_QUALITY_HIGHEST = 3 };

/* Builds every lookup table (icbc init_dxt1 + initSingleColorLookup). Idempotent. */
void ntxo_init(void);

/* ---- tables, exposed so tests can pin them ---- */
int  ntxo_cluster_table4(uint8_t *out_c012 /* 976*3 */, int *totals /* 16 */);
int  ntxo_cluster_table3(uint8_t *out_c01 /* 160*2 */, int *totals /* 16 */);
void ntxo_icbc_match_tables(uint8_t *m5 /* 256*2 */, uint8_t *m6 /* 256*2 */);
void ntxo_omatch_tables(uint8_t *m5 /* 256*2 */, uint8_t *m6 /* 256*2 */);
void ntxo_midpoints(float *mid5 /* 32 */, float *mid6 /* 64 */);

/* ---- BC1 (icbc Level8) ----
 * colors: 16 texels x (r,g,b,a) floats; weights: 16 floats; cw: colour weights.
 * Writes 8 bytes.  Returns icbc's weighted MSE.  *n_out = reduce_colors count,
 * *m_out = count seen by the three-colour fit (or -1 if it did not run). */
float ntxo_bc1_block(const float *colors, const float *weights, const float *cw,
                     uint8_t *out, int *n_out, int *m_out);
/* The same at a quality: Production is icbc Quality_Level9 (Level8 plus
 * refine_endpoints, icbc.h:3329-3406), Fastest Quality_Level1 (box fit +
 * least squares), Normal and Highest Level8. */
float ntxo_bc1_block_q(const float *colors, const float *weights, const float *cw, int quality, uint8_t *out);
/* D3D10 decode of one BC1 block into 16 RGBA8 texels. */
void ntxo_bc1_decode(const uint8_t *blk, uint8_t *rgba);

/* ---- BC3 colour (squish) / BC3 / BC4 / BC5 on 8-bit blocks ----
 * bgra: 16 texels in nv::Color32 byte order (b,g,r,a). */
void ntxo_bc3_color_block(const uint8_t *bgra, int weight_by_alpha, const float *cw, uint8_t *out8);
void ntxo_bc4_channel(const uint8_t *v16, uint8_t *out8);
/* n channel blocks of 16 values: ntxo_bc4_channel on each. */
void ntxo_bc4_channels(const uint8_t *v, size_t n, uint8_t *out);
void ntxo_bc4_decode(const uint8_t *blk, uint8_t *v16);
void ntxo_bc1a_block(const uint8_t *bgra, int weight_by_alpha, const float *cw, uint8_t *out);
void ntxo_omatch_alpha_tables(uint8_t *m5, uint8_t *m6);
void ntxo_bc1a_squish(const uint8_t *bgra, int weight_by_alpha, const float *cw, uint8_t *out);
void ntxo_bc3n_block(const uint8_t *bgra, int weight_by_alpha, uint8_t *out);
void ntxo_bc3_block(const uint8_t *bgra, int weight_by_alpha, const float *cw, uint8_t *out16);
/* OptimalCompress::compressDXT5A (OptimalCompressDXT.cpp:512-607) on one
 * channel block, first error seed = the error of (max, min). */
void ntxo_bc4_channel_optimal(const uint8_t *v16, uint8_t *out8);
void ntxo_bc4_channels_optimal(const uint8_t *v, size_t n, uint8_t *out);
/* OptimalCompress::compressDXT1G(ColorBlock) (OptimalCompressDXT.cpp:315-390). */
void ntxo_bc1_green_optimal(const uint8_t *bgra, uint8_t *out8);
/* CompressorDXT5(n) / FastCompressorDXT5(n) at any quality (CompressorDX9.cpp:67-80, 145-222). */
void ntxo_bc3_block_q(const uint8_t *bgra, int quality, int weight_by_alpha, const float *cw, uint8_t *out16);
void ntxo_bc3n_block_q(const uint8_t *bgra, int quality, int weight_by_alpha, uint8_t *out16);

/* ---- whole-image compressors (planar RGBA float input, block-raster output) ---- */
size_t ntxo_output_size(int format, int w, int h);
void ntxo_gather_u8(int w, int h, const float *planar, int bx, int by, uint8_t *bgra);
void ntxo_compress_image_q(int format, int quality, int alpha_mode, const float *cw4, int w, int h,
                           const float *planar, uint8_t *out, int threads);
void ntxo_compress_image(int format, int alpha_mode, const float *cw4, int w, int h,
                         const float *planar, uint8_t *out, int threads);
/* BC1 image compression that also reports n,m per block (flop accounting). */
void ntxo_compress_bc1_counts(int alpha_mode, const float *cw4, int w, int h,
                              const float *planar, uint8_t *out, int16_t *n_m, int threads);

/* ---- mipmaps ---- */
void ntxo_fast_downsample(int W, int H, const float *src, float *dst);
int  ntxo_kernel_window(int filter, float width, int src_len, int dst_len, float *kwidth);
void ntxo_polyphase_kernel(int filter, float width, const float *params, int src_len, int dst_len,
                           float *weights /* dst_len * window */);
void ntxo_downsample(int filter, float width, const float *params, int wrap,
                     int W, int H, const float *src, float *dst);
/* Surface::buildNextMipmap for AlphaMode_None, depth 1. */
void ntxo_build_next_mipmap(int filter, float width, const float *params, int wrap,
                            int W, int H, const float *src, float *dst);
void ntxo_srgb_to_linear(size_t n, float *data);

/* ---- gamma (FloatImage.cpp:258-297, Gamma.cpp:30-441) ----
 * The exponent tables of the gamma-2.2 forms (which: 0 = 5/11, 1 = 11/5),
 * and FloatImage::toLinear / toGamma on n floats (the R, G, B planes). */
void ntxo_gamma_table(int which, float *out512);
float ntxo_powf_5_11(float x);
float ntxo_powf_11_5(float x);
void ntxo_to_linear(size_t n, float *data, float gamma);
void ntxo_to_gamma(size_t n, float *data, float gamma);
/* Context.cpp:329-336 on a generated normal-map level: expandNormals()
 * (Surface.cpp:2959-2963), normalizeNormalMap() (FloatImage.cpp:201-219),
 * packNormals() (Surface.cpp:2952-2956); planar RGBA of n texels, in place. */
void ntxo_normalize_normals(size_t n, float *planar);
/* FloatImage.cpp:201-219 normalize(0) on the planes as they are (KTX branch). */
void ntxo_normalize_vectors(size_t n, float *planar);

/* ---- decode and error metrics (Surface.cpp:1022-1107, ErrorMetric.cpp:13-69) ---- */
void ntxo_decode_image(int format, int d3d9, const uint8_t *blocks, int w, int h, float *planar);
float ntxo_rms_error(int metric, const float *ref, const float *img, size_t n);

#ifdef __cplusplus
}
#endif
#endif
