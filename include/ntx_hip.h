/*
 * ntx_hip.h -- C ABI of the MI355X (gfx950) backend for NVTT's BC1/BC3/BC4/BC5
 * block compressors and mipmap downsampler.
 *
 * This is the seam that replaces NVTT 2.1.2's CUDA backend
 * (src/nvtt/cuda/CudaCompressorDXT.{h,cpp}, src/nvtt/cuda/CudaUtils.cpp) and
 * the single-threaded FloatImage mip filters behind Surface::buildNextMipmap.
 * Plain C types only: pointers, sizes and ints.  Every call returns an
 * ntx_status (0 = success).
 *
 * Threading: one context owns one HIP stream.  Calls on one context are
 * serialised by the caller (as Compressor::compress is in the reference,
 * whose CUDA compressor shares one result buffer: CudaCompressorDXT.cpp:67-118).
 * Entry points suffixed _device take device pointers and only enqueue work on
 * the context's stream; the others take host pointers and return when the
 * result is in host memory.
 */
#ifndef NTX_HIP_H
#define NTX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define NTX_API __attribute__((visibility("default")))
#else
#define NTX_API
#endif

typedef enum {
    NTX_OK = 0,
    NTX_ERR_INVALID_ARGUMENT = 1, /* maps to nvtt::Error_InvalidInput */
    NTX_ERR_UNSUPPORTED = 2,      /* maps to nvtt::Error_UnsupportedFeature */
    NTX_ERR_DEVICE = 3,           /* a HIP runtime error; maps to nvtt::Error_CudaError (value 3, kept for ABI) */
    NTX_ERR_NO_DEVICE = 4,
    NTX_ERR_OUT_OF_MEMORY = 5
} ntx_status;

/* nvtt::Format values (nvtt.h:77-97) accepted by the compressor.  BC1A is
 * Format_BC1a (DXT1a, CompressorDX9.cpp:83-111): BC1 whose texels of alpha 0
 * take the transparent index.  BC3N is Format_BC3n (DXT5n,
 * CompressorDX9.cpp:179-222): a normal map's X (red) in the alpha block and
 * Y (green) in the colour block. */
enum { NTX_FORMAT_BC1 = 1, NTX_FORMAT_BC1A = 2, NTX_FORMAT_BC3 = 4, NTX_FORMAT_BC3N = 5, NTX_FORMAT_BC4 = 6,
       NTX_FORMAT_BC5 = 7 };
/* nvtt::Quality (nvtt.h:148-154). */
enum { NTX_QUALITY_FASTEST = 0, NTX_QUALITY_NORMAL = 1, NTX_QUALITY_PRODUCTION = 2, NTX_QUALITY_HIGHEST = 3 };
/* nvtt::AlphaMode (nvtt.h:274-279). */
enum { NTX_ALPHA_NONE = 0, NTX_ALPHA_TRANSPARENCY = 1, NTX_ALPHA_PREMULTIPLIED = 2 };
/* nvtt::MipmapFilter (nvtt.h:236-241) and nvtt::WrapMode (nvtt.h:209-214). */
enum { NTX_FILTER_BOX = 0, NTX_FILTER_TRIANGLE = 1, NTX_FILTER_KAISER = 2 };
enum { NTX_WRAP_CLAMP = 0, NTX_WRAP_REPEAT = 1, NTX_WRAP_MIRROR = 2 };

typedef struct ntx_hip_context ntx_hip_context;

/* ---------------------------------------------------------------- devices --
 * Replace cuda::isHardwarePresent / cuda::getFastestDevice
 * (src/nvtt/cuda/CudaUtils.cpp:136-224, called from Context.cpp:61-73). */
NTX_API int ntx_hip_is_hardware_present(void);
NTX_API ntx_status ntx_hip_device_count(int *count);
NTX_API const char *ntx_hip_status_string(ntx_status status);

/* --------------------------------------------------------------- contexts --
 * Replace CudaContext (CudaCompressorDXT.cpp:67-118), created by
 * Compressor::enableCudaAcceleration (Context.cpp:81-98).  Uploads the cluster
 * and single-colour tables and creates the stream. */
NTX_API ntx_status ntx_hip_create(int device, ntx_hip_context **ctx);
NTX_API ntx_status ntx_hip_destroy(ntx_hip_context *ctx);
/* Text of the last HIP error seen by this context ("" if none). */
NTX_API const char *ntx_hip_last_error(const ntx_hip_context *ctx);
/* The hipStream_t the context launches on (as a void*). */
NTX_API void *ntx_hip_stream(ntx_hip_context *ctx);

/* Device memory staging for callers that do not own a HIP allocator. */
NTX_API ntx_status ntx_hip_malloc(ntx_hip_context *ctx, size_t bytes, void **dptr);
NTX_API ntx_status ntx_hip_free(ntx_hip_context *ctx, void *dptr);
NTX_API ntx_status ntx_hip_upload(ntx_hip_context *ctx, void *dst, const void *src, size_t bytes);
NTX_API ntx_status ntx_hip_download(ntx_hip_context *ctx, void *dst, const void *src, size_t bytes);
NTX_API ntx_status ntx_hip_synchronize(ntx_hip_context *ctx);
/* Device-to-device copy and byte fill on the context's stream; both wait for
 * completion.  Used by the Surface front end for copy-on-write
 * (Surface.cpp:354-363 detach) and for Surface::setImage(w, h, d). */
NTX_API ntx_status ntx_hip_copy_device(ntx_hip_context *ctx, void *dst, const void *src, size_t bytes);
NTX_API ntx_status ntx_hip_memset_device(ntx_hip_context *ctx, void *dst, int value, size_t bytes);

/* Stream-ordered timer: records an event pair around the work enqueued
 * between begin and end; end waits and returns the elapsed milliseconds. */
NTX_API ntx_status ntx_hip_timer_begin(ntx_hip_context *ctx);
NTX_API ntx_status ntx_hip_timer_end(ntx_hip_context *ctx, float *ms);

/* Per-kernel timing: while enabled, every kernel the context launches is
 * bracketed by a pair of HIP events on the context's stream.  read()
 * waits for the stream and returns, per kernel, the summed milliseconds and
 * launch count since the last reset.  No counterpart in the reference (its
 * CUDA timing is commented out, CudaCompressorDXT.cpp:173-174). */
enum {
    NTX_KERNEL_BC1 = 0,
    NTX_KERNEL_BC3_SEARCH = 1, /* BC3 colour: weighted cluster fit over bucketed blocks, one per lane */
    NTX_KERNEL_CHANNEL = 2, /* BC4 / BC5 blocks */
    NTX_KERNEL_BOX = 3,
    NTX_KERNEL_POLY_X = 4,
    NTX_KERNEL_POLY_Y = 5,
    NTX_KERNEL_SRGB = 6,
    NTX_KERNEL_UNPACK = 7,
    NTX_KERNEL_POLY_FUSED = 8, /* X and Y polyphase taps in one pass through LDS */
    NTX_KERNEL_BC3_CLASSIFY = 9, /* BC3 pass 1: alpha halves, single-colour blocks, bucketing by colour count */
    NTX_KERNEL_BATCH = 10,       /* one batch run: mip stages + persistent compression kernel + two-pass BC1 */
    NTX_KERNEL_BC1_MLP = 11,     /* BC1 from network proposals + least squares */
    NTX_KERNEL_BC1_LIST = 12,    /* BC1 Level8 over a block list (the learned form's fallback), one block per 16-lane group */
    NTX_KERNEL_MLP_TRAIN = 13,   /* network training step (forward, backward, Adam) */
    NTX_KERNEL_BC3_SEARCH_SPLIT = 14, /* BC3 colour search of a small image: 2..64 lanes per block */
    NTX_KERNEL_GAMMA = 15,       /* gamma transforms of R, G, B (one launch per up to 16 levels) */
    NTX_KERNEL_NORMALS = 16,     /* normal-map renormalisation of a generated level */
    NTX_KERNEL_BC1_CLASSIFY = 17, /* BC1 pass 1 (large images): colour counts, 0/1-colour blocks, bucketing */
    NTX_KERNEL_BC1_SEARCH = 18,  /* BC1 pass 2: cluster fit, one block per lane, over the buckets */
    NTX_KERNEL_OPTIMAL_CHANNEL = 19, /* OptimalCompress::compressDXT5A, one wavefront per channel block */
    NTX_KERNEL_OPTIMAL_GREEN = 20,   /* OptimalCompress::compressDXT1G (BC3n Highest colour), one wavefront per block */
    NTX_KERNEL_BC3_FAST = 21,        /* QuickCompress::compressDXT5 (BC3 / BC3n Fastest), one thread per block */
    NTX_KERNEL_BC1A = 22,            /* BC1a: bucketing front pass + kDxt1 cluster fit, one block per lane */
    NTX_KERNEL_MIP_TAIL = 23,        /* the small last levels of a mip chain, one launch, one workgroup per plane */
    NTX_KERNEL_POLY_SLIDE = 24,      /* exact halving, same taps for every output: one wavefront per column strip */
    NTX_KERNEL_COUNT = 25
};
NTX_API const char *ntx_hip_kernel_name(int kernel);
NTX_API ntx_status ntx_hip_profile_enable(ntx_hip_context *ctx, int enable);
NTX_API ntx_status ntx_hip_profile_read(ntx_hip_context *ctx, float *total_ms /* NTX_KERNEL_COUNT */,
                                        int *launches /* NTX_KERNEL_COUNT */, int reset);

/* ------------------------------------------------------------- compression --
 * Size in bytes of the block-compressed image (Surface.cpp:210-218
 * computeImageSize for block formats): ceil(w/4)*ceil(h/4)*{8,16}. */
NTX_API size_t ntx_hip_compressed_size(int format, int w, int h);

/* Replaces CudaCompressor::compress (CudaCompressorDXT.cpp:128-206) behind
 * nv::CompressorInterface::compress (src/nvtt/Compressor.h:34-38), extended
 * from BC1 to BC1/BC3/BC3n/BC4/BC5, each at the qualities whose CPU
 * compressor it reproduces byte for byte:
 *   BC1 (every quality; BlockCompressor.cpp:116-222): icbc compress_dxt1 per
 *     block at Level1 (Fastest), Level8 (Normal, Highest) or Level9
 *     (Production);
 *   BC1a (Normal, Production, Highest; CompressorDX9.cpp:83-111):
 *     OptimalCompress::compressDXT1a for blocks of one RGB, otherwise the
 *     squish weighted cluster fit with kDxt1 (texels of alpha 0 take index
 *     3; Compress3, then Compress4 unless a texel is transparent).  Fastest
 *     (QuickCompress::compressDXT1a) returns NTX_ERR_UNSUPPORTED: its
 *     three-colour index loop reads 16 entries of an array that holds only
 *     the opaque texels (QuickCompressDXT.cpp:378-403, 740-760), so its
 *     output depends on uninitialised memory;
 *   BC3 / BC3n (every quality; Context.cpp:1080-1097 picks the compressor):
 *     Fastest: QuickCompress::compressDXT5 (FastCompressorDXT5 / DXT5n,
 *     CompressorDX9.cpp:67-80); Normal, Production: QuickCompress alpha +
 *     squish weighted cluster fit colour (CompressorDX9.cpp:145-222);
 *     Highest: OptimalCompress::compressDXT5A alpha (BC3n: red) and, for
 *     BC3n, OptimalCompress::compressDXT1G colour;
 *   BC4 / BC5 (every quality; Context.cpp:1098-1116): Fastest, Normal:
 *     QuickCompress::compressDXT5A on R / R,G (CompressorDX10.cpp:42-60);
 *     Production, Highest: OptimalCompress::compressDXT5A
 *     (CompressorDX10.cpp:65-87).
 * OptimalCompress::compressDXT5A reads its first error seed from the output
 * block before writing it (OptimalCompressDXT.cpp:546), so the reference's
 * result depends on what that memory held; here the seed is the error of
 * the (max, min) endpoints it falls back to, a fixed choice.
 * rgba: planar FP32, four planes R,G,B,A of w*h floats (Surface::data(),
 * Surface.cpp:492-495).  out: ntx_hip_compressed_size() bytes in block-raster
 * order.  color_weights: CompressionOptions::setColorWeights (r,g,b,a).
 * Other (format, quality) pairs return NTX_ERR_UNSUPPORTED. */
NTX_API ntx_status ntx_hip_compress(ntx_hip_context *ctx, int format, int quality, int alpha_mode,
                                    const float *color_weights, const float *rgba, int w, int h,
                                    void *out, size_t out_bytes);
NTX_API ntx_status ntx_hip_compress_device(ntx_hip_context *ctx, int format, int quality, int alpha_mode,
                                           const float *color_weights, const float *d_rgba, int w, int h,
                                           void *d_out);
/* ntx_hip_compress_device for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, hosts that forbid
 * hipMalloc); no counterpart in the reference.  Same arguments, same bytes;
 * calls no allocating routine.  The scratch of the encoders that need one
 * (BC3 / BC3n, and BC1 at Level8 when the two-pass encoder applies) is a
 * device array of the library, 8.7 MB per GPU, held until the process ends:
 * such images may have at most 65,536 4x4 blocks (1024 x 1024 texels), else
 * NTX_ERR_INVALID_ARGUMENT (BC1 above that uses its one-pass encoder, which
 * needs none).  Every call on a GPU shares its array: calls that overlap on
 * one GPU (several threads or contexts) race. */
NTX_API ntx_status ntx_hip_compress_device_noalloc(ntx_hip_context *ctx, int format, int quality, int alpha_mode,
                                                   const float *color_weights, const float *d_rgba, int w, int h,
                                                   void *d_out);

/* Caps the device memory the context allocates for its own scratch and
 * staging (host-pointer staging, bucket lists, chain regions, metric
 * partials); a call that would exceed it allocates nothing and returns
 * NTX_ERR_OUT_OF_MEMORY (Error_CudaError in the nvtt front end, which then
 * goes on with the next image, as CudaCompressorDXT.cpp:183-189 does).
 * 0 = no cap (the default).  No counterpart in the reference. */
NTX_API ntx_status ntx_hip_set_scratch_limit(ntx_hip_context *ctx, size_t bytes);

/* BC1 at icbc Level8 (Normal / Highest) has two encoders with identical
 * output: one block per 16-lane group in one pass, and, for images of at
 * least `min_blocks` 4x4 blocks, two passes with one block per lane (colour
 * counts, then the cluster fit over blocks bucketed by count).  Sets the
 * context's threshold; 0 restores the default (2^17 blocks).  No
 * counterpart in the reference. */
NTX_API ntx_status ntx_hip_set_bc1_lanes_min(ntx_hip_context *ctx, int min_blocks);

/* BC1 only: also writes, per block, the colour count after reduce_colors (n)
 * and the count seen by the three-colour fit (m, -1 if not run) as int16
 * pairs; these fix the algorithmic flop count of each block. */
NTX_API ntx_status ntx_hip_compress_bc1_stats_device(ntx_hip_context *ctx, int alpha_mode,
                                                     const float *color_weights, const float *d_rgba,
                                                     int w, int h, void *d_out, int16_t *d_nm);

/* ---------------------------------------------------------------- mipmaps --
 * Replaces Surface::buildNextMipmap's filtering (Surface.cpp:1344-1406):
 * AlphaMode_None with Box of width 0.5 -> FloatImage::fastDownSample
 * (FloatImage.cpp:559-737); every other case -> FloatImage::downSample ->
 * resize with PolyphaseKernel weights (FloatImage.cpp:740-808, 1115-1176;
 * Filter.cpp:563-608).  With AlphaMode_Transparency the reference filters the
 * alpha plane first (FloatImage.cpp:877-925), which changes no value.
 * params: Kaiser (alpha, stretch), may be NULL.  Output is
 * max(1,W/2) x max(1,H/2), four planes.  1x1 input is rejected (the reference
 * makes no further mipmap: canMakeNextMipmap, Surface.cpp:155-167). */
NTX_API ntx_status ntx_hip_downsample(ntx_hip_context *ctx, int filter, float width, const float *params,
                                      int wrap, int alpha_mode, const float *rgba, int W, int H, float *out);
NTX_API ntx_status ntx_hip_downsample_device(ntx_hip_context *ctx, int filter, float width, const float *params,
                                             int wrap, int alpha_mode, const float *d_rgba, int W, int H,
                                             float *d_out);

/* Surface::buildNextMipmap repeated (Surface.cpp:1344-1406): d_levels[0] is
 * level 0 (w x h, read only); d_levels[k], 1 <= k < levels, receives level k
 * (max(1, w >> k) x max(1, h >> k)), each exactly as ntx_hip_downsample_device
 * makes it from level k - 1. */
NTX_API ntx_status ntx_hip_downsample_chain_device(ntx_hip_context *ctx, int filter, float filter_width,
                                                   const float *filter_params, int wrap, int alpha_mode, int w, int h,
                                                   int levels, float *const *d_levels);

/* A whole mip chain as the per-mipmap loop of Compressor::Private::compress
 * runs it (Context.cpp:288-343: compress the level, then build the next one
 * with Surface::buildNextMipmap), with the gamma and quantisation steps left
 * to the caller.  d_levels[0] is level 0 (w x h, read only); for
 * 1 <= k < levels, d_levels[k] receives level k (max(1, w >> k) x
 * max(1, h >> k), as ntx_hip_downsample_device makes it from level k - 1);
 * d_out[k] receives level k compressed as ntx_hip_compress_device does it.
 * levels: 1 .. the full chain length (nv::countMipmaps).  The bytes are those
 * of the per-level calls; the work is scheduled differently: the levels'
 * compressions run concurrently on the context's stream and two internal
 * streams (the small levels' latency-bound launches overlap the large ones),
 * and all of it is ordered before later work on the context's stream. */
NTX_API ntx_status ntx_hip_compress_chain_device(ntx_hip_context *ctx, int format, int quality, int alpha_mode,
                                                 const float *color_weights, int filter, float filter_width,
                                                 const float *filter_params, int wrap, int w, int h, int levels,
                                                 float *const *d_levels, void *const *d_out);

/* n independent images (d_rgba[k]: w[k] x h[k] planar FP32), each compressed
 * into d_out[k] exactly as ntx_hip_compress_device does it, the way
 * Compressor::compress would be called once per image (Context.cpp:486-516)
 * but scheduled together: the largest on the context's stream, the others on
 * two internal streams beside it (the small images' latency-bound launches
 * overlap the large ones).  All of it is ordered before later work on the
 * context's stream. */
NTX_API ntx_status ntx_hip_compress_images_device(ntx_hip_context *ctx, int format, int quality, int alpha_mode,
                                                  const float *color_weights, int n, const float *const *d_rgba,
                                                  const int *w, const int *h, void *const *d_out);

/* Surface::toLinearFromSrgb (Surface.cpp:1562-1592) on the R,G,B planes, in place. */
NTX_API ntx_status ntx_hip_srgb_to_linear_device(ntx_hip_context *ctx, float *d_rgba, int w, int h);
/* Surface::setImage(InputFormat_BGRA_8UB, ...) (Surface.cpp:745-761):
 * 8-bit B,G,R,A texels to planar FP32 (value / 255). */
NTX_API ntx_status ntx_hip_unpack_bgra8_device(ntx_hip_context *ctx, const uint8_t *d_bgra, int w, int h,
                                               float *d_rgba);
/* Surface::setImage(InputFormat_BGRA_8UB, ...) followed by
 * Surface::toLinearFromSrgb() (Surface.cpp:745-761, 1562-1592) in one pass
 * over the texels; the same values as the two calls above. */
NTX_API ntx_status ntx_hip_unpack_bgra8_srgb_device(ntx_hip_context *ctx, const uint8_t *d_bgra, int w, int h,
                                                    float *d_rgba);

/* FloatImage::toLinear(0, 3, gamma) (FloatImage.cpp:258-270) when
 * to_gamma == 0, FloatImage::toGamma(0, 3, gamma) (FloatImage.cpp:272-284)
 * when to_gamma == 1, on the R, G, B planes of a w x h planar FP32 RGBA image;
 * the A plane is copied.  gamma == 2.2f takes the reference's table-and-
 * polynomial powf(x, 11/5) / powf(x, 5/11) (Gamma.cpp:311-441), bit-exact;
 * any other gamma takes powf(max(0, x), gamma) or powf(max(0, x), 1 / gamma)
 * (FloatImage.cpp:286-297) correctly rounded (exp2 and log2 in double), which
 * differs by one unit in the last place from the host C library's powf where
 * that is not correctly rounded (about 5 in 10^4 inputs).  d_dst may equal d_src.  Unlike
 * Surface::toLinear / toGamma (Surface.cpp:1470-1487) there is no shortcut for
 * a gamma of 1. */
NTX_API ntx_status ntx_hip_gamma_device(ntx_hip_context *ctx, int to_gamma, float gamma, const float *d_src, int w,
                                        int h, float *d_dst);
/* The same over n images (the levels of a chain) in one launch per 16. */
NTX_API ntx_status ntx_hip_gamma_levels_device(ntx_hip_context *ctx, int to_gamma, float gamma, int n,
                                               const float *const *d_src, const int *w, const int *h,
                                               float *const *d_dst);
/* Surface::setImage(InputFormat_BGRA_8UB, ...) followed by
 * Surface::toLinear(gamma) (Surface.cpp:745-761, 1470-1477) in one pass: each
 * 8-bit R, G, B value goes through a 256-entry table built on the host with
 * the reference's operations (the table form for 2.2f, the C library's powf
 * otherwise), so the result is bit-identical to the two calls; a gamma within
 * nv::equal of 1 leaves the values as setImage gives them. */
NTX_API ntx_status ntx_hip_unpack_bgra8_gamma_device(ntx_hip_context *ctx, const uint8_t *d_bgra, int w, int h,
                                                     float gamma, float *d_rgba);

/* What Compressor::process does to each generated level of a normal map
 * with normalizeMipmaps set (Context.cpp:329-336): Surface::expandNormals()
 * (scale 2, bias -2 * 127 / 255; Surface.cpp:2959-2963), normalizeNormalMap()
 * (FloatImage.cpp:201-219: v / |v|, zero for a zero vector) and packNormals()
 * (scale 0.5, bias 0.5; Surface.cpp:2952-2956), in place on R, G, B. */
NTX_API ntx_status ntx_hip_normalize_normals_device(ntx_hip_context *ctx, float *d_rgba, int w, int h);
/* Surface::normalizeNormalMap() alone (FloatImage::normalize(0),
 * FloatImage.cpp:201-219), in place on R, G, B: what the KTX branch of
 * Compressor::process applies to generated normal-map levels
 * (Context.cpp:446-449). */
NTX_API ntx_status ntx_hip_normalize_vectors_device(ntx_hip_context *ctx, float *d_rgba, int w, int h);

/* nvtt::Decoder values (nvtt.h:149-155). */
enum { NTX_DECODER_D3D10 = 0, NTX_DECODER_D3D9 = 1, NTX_DECODER_NV5X = 2 };
/* Surface::setImage2D(format, decoder, w, h, blocks) (Surface.cpp:908-1118)
 * for BC1, BC3, BC3n (decoded as BC3, unswizzled), BC4 and BC5 with the
 * D3D10 or D3D9 decoder: every block
 * decoded to 8-bit RGBA as nv::BlockDXT1 / BlockDXT5 / BlockATI1 / BlockATI2
 * do (BlockDXT.cpp:43-98, 200-217, 336-432, 475-566; BC4 gives R = G = B,
 * BC5 R, G and B = 0), then value / 255 into the four planes of d_rgba.
 * NV5x: NTX_ERR_UNSUPPORTED. */
NTX_API ntx_status ntx_hip_decode_device(ntx_hip_context *ctx, int format, int decoder, const void *d_blocks, int w, int h,
                                         float *d_rgba);
/* Error metrics of ErrorMetric.cpp, two planar RGBA images of w x h. */
enum {
    NTX_METRIC_RMS_COLOR = 0,          /* nvtt::rmsError, reference without Transparency (ErrorMetric.cpp:13-46) */
    NTX_METRIC_RMS_COLOR_ALPHA_WEIGHTED = 1, /* nvtt::rmsError, reference with AlphaMode_Transparency */
    NTX_METRIC_RMS_ALPHA = 2           /* nvtt::rmsAlphaError (ErrorMetric.cpp:50-69) */
};
/* The per-texel terms are the reference's float operations; they are summed
 * in double as the reference does, but not one after another, so the
 * double sum may differ in its last bits and the float result by one unit in
 * the last place at most in practice.  Waits for the device. */
NTX_API ntx_status ntx_hip_rms_error_device(ntx_hip_context *ctx, int metric, const float *d_reference,
                                            const float *d_image, int w, int h, float *result);

/* Error of a BC1 level against its FP32 source: D3D10 decode
 * (BlockDXT.cpp:43-98) and the sum over texels of the squared R, G, B
 * differences in 8-bit units (bc1enc.cpp:73-81); PSNR = 20 log10(255 /
 * sqrt(sse / (w h))).  Deterministic; waits for the device. */
NTX_API ntx_status ntx_hip_bc1_sse_device(ntx_hip_context *ctx, const float *d_rgba, int w, int h, const void *d_blocks,
                                          double *sse);

/* ---------------------------------------------------------------- batches --
 * No counterpart in the reference, which compresses one image per
 * Compressor::compress call (Context.cpp:146-190, 486-516).  A batch is many
 * textures of any sizes, each with a full or partial mip chain
 * (Surface::buildNextMipmap semantics, Surface.cpp:1335-1406) and every level
 * block-compressed (Compressor::compress semantics), run as one HIP graph in
 * one of two schedules (ntx_batch_tuning::schedule): a launch per mip stage
 * (level s + 1 of every texture) then one persistent launch over the
 * compression items, or one persistent launch taking mip and compression
 * items from one queue, ordered by per-level ready counters; after either,
 * the two-pass BC1 encoder over every BC1 level at icbc Level8.  Every output
 * byte equals what the per-image entry points above produce for the same
 * level.
 *
 * Formats: BC1 at every quality (icbc Level8 one block per lane through the
 * two-pass encoder, Level1 / Level9 one block per 16-lane group), BC4 and
 * BC5 (Fastest/Normal).  BC3 / BC3n, and
 * BC4 / BC5 at Production / Highest, return NTX_ERR_UNSUPPORTED here; use
 * ntx_hip_compress_chain_device or ntx_hip_compress_images_device for them. */
typedef struct {
    const float *d_rgba;   /* level 0: planar FP32 R,G,B,A planes of width*height, device memory */
    int width, height;
    int format;            /* NTX_FORMAT_BC1 / BC4 / BC5 */
    int quality;           /* NTX_QUALITY_* */
    int alpha_mode;        /* NTX_ALPHA_*: BC1 weights and the mip filter choice */
    float color_weights[4];/* CompressionOptions::setColorWeights; all zero = (1,1,1,1) */
    int mip_levels;        /* levels to produce and compress, 0 = full chain (nv::countMipmaps) */
    int filter;            /* NTX_FILTER_* */
    float filter_width;    /* <= 0: the filter's default (Surface.cpp:1121-1141) */
    float filter_params[2];/* Kaiser alpha, stretch (ignored unless filter_width > 0) */
    int wrap;              /* NTX_WRAP_* */
} ntx_texture_desc;

/* Work decomposition of the mip stages and the persistent compression
 * kernel; ntx_hip_batch_autotune chooses it by timing candidates on the
 * batch itself. */
typedef struct {
    int bc1_blocks_per_item;   /* multiple of 16 */
    int channel_jobs_per_item; /* multiple of 256 (a BC5 block is two jobs) */
    int mip_tiles_per_item;
    int mip_tile_width;        /* 16, 32, 64, 128 or 256 outputs; tile height = 256 / width */
    int kernel_variant;        /* 0: 4 waves/SIMD, 1: 5, 2: 6 (register budget of the kernel) */
    int workgroups;            /* persistent workgroups, 0 = as many as are resident */
    int schedule;              /* 0: a launch per mip stage, then the persistent compression launch;
                                  1: mips and compression in one persistent launch */
} ntx_batch_tuning;

typedef struct ntx_hip_batch ntx_hip_batch;

/* Plans a batch: level buffers for levels >= 1, output buffers, polyphase
 * tables, the work list and the graph.  descs[i].d_rgba must stay valid and
 * unchanged while the batch is used.  The tuning last chosen by
 * ntx_hip_batch_autotune for an identical list of shapes on this context is
 * reused, otherwise defaults apply. */
NTX_API ntx_status ntx_hip_batch_create(ntx_hip_context *ctx, const ntx_texture_desc *descs, int n,
                                        ntx_hip_batch **batch);
NTX_API ntx_status ntx_hip_batch_destroy(ntx_hip_batch *batch);
/* Enqueues one run of the whole batch on the context's stream. */
NTX_API ntx_status ntx_hip_batch_submit(ntx_hip_batch *batch);
/* Waits for the submitted runs; NTX_ERR_DEVICE if the one-launch schedule's
 * kernel reported a dependency wait that did not complete. */
NTX_API ntx_status ntx_hip_batch_wait(ntx_hip_batch *batch);
/* Number of levels texture `tex` has in this batch. */
NTX_API int ntx_hip_batch_levels(const ntx_hip_batch *batch, int tex);
/* Device pointers to one level: its planar FP32 image (w x h) and its
 * compressed blocks (bytes long), valid until the batch is destroyed. */
NTX_API ntx_status ntx_hip_batch_level(const ntx_hip_batch *batch, int tex, int level, const float **d_rgba, int *w,
                                       int *h, const void **d_blocks, size_t *bytes);
NTX_API ntx_status ntx_hip_batch_get_tuning(const ntx_hip_batch *batch, ntx_batch_tuning *tuning);
/* Replaces the work decomposition (rebuilds the work list and graph). */
NTX_API ntx_status ntx_hip_batch_set_tuning(ntx_hip_batch *batch, const ntx_batch_tuning *tuning);
/* Times candidate decompositions on this batch (each run `reps` times after
 * one warm-up; coordinate descent over the fields of ntx_batch_tuning),
 * keeps the fastest, records it on the context for identical batches, and
 * returns it with its time in milliseconds per run. */
NTX_API ntx_status ntx_hip_batch_autotune(ntx_hip_batch *batch, int reps, ntx_batch_tuning *best, float *best_ms);

/* ------------------------------------------------ learned BC1 endpoints --
 * No counterpart in the reference.  A 48-64-64-25 ReLU network (BF16
 * weights and activations, FP32 accumulation on the matrix cores) proposes
 * four candidate endpoint pairs per 4x4 block, as offsets from the four
 * diagonals of the block's colour bounding box, plus the RMSE the full
 * cluster fit would reach; each candidate is refined by least squares and
 * quantised as icbc does, the best is kept, and blocks the network is unsure
 * about are encoded by the full cluster fit (the ntx_hip_compress BC1 path),
 * which is therefore both the fallback and the measure the proposals are
 * trained on.  Parameters: NTX_MLP_PARAMS floats, W1[64][48] b1[64]
 * W2[64][64] b2[64] W3[25][64] b3[25].  Features: per block 48 BF16 values
 * (R, G, B planes of the 16 texels, centred on the block mean and divided by
 * the largest deviation or 1/32, whichever is larger) and 4 floats (mean R,
 * G, B, scale).  Labels: 8 floats per block (cluster-fit endpoints A, B and
 * RMSE in that frame). */
#define NTX_MLP_PARAMS 8921
typedef struct ntx_hip_mlp ntx_hip_mlp;

/* params NULL: He-uniform hidden layers from `seed`, zero biases and zero
 * output layer (the untrained network proposes the box diagonals). */
NTX_API ntx_status ntx_hip_mlp_create(ntx_hip_context *ctx, const float *params, uint32_t seed, ntx_hip_mlp **mlp);
NTX_API ntx_status ntx_hip_mlp_destroy(ntx_hip_mlp *mlp);
NTX_API ntx_status ntx_hip_mlp_get_params(ntx_hip_mlp *mlp, float *params);
/* Also resets the optimiser state. */
NTX_API ntx_status ntx_hip_mlp_set_params(ntx_hip_mlp *mlp, const float *params);
/* Features of every block of a w x h planar FP32 image (BC1 block order). */
NTX_API ntx_status ntx_hip_mlp_features_device(ntx_hip_context *ctx, const float *d_rgba, int w, int h,
                                               void *d_features, float *d_aux);
/* Training labels from the image's BC1 cluster-fit blocks. */
NTX_API ntx_status ntx_hip_mlp_labels_device(ntx_hip_context *ctx, const float *d_rgba, int w, int h, const void *d_bc1,
                                             const float *d_aux, float *d_labels);
/* Network outputs for n blocks: 32 floats per block, the first 25 used. */
NTX_API ntx_status ntx_hip_mlp_forward_device(ntx_hip_mlp *mlp, const void *d_features, int n, float *d_out);
/* One training step on n blocks (min-of-4 endpoint loss + RMSE head): the
 * mean loss per block and, when grad is not NULL, the mean gradient
 * (NTX_MLP_PARAMS floats, host memory).  lr > 0 applies an Adam step. */
NTX_API ntx_status ntx_hip_mlp_train_step(ntx_hip_mlp *mlp, const void *d_features, const float *d_labels, int n,
                                          float lr, float *loss, float *grad);
/* BC1 (AlphaMode_None, colour weights 1) of a w x h planar FP32 image from
 * the network's proposals.  A block falls back to the full cluster fit when
 * its predicted cluster-fit RMSE exceeds tau_pred, or when its best refined
 * candidate's RMSE exceeds that prediction by more than tau_gap (disabled
 * if negative); both in 8-bit units.  refine_iters: least-squares passes
 * per candidate; greedy_rounds: rounds of the search that moves one 565
 * channel of one endpoint by one step while the block error drops.  flags:
 * NTX_MLP_BOX_CANDIDATES starts from the four diagonals of each block's
 * colour bounding box instead of the network (the same refinement without
 * the learned part; only tau_pred < 0, which sends every block to the
 * fallback, applies).  fallback_blocks (may be NULL) receives the count,
 * which makes the call wait for the device. */
#define NTX_MLP_BOX_CANDIDATES 1
NTX_API ntx_status ntx_hip_compress_bc1_mlp_device(ntx_hip_mlp *mlp, const float *d_rgba, int w, int h, float tau_pred,
                                                   float tau_gap, int refine_iters, int greedy_rounds, int flags,
                                                   void *d_out, int *fallback_blocks);

/* The learned form over n images (the levels of mip chains, cube faces) in
 * one call, each image with its own settings: images whose settings set
 * `plain` take the plain cluster fit (the bytes of ntx_hip_compress_device
 * at BC1 Normal) concurrently on a side stream; the others go through
 * ntx_hip_compress_bc1_mlp_device with their settings one after another on
 * the context's stream.  No counterpart in the reference. */
typedef struct {
    float tau_pred, tau_gap;
    int refine_iters, greedy_rounds, flags;
    int plain;
} ntx_mlp_settings;
NTX_API ntx_status ntx_hip_compress_bc1_mlp_images_device(ntx_hip_mlp *mlp, int n, const float *const *d_rgba,
                                                          const int *w, const int *h,
                                                          const ntx_mlp_settings *settings, void *const *d_out);

#ifdef __cplusplus
}
#endif
#endif /* NTX_HIP_H */
