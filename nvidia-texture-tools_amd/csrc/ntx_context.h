// ntx_context.h -- internal: the context behind the C ABI handle and the
// helpers shared by the translation units of libntx_hip (ntx_api.cpp,
// batch_api.cpp).  Not installed; the public interface is include/ntx_hip.h.
#pragma once

#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "ntx_common.h"

namespace ntx {
hipError_t launch_bc1(const float *d_src, int w, int h, int alpha_mode, const float *cw, const Tables *d_tab, void *d_out,
                      int16_t *d_nm, hipStream_t stream, bool urgent = false, int level = 8);
// icbc Level8 in two passes, one block per lane (large images); d_counts:
// bucket scratch of bc1_scratch_ints(blocks) ints with zero tallies.
hipError_t launch_bc1_classify(const float *d_src, int w, int h, int alpha_mode, const Tables *d_tab, void *d_out,
                               int16_t *d_nm, const int *d_sel, const int *d_sel_count, int *d_counts, int *d_lists,
                               int cap, int max_groups, hipStream_t stream, bool urgent);
hipError_t launch_bc1_list(const float *d_src, int w, int h, int alpha_mode, const float *cw, const Tables *d_tab,
                           void *d_out, const int *d_list, const int *d_count, int cap, int max_groups,
                           hipStream_t stream);
hipError_t launch_bc1_search(const float *d_src, int w, int h, int alpha_mode, const float *cw, const Tables *d_tab,
                             void *d_out, int16_t *d_nm, int *d_counts, const int *d_lists, int cap, int max_groups,
                             hipStream_t stream);
// The multi-image two-pass BC1 encoder (Level8) over a table of images:
// at most bc1_multi_max_images() images of at most bc1_multi_max_blocks()
// blocks; d_map: (image, first block) per 256-block chunk.
int bc1_multi_max_images();
int bc1_multi_max_blocks();
hipError_t launch_bc1_classify_multi(const Bc1Image *d_imgs, const void *d_map, int nchunks, const Tables *d_tab,
                                     int *d_counts, int *d_lists, int cap, hipStream_t stream);
hipError_t launch_bc1_search_multi(const Bc1Image *d_imgs, int nimgs, const Tables *d_tab, int *d_counts,
                                   const int *d_lists, int cap, int max_groups, hipStream_t stream);
// Workgroups of bc1_search_kernel resident at once per compute unit.
constexpr int kBc1SearchGroupsPerCU = 8;
// Workgroups (four wavefronts each) per compute unit for the wavefront-per-
// block OptimalCompress searches: a grid-stride loop over the blocks.
constexpr int kOptimalGroupsPerCU = 8;
int bc1_scratch_ints(int nblocks);
hipError_t launch_channels(const float *d_src, int w, int h, int format, void *d_out, hipStream_t stream,
                           bool urgent = false);
hipError_t launch_bc3_front(const float *d_src, int w, int h, const Tables *d_tab, void *d_out, int *d_counts,
                            int *d_lists, int cap, int weight_by_alpha, uint32_t *d_packed, bool urgent,
                            hipStream_t stream, bool normal = false);
hipError_t launch_bc3_clear_counts(int *d_counts, hipStream_t stream);
hipError_t launch_bc1a(const float *d_src, int w, int h, int weight_by_alpha, const float *cw, const Tables *d_tab,
                       void *d_out, int *d_counts, int *d_lists, int cap, uint32_t *d_packed, hipStream_t stream);
hipError_t launch_optimal_channels(const float *d_src, int w, int h, int format, void *d_out, int groups,
                                   hipStream_t stream);
hipError_t launch_optimal_green(const float *d_src, int w, int h, const Tables *d_tab, void *d_out, int groups,
                                hipStream_t stream);
hipError_t launch_bc3_fast(const float *d_src, int w, int h, bool normal, const Tables *d_tab, void *d_out,
                           hipStream_t stream);
hipError_t launch_bc3_search(const uint32_t *d_packed, int w, int h, int weight_by_alpha, const float *cw, void *d_out,
                             int *d_counts, const int *d_lists, int cap, const Tables *d_tab, int lpb_shift,
                             bool urgent, hipStream_t stream);
int bc3_scratch_ints(int nblocks);
int bc3_lists_offset();
int bc3_packed_offset(int nblocks);
hipError_t launch_box(const float *d_src, int W, int H, float *d_dst, hipStream_t stream);
hipError_t launch_poly_x(const float *d_src, int W, int H, float *d_tmp, const float *d_kx, const int *d_lx, int wx, int wrap,
                         hipStream_t stream);
hipError_t launch_poly_y(const float *d_tmp, int W, int H, float *d_dst, const float *d_ky, const int *d_ly, int wy, int wrap,
                         hipStream_t stream);
hipError_t launch_mip_tail(const float *d_src, const MipTail &tail, int wrap, hipStream_t stream);
// Exact halving with the same taps for every output (PolyDev::uniform) in
// both directions, an odd window of 3 to 15 taps: host weights kx, ky,
// first taps lx0, ly0 of output 0, strips of `rows` output rows
// (poly_slide_rows: 0 when the tiled kernel suits the size better).
int poly_slide_rows(int W, int H, int cus);
hipError_t launch_poly_slide(const float *d_src, int W, int H, float *d_dst, const float *kx, int lx0, const float *ky,
                             int ly0, int win, int wrap, int rows, hipStream_t stream);
hipError_t launch_poly_fused(const float *d_src, int W, int H, float *d_dst, const float *d_kx, const int *d_lx, int wx,
                             bool pair_x, const float *d_ky, const int *d_ly, int wy, int wrap, hipStream_t stream);
bool poly_fused_fits(int max_sw, int max_sh, int wx, int wy);
// The fused kernel without tables, for exact halving with uniform taps
// (PolyDev::uniform) of a window poly_fused_uniform_window accepts.
bool poly_fused_uniform_window(int win);
hipError_t launch_poly_fused_uniform(const float *d_src, int W, int H, float *d_dst, const float *kx, int lx0,
                                     const float *ky, int ly0, int win, int wrap, hipStream_t stream);

int poly_tile_x();
int poly_tile_y();
hipError_t launch_srgb(float *d, size_t n, const float *d_lut, hipStream_t stream);
// d_srgb_lut: NULL, or the 256-entry sRGB->linear table to convert R, G, B as well.
hipError_t launch_unpack(const uint8_t *d_bgra, size_t n, float *d_dst, const float *d_srgb_lut, hipStream_t stream);
// Levels for one gamma launch: texels[l] = w * h of level l; start[] is
// filled by launch_gamma.
constexpr int kMaxGammaLevels = 16;
struct GammaJob {
    const float *src[kMaxGammaLevels];
    float *dst[kMaxGammaLevels];
    unsigned long long texels[kMaxGammaLevels];
    unsigned long long start[kMaxGammaLevels + 1];
    int levels;
};
// mode 0: gamma 2.2 -> linear (table form), 1: linear -> gamma 2.2 (table
// form), 2: powf(max(0, x), power).
hipError_t launch_gamma(const GammaJob &job, int mode, float power, const Tables *d_tab, hipStream_t stream);
// [expandNormals +] normalizeNormalMap [+ packNormals] on the R, G, B planes of n texels.
hipError_t launch_normals(float *d_rgba, size_t n, bool expand_pack, hipStream_t stream);
hipError_t launch_bc1_sse(const float *d_src, int w, int h, const void *d_blocks, double *d_partial, double *d_out,
                          hipStream_t stream);
int bc1_sse_partials(int w, int h);
hipError_t launch_decode(int format, int d3d9, const void *d_blocks, int w, int h, float *d_dst, const float *d_div255,
                         hipStream_t stream);
int rms_partials();
hipError_t launch_rms(const float *d_ref, const float *d_img, size_t n, int mode, double *d_partial, double *d_out,
                      hipStream_t stream);

struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
};

struct PolyDev {
    float *weights = nullptr;
    int *left = nullptr;
    int window = 0;
    int span_x = 0;  // largest source span of a run of poly_tile_x() outputs
    int span_y = 0;  // same for poly_tile_y() outputs
    bool stride2 = false;  // left[i] == left[0] + 2 i for every output i
    // stride2 and every output's weights equal output 0's (bit for bit):
    // exact halving, whose taps left0, w0[0 .. window) the sliding-window
    // kernel takes as arguments.
    bool uniform = false;
    int left0 = 0;
    float w0[16] = {};
    // Host copies of the weights and first taps (the tail's packed taps).
    std::vector<float> host_weights;
    std::vector<int> host_left;
};

using PolyKey = std::tuple<int, uint32_t, uint32_t, uint32_t, int, int>;
// Tail taps: filter, width, alpha, stretch bits, source W, H, levels, wrap.
using TailKey = std::tuple<int, uint32_t, uint32_t, uint32_t, int, int, int, int>;
struct TailTaps {
    float *packed = nullptr;  // device, 2 floats per tap
    int count = 0;
};

}  // namespace ntx

struct ntx_hip_context {
    int device = 0;
    int cus = 256;  // compute units of the device
    hipStream_t stream = nullptr;
    ntx::Tables *d_tab = nullptr;
    float *d_srgb = nullptr;
    // 8-bit -> linear table of ntx_hip_unpack_bgra8_gamma_device for the
    // gamma whose float bits are gamma_lut_key (valid when d_gamma_lut set).
    float *d_gamma_lut = nullptr;
    uint32_t gamma_lut_key = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    ntx::DevBuf in, out, tmp, dst, bucket, metric;
    // Bytes of the DevBuf scratch above (and `chain`) now allocated, and the
    // cap set by ntx_hip_set_scratch_limit (0: none).
    size_t scratch_bytes = 0, scratch_limit = 0;
    // BC1 images of at least this many blocks take the two-pass encoder.
    static constexpr int kBc1LanesMinDefault = 1 << 17;
    int bc1_lanes_min = kBc1LanesMinDefault;
    // The BC3 tallies at the start of `bucket` are zero: set after a complete
    // front + search sequence (the search clears them), cleared before one.
    bool bc3_counts_zero = false;
    // Concurrent compression of several images (mip chains, image sets):
    // the side streams, fork/join events, and per-image BC3 scratch (every
    // region's tallies zero when chain_counts_zero).
    static constexpr int kSideStreams = 2;
    hipStream_t side[kSideStreams] = {};
    std::vector<hipEvent_t> sync_events;
    ntx::DevBuf chain;
    bool chain_counts_zero = false;
    std::vector<size_t> chain_layout;  // offsets (ints) of the regions with tallies in the last complete run
    std::map<ntx::PolyKey, ntx::PolyDev> poly;
    std::map<ntx::TailKey, ntx::TailTaps> tail_taps;
    std::map<std::string, ntx_batch_tuning> batch_tuning;  // autotuned, keyed by batch shape list
    std::string err;
    // per-kernel event timing (ntx_hip_profile_*)
    bool profiling = false;
    struct Mark { int kernel; hipEvent_t start, stop; };
    std::vector<Mark> marks;
    std::vector<hipEvent_t> spare;
};

namespace ntx {

ntx_status fail(ntx_hip_context *ctx, hipError_t e, const char *what);
ntx_status bind(ntx_hip_context *ctx);
ntx_status ensure(ntx_hip_context *ctx, DevBuf &b, size_t bytes);
// Frees a buffer ensure() allocated (and takes it off the scratch count).
void release(ntx_hip_context *ctx, DevBuf &b);
hipEvent_t take_event(ntx_hip_context *ctx);
// Quality levels handled exactly by the kernels.
bool supported(int format, int quality);
// icbc quality level of a BC1 nvtt quality (1, 8 or 9).
int bc1_level(int quality);
// BC1 at this quality and size takes the two-pass (lane per block) encoder.
bool bc1_two_pass(const ntx_hip_context *ctx, int quality, int w, int h);
// Bucket scratch (ints) one image's compression needs: BC3, two-pass BC1.
size_t scratch_ints(const ntx_hip_context *ctx, int format, int quality, int w, int h);
// One image's compression on the context's stream with the given bucket
// scratch (null: BC1 takes its one-pass encoder); the tallies are cleared by
// a fill kernel first.
ntx_status side_stream_after(ntx_hip_context *ctx, int i, hipStream_t *out);
ntx_status side_stream_done(ntx_hip_context *ctx, int i, hipEvent_t *out);
ntx_status side_plain_bc1(ntx_hip_context *ctx, int n, const float *const *d_in, const int *w, const int *h,
                          void *const *d_out, hipEvent_t *join);
hipError_t compress_scratch(ntx_hip_context *ctx, int format, int quality, int alpha_mode, const float *cw,
                            const float *d_rgba, int w, int h, void *d_out, int *bucket);
// Looks up the device address of the no-allocation entry point's scratch
// array on the context's (current) device; once per device.
ntx_status noalloc_init(ntx_hip_context *ctx);
// Polyphase weights and first-tap indices for src -> dst samples, cached per context.
ntx_status poly_tables(ntx_hip_context *ctx, int filter, float width, const float *params, int src, int dst, PolyDev **out);

// Runs one launch; brackets it with events when the context is profiling.
// Runs `launch` (which enqueues on `stream`), bracketed by profiling events
// on that stream when profiling is on.
template <typename F>
hipError_t timed_on(ntx_hip_context *ctx, hipStream_t stream, int kernel, F &&launch)
{
    if (!ctx->profiling) return launch();
    hipEvent_t a = take_event(ctx), b = take_event(ctx);
    if (!a || !b) return hipErrorOutOfMemory;
    hipError_t e = hipEventRecord(a, stream);
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(b, stream);
    ctx->marks.push_back({ kernel, a, b });
    return e;
}

template <typename F>
hipError_t timed(ntx_hip_context *ctx, int kernel, F &&launch)
{
    return timed_on(ctx, ctx->stream, kernel, static_cast<F &&>(launch));
}

}  // namespace ntx

#define NTX_HIP(ctx, call)                                      \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return ntx::fail(ctx, e_, #call); \
    } while (0)
