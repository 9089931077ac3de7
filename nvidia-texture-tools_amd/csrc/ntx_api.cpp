// ntx_api.cpp -- the C ABI declared in include/ntx_hip.h: contexts, staging,
// dispatch to the gfx950 kernels.  No CPU compression path exists here: when
// the device is missing or a HIP call fails the entry points return an error
// status, which the nvtt front end reports as Error_CudaError.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "gamma.hpp"
#include "ntx_context.h"

namespace {

int max_span(const std::vector<int> &left, int window, int tile)
{
    int best = 0;
    for (size_t i = 0; i < left.size(); i += tile) {
        const size_t last = std::min(left.size(), i + size_t(tile)) - 1;
        best = std::max(best, left[last] - left[i] + window);
    }
    return best;
}

uint32_t fbits(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// Surface.cpp:1562-1569 fromSrgb, evaluated on the host for the 256 values an
// 8-bit source can hold.
float from_srgb(float f)
{
    if (f < 0.0f) f = 0.0f;
    else if (f < 0.04045f) f = f / 12.92f;
    else if (f <= 1.0f) f = std::pow((f + 0.055f) / 1.055f, 2.4f);
    else f = 1.0f;
    return f;
}

}  // namespace

namespace ntx {

ntx_status fail(ntx_hip_context *ctx, hipError_t e, const char *what)
{
    if (ctx) ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? NTX_ERR_OUT_OF_MEMORY : NTX_ERR_DEVICE;
}


hipEvent_t take_event(ntx_hip_context *ctx)
{
    if (!ctx->spare.empty()) {
        hipEvent_t e = ctx->spare.back();
        ctx->spare.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}


ntx_status ensure(ntx_hip_context *ctx, DevBuf &b, size_t bytes)
{
    if (b.cap >= bytes) return NTX_OK;
    // The context's scratch stays within its limit (ntx_hip_set_scratch_limit).
    if (ctx->scratch_limit && ctx->scratch_bytes - b.cap + bytes > ctx->scratch_limit) {
        ctx->err = "scratch of " + std::to_string(ctx->scratch_bytes - b.cap + bytes) + " bytes exceeds the limit of " +
                   std::to_string(ctx->scratch_limit);
        return NTX_ERR_OUT_OF_MEMORY;
    }
    if (b.ptr) NTX_HIP(ctx, hipFree(b.ptr));
    ctx->scratch_bytes -= b.cap;
    b.ptr = nullptr;
    b.cap = 0;
    NTX_HIP(ctx, hipMalloc(&b.ptr, bytes));
    b.cap = bytes;
    ctx->scratch_bytes += bytes;
    return NTX_OK;
}

void release(ntx_hip_context *ctx, DevBuf &b)
{
    if (b.ptr) (void)hipFree(b.ptr);
    ctx->scratch_bytes -= b.cap;
    b.ptr = nullptr;
    b.cap = 0;
}

ntx_status bind(ntx_hip_context *ctx)
{
    if (!ctx) return NTX_ERR_INVALID_ARGUMENT;
    NTX_HIP(ctx, hipSetDevice(ctx->device));
    return NTX_OK;
}

// Quality levels handled exactly by the kernels (qualityLevel in
// BlockCompressor.cpp:202-209; chooseCpuCompressor, Context.cpp:1038-1163).
// BlockCompressor.cpp:211-217 qualityLevel: Fastest -> icbc Level1 (box fit +
// least squares), Production -> Level9 (cluster fit + refine_endpoints),
// Normal and Highest -> Level8.
int bc1_level(int quality)
{
    return quality == NTX_QUALITY_FASTEST ? 1 : quality == NTX_QUALITY_PRODUCTION ? 9 : 8;
}

bool supported(int format, int quality)
{
    switch (format) {
    // BlockCompressor.cpp:211-217: every quality maps to an icbc level.
    case NTX_FORMAT_BC1:
        return quality >= NTX_QUALITY_FASTEST && quality <= NTX_QUALITY_HIGHEST;
    // Context.cpp:1058-1066: CompressorDXT1a above Fastest.  FastCompressorDXT1a
    // reads uninitialised memory (QuickCompressDXT.cpp:378-403, 740-760).
    case NTX_FORMAT_BC1A: return quality >= NTX_QUALITY_NORMAL && quality <= NTX_QUALITY_HIGHEST;
    // Context.cpp:1080-1116: Fastest takes FastCompressorDXT5 / DXT5n,
    // Highest OptimalCompress for the BC3 alpha and both BC3n halves
    // (CompressorDX9.cpp:67-80, 145-222); BC4 / BC5 take the Production
    // compressors at Production and Highest (CompressorDX10.cpp:42-87).
    case NTX_FORMAT_BC3:
    case NTX_FORMAT_BC3N:
    case NTX_FORMAT_BC4:
    case NTX_FORMAT_BC5: return quality >= NTX_QUALITY_FASTEST && quality <= NTX_QUALITY_HIGHEST;
    default: return false;
    }
}

// Lanes per block of the BC3 colour search (as a power of two): one block
// per lane while the image fills 2^target_log2 lanes, otherwise enough lanes
// per block to reach about that many, up to a whole wavefront per block.  The
// search time of a small image is set by the longest block, so splitting each
// block's candidates over more lanes shortens it; each extra lane repeats the
// block's set-up, so a level that runs beside other work (a chain) aims lower.
int bc3_lanes_shift(int nblocks, int target_log2)
{
    int s = 0;
    while (s < 6 && (size_t(nblocks) << s) < (size_t(1) << target_log2)) s++;
    return s;
}

// Whether a BC3 / BC3n level's colour half is the squish cluster fit.
bool bc3_cluster_fit(int format, int quality)
{
    return quality != NTX_QUALITY_FASTEST && !(format == NTX_FORMAT_BC3N && quality == NTX_QUALITY_HIGHEST);
}

bool bc1_two_pass(const ntx_hip_context *ctx, int quality, int w, int h)
{
    const size_t nblocks = size_t((w + 3) / 4) * size_t((h + 3) / 4);
    return bc1_level(quality) == 8 && nblocks >= size_t(ctx->bc1_lanes_min);
}

size_t scratch_ints(const ntx_hip_context *ctx, int format, int quality, int w, int h)
{
    const int nblocks = ((w + 3) / 4) * ((h + 3) / 4);
    // The weighted-cluster-fit colour of BC3 / BC3n buckets its blocks;
    // Fastest and the BC3n Highest colour (compressDXT1G) need no scratch.
    if (format == NTX_FORMAT_BC3 || format == NTX_FORMAT_BC3N)
        return bc3_cluster_fit(format, quality) ? size_t(bc3_scratch_ints(nblocks)) : 0;
    if (format == NTX_FORMAT_BC1A) return size_t(bc3_scratch_ints(nblocks));
    if (format == NTX_FORMAT_BC1 && bc1_two_pass(ctx, quality, w, h)) return size_t(bc1_scratch_ints(nblocks));
    return 0;
}

ntx_status poly_tables(ntx_hip_context *ctx, int filter, float width, const float *params, int src, int dst, PolyDev **out)
{
    const float alpha = (filter == NTX_FILTER_KAISER && params) ? params[0] : 4.0f;
    const float stretch = (filter == NTX_FILTER_KAISER && params) ? params[1] : 1.0f;
    PolyKey key(filter, fbits(width), fbits(alpha), fbits(stretch), src, dst);
    auto it = ctx->poly.find(key);
    if (it != ctx->poly.end()) { *out = &it->second; return NTX_OK; }
    ntx::Kernel1D k;
    const float p[2] = { alpha, stretch };
    ntx::build_polyphase(filter, width, p, src, dst, &k);
    // applyKernelX/Y (FloatImage.cpp:1118-1133): first source tap per output.
    std::vector<int> left(dst);
    for (int i = 0; i < dst; i++) left[i] = int(std::floor((0.5f + i) * k.iscale - k.width));
    PolyDev d;
    d.window = k.window;
    d.span_x = max_span(left, k.window, ntx::poly_tile_x());
    d.span_y = max_span(left, k.window, ntx::poly_tile_y());
    d.stride2 = true;
    for (int i = 0; i < dst; i++) d.stride2 = d.stride2 && left[i] == left[0] + 2 * i;
    d.uniform = d.stride2 && k.window <= 16;
    for (int i = 1; d.uniform && i < dst; i++)
        d.uniform = std::memcmp(k.weights + size_t(i) * k.window, k.weights, sizeof(float) * size_t(k.window)) == 0;
    d.left0 = left[0];
    if (d.uniform) std::memcpy(d.w0, k.weights, sizeof(float) * size_t(k.window));
    d.host_weights.assign(k.weights, k.weights + size_t(k.window) * dst);
    d.host_left = left;
    hipError_t e = hipMalloc(&d.weights, sizeof(float) * size_t(k.window) * dst);
    if (e == hipSuccess) e = hipMalloc(&d.left, sizeof(int) * dst);
    if (e == hipSuccess) e = hipMemcpy(d.weights, k.weights, sizeof(float) * size_t(k.window) * dst, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.left, left.data(), sizeof(int) * dst, hipMemcpyHostToDevice);
    ntx::free_polyphase(&k);
    if (e != hipSuccess) {
        if (d.weights) (void)hipFree(d.weights);
        if (d.left) (void)hipFree(d.left);
        return fail(ctx, e, "polyphase table upload");
    }
    auto res = ctx->poly.emplace(key, d);
    *out = &res.first->second;
    return NTX_OK;
}

}  // namespace ntx

using namespace ntx;

extern "C" {

NTX_API int ntx_hip_is_hardware_present(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0;
}

NTX_API ntx_status ntx_hip_device_count(int *count)
{
    if (!count) return NTX_ERR_INVALID_ARGUMENT;
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) { *count = 0; return NTX_ERR_NO_DEVICE; }
    if (e != hipSuccess) return NTX_ERR_DEVICE;
    return *count > 0 ? NTX_OK : NTX_ERR_NO_DEVICE;
}

NTX_API const char *ntx_hip_status_string(ntx_status s)
{
    switch (s) {
    case NTX_OK: return "ok";
    case NTX_ERR_INVALID_ARGUMENT: return "invalid argument";
    case NTX_ERR_UNSUPPORTED: return "unsupported format or quality";
    case NTX_ERR_DEVICE: return "HIP device error";
    case NTX_ERR_NO_DEVICE: return "no HIP device";
    case NTX_ERR_OUT_OF_MEMORY: return "out of device memory";
    }
    return "unknown status";
}

NTX_API ntx_status ntx_hip_create(int device, ntx_hip_context **out)
{
    if (!out) return NTX_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return NTX_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return NTX_ERR_INVALID_ARGUMENT;
    ntx_hip_context *ctx = new ntx_hip_context;
    ctx->device = device;
    auto cleanup = [&](ntx_status s) { ntx_hip_destroy(ctx); return s; };
    if (hipSetDevice(device) != hipSuccess) return cleanup(NTX_ERR_DEVICE);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->cus = cus;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(NTX_ERR_DEVICE);
    if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) return cleanup(NTX_ERR_DEVICE);
    ntx::Tables *tab = new ntx::Tables;
    ntx::build_tables(tab);
    float lut[256];
    for (int k = 0; k < 256; k++) lut[k] = from_srgb(float(k) / 255.0f);
    hipError_t e = hipMalloc(&ctx->d_tab, sizeof(ntx::Tables));
    if (e == hipSuccess) e = hipMemcpy(ctx->d_tab, tab, sizeof(ntx::Tables), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&ctx->d_srgb, sizeof(lut));
    if (e == hipSuccess) e = hipMemcpy(ctx->d_srgb, lut, sizeof(lut), hipMemcpyHostToDevice);
    delete tab;
    if (e != hipSuccess) return cleanup(e == hipErrorOutOfMemory ? NTX_ERR_OUT_OF_MEMORY : NTX_ERR_DEVICE);
    if (ntx_status s = ntx::noalloc_init(ctx)) return cleanup(s);
    *out = ctx;
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_destroy(ntx_hip_context *ctx)
{
    if (!ctx) return NTX_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->poly) {
        (void)hipFree(kv.second.weights);
        (void)hipFree(kv.second.left);
    }
    for (auto &kv : ctx->tail_taps) (void)hipFree(kv.second.packed);
    for (hipStream_t st : ctx->side)
        if (st) (void)hipStreamSynchronize(st);
    for (DevBuf *b : { &ctx->in, &ctx->out, &ctx->tmp, &ctx->dst, &ctx->bucket, &ctx->metric, &ctx->chain })
        if (b->ptr) (void)hipFree(b->ptr);
    for (hipEvent_t e : ctx->sync_events) (void)hipEventDestroy(e);
    for (hipStream_t st : ctx->side)
        if (st) (void)hipStreamDestroy(st);
    if (ctx->d_tab) (void)hipFree(ctx->d_tab);
    if (ctx->d_srgb) (void)hipFree(ctx->d_srgb);
    if (ctx->d_gamma_lut) (void)hipFree(ctx->d_gamma_lut);
    for (auto &m : ctx->marks) { (void)hipEventDestroy(m.start); (void)hipEventDestroy(m.stop); }
    for (hipEvent_t e : ctx->spare) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return NTX_OK;
}

NTX_API const char *ntx_hip_last_error(const ntx_hip_context *ctx) { return ctx ? ctx->err.c_str() : ""; }

NTX_API void *ntx_hip_stream(ntx_hip_context *ctx) { return ctx ? reinterpret_cast<void *>(ctx->stream) : nullptr; }

NTX_API ntx_status ntx_hip_malloc(ntx_hip_context *ctx, size_t bytes, void **dptr)
{
    if (!dptr) return NTX_ERR_INVALID_ARGUMENT;
    ntx_status s = bind(ctx);
    if (s) return s;
    NTX_HIP(ctx, hipMalloc(dptr, bytes ? bytes : 1));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_free(ntx_hip_context *ctx, void *dptr)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (dptr) NTX_HIP(ctx, hipFree(dptr));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_upload(ntx_hip_context *ctx, void *dst, const void *src, size_t bytes)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    NTX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_download(ntx_hip_context *ctx, void *dst, const void *src, size_t bytes)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    NTX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_copy_device(ntx_hip_context *ctx, void *dst, const void *src, size_t bytes)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (bytes && (!dst || !src)) return NTX_ERR_INVALID_ARGUMENT;
    NTX_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_memset_device(ntx_hip_context *ctx, void *dst, int value, size_t bytes)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (bytes && !dst) return NTX_ERR_INVALID_ARGUMENT;
    NTX_HIP(ctx, hipMemsetAsync(dst, value, bytes, ctx->stream));
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_synchronize(ntx_hip_context *ctx)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_timer_begin(ntx_hip_context *ctx)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    NTX_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_timer_end(ntx_hip_context *ctx, float *ms)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    NTX_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    NTX_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float t = 0.0f;
    NTX_HIP(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    if (ms) *ms = t;
    return NTX_OK;
}

NTX_API const char *ntx_hip_kernel_name(int kernel)
{
    static const char *names[NTX_KERNEL_COUNT] = { "bc1_kernel",    "bc3_search_kernel", "channel_kernel",
                                                   "box_kernel",    "poly_x_kernel",    "poly_y_kernel",
                                                   "srgb_kernel",   "unpack_kernel",    "poly_fused_kernel",
                                                   "bc3_front_kernel", "batch_kernel", "bc1_mlp_kernel",
                                                   "bc1_list_kernel", "mlp_train_kernel",
                                                   "bc3_search_split_kernel", "gamma_kernel",
                                                   "normal_kernel", "bc1_classify_kernel",
                                                   "bc1_search_kernel", "optimal_channel_kernel",
                                                   "optimal_green_kernel", "bc3_fast_kernel",
                                                   "bc1a_front_kernel+bc3_search_kernel", "mip_tail_kernel",
                                                   "poly_slide_kernel" };
    return (kernel >= 0 && kernel < NTX_KERNEL_COUNT) ? names[kernel] : "";
}

NTX_API ntx_status ntx_hip_profile_enable(ntx_hip_context *ctx, int enable)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    ctx->profiling = enable != 0;
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_profile_read(ntx_hip_context *ctx, float *total_ms, int *launches, int reset)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < NTX_KERNEL_COUNT; k++) {
        if (total_ms) total_ms[k] = 0.0f;
        if (launches) launches[k] = 0;
    }
    for (auto &m : ctx->marks) {
        float ms = 0.0f;
        NTX_HIP(ctx, hipEventElapsedTime(&ms, m.start, m.stop));
        if (total_ms) total_ms[m.kernel] += ms;
        if (launches) launches[m.kernel] += 1;
    }
    if (reset) {
        for (auto &m : ctx->marks) { ctx->spare.push_back(m.start); ctx->spare.push_back(m.stop); }
        ctx->marks.clear();
    }
    return NTX_OK;
}

NTX_API size_t ntx_hip_compressed_size(int format, int w, int h)
{
    if (w <= 0 || h <= 0) return 0;
    const size_t bs = (format == NTX_FORMAT_BC1 || format == NTX_FORMAT_BC1A || format == NTX_FORMAT_BC4) ? 8 : 16;
    return size_t((w + 3) / 4) * size_t((h + 3) / 4) * bs;
}

}  // extern "C"

namespace ntx {

// One level's compression enqueued on `stream`.  BC3 uses `bucket`
// (bc3_scratch_ints(blocks) ints) whose tallies are zero when *counts_zero;
// the flag is left true only after a complete front + search sequence.
// lanes_log2: the BC3 search's lane target (bc3_lanes_shift); urgent: its
// small launches raise their wave priority.
static hipError_t compress_on(ntx_hip_context *ctx, hipStream_t stream, int format, int quality, int alpha_mode,
                              const float *cw, const float *d_rgba, int w, int h, void *d_out, int *bucket,
                              bool *counts_zero, int lanes_log2 = 18, bool urgent = false, int16_t *d_nm = nullptr)
{
    hipError_t e = hipSuccess;
    if (format == NTX_FORMAT_BC1 && bucket && bc1_two_pass(ctx, quality, w, h)) {
        const int nblocks = ((w + 3) / 4) * ((h + 3) / 4);
        int *counts = bucket, *lists = counts + bc3_lists_offset();
        if (!*counts_zero) e = launch_bc3_clear_counts(counts, stream);
        *counts_zero = false;
        if (e == hipSuccess)
            e = timed_on(ctx, stream, NTX_KERNEL_BC1_CLASSIFY, [&] {
                return launch_bc1_classify(d_rgba, w, h, alpha_mode, ctx->d_tab, d_out, d_nm, nullptr, nullptr, counts,
                                           lists, nblocks, 0, stream, urgent);
            });
        if (e == hipSuccess)
            e = timed_on(ctx, stream, NTX_KERNEL_BC1_SEARCH, [&] {
                return launch_bc1_search(d_rgba, w, h, alpha_mode, cw, ctx->d_tab, d_out, d_nm, counts, lists, nblocks,
                                         kBc1SearchGroupsPerCU * ctx->cus, stream);
            });
        if (e == hipSuccess) *counts_zero = true;
    } else if (format == NTX_FORMAT_BC1) {
        e = timed_on(ctx, stream, NTX_KERNEL_BC1, [&] {
            return launch_bc1(d_rgba, w, h, alpha_mode, cw, ctx->d_tab, d_out, d_nm, stream, urgent,
                              bc1_level(quality));
        });
    } else if (format == NTX_FORMAT_BC1A) {
        const int nblocks = ((w + 3) / 4) * ((h + 3) / 4);
        int *counts = bucket, *lists = counts + bc3_lists_offset();
        uint32_t *packed = reinterpret_cast<uint32_t *>(counts + bc3_packed_offset(nblocks));
        if (!*counts_zero) e = launch_bc3_clear_counts(counts, stream);
        *counts_zero = false;
        if (e == hipSuccess)
            e = timed_on(ctx, stream, NTX_KERNEL_BC1A, [&] {
                return launch_bc1a(d_rgba, w, h, alpha_mode == NTX_ALPHA_TRANSPARENCY ? 1 : 0, cw, ctx->d_tab, d_out,
                                   counts, lists, nblocks, packed, stream);
            });
        if (e == hipSuccess) *counts_zero = true;
    } else if ((format == NTX_FORMAT_BC3 || format == NTX_FORMAT_BC3N) && quality == NTX_QUALITY_FASTEST) {
        e = timed_on(ctx, stream, NTX_KERNEL_BC3_FAST, [&] {
            return launch_bc3_fast(d_rgba, w, h, format == NTX_FORMAT_BC3N, ctx->d_tab, d_out, stream);
        });
    } else if (format == NTX_FORMAT_BC3N && quality == NTX_QUALITY_HIGHEST) {
        // CompressorDXT5n at Highest: compressDXT1G colour, OptimalCompress red.
        e = timed_on(ctx, stream, NTX_KERNEL_OPTIMAL_GREEN, [&] {
            return launch_optimal_green(d_rgba, w, h, ctx->d_tab, d_out, kOptimalGroupsPerCU * ctx->cus, stream);
        });
        if (e == hipSuccess)
            e = timed_on(ctx, stream, NTX_KERNEL_OPTIMAL_CHANNEL, [&] {
                return launch_optimal_channels(d_rgba, w, h, format, d_out, kOptimalGroupsPerCU * ctx->cus, stream);
            });
    } else if ((format == NTX_FORMAT_BC4 || format == NTX_FORMAT_BC5) && quality >= NTX_QUALITY_PRODUCTION) {
        e = timed_on(ctx, stream, NTX_KERNEL_OPTIMAL_CHANNEL, [&] {
            return launch_optimal_channels(d_rgba, w, h, format, d_out, kOptimalGroupsPerCU * ctx->cus, stream);
        });
    } else if (format == NTX_FORMAT_BC3 || format == NTX_FORMAT_BC3N) {
        // CompressorDXT5n: metric (0, 1, 0) whatever the options say.
        const float green[3] = { 0.0f, 1.0f, 0.0f };
        const bool normal = format == NTX_FORMAT_BC3N;
        if (normal) cw = green;
        const int nblocks = ((w + 3) / 4) * ((h + 3) / 4);
        int *counts = bucket, *lists = counts + bc3_lists_offset();
        uint32_t *packed = reinterpret_cast<uint32_t *>(counts + bc3_packed_offset(nblocks));
        const int wba = alpha_mode == NTX_ALPHA_TRANSPARENCY ? 1 : 0;
        const int shift = bc3_lanes_shift(nblocks, lanes_log2);
        if (!*counts_zero) e = launch_bc3_clear_counts(counts, stream);
        *counts_zero = false;
        if (e == hipSuccess)
            e = timed_on(ctx, stream, NTX_KERNEL_BC3_CLASSIFY, [&] {
                return launch_bc3_front(d_rgba, w, h, ctx->d_tab, d_out, counts, lists, nblocks, wba, packed, urgent,
                                        stream, normal);
            });
        if (e == hipSuccess)
            e = timed_on(ctx, stream, shift ? NTX_KERNEL_BC3_SEARCH_SPLIT : NTX_KERNEL_BC3_SEARCH, [&] {
                return launch_bc3_search(packed, w, h, wba, cw, d_out, counts, lists, nblocks, ctx->d_tab, shift, urgent,
                                         stream);
            });
        if (e == hipSuccess) *counts_zero = true;
        // CompressorDXT5 at Highest: the alpha half by OptimalCompress,
        // over the QuickCompress half the front wrote.
        if (e == hipSuccess && quality == NTX_QUALITY_HIGHEST)
            e = timed_on(ctx, stream, NTX_KERNEL_OPTIMAL_CHANNEL, [&] {
                return launch_optimal_channels(d_rgba, w, h, format, d_out, kOptimalGroupsPerCU * ctx->cus, stream);
            });
    } else {
        e = timed_on(ctx, stream, NTX_KERNEL_CHANNEL,
                     [&] { return launch_channels(d_rgba, w, h, format, d_out, stream, urgent); });
    }
    return e;
}

hipError_t compress_scratch(ntx_hip_context *ctx, int format, int quality, int alpha_mode, const float *cw,
                            const float *d_rgba, int w, int h, void *d_out, int *bucket)
{
    bool counts_zero = false;  // cleared by a fill kernel first
    return compress_on(ctx, ctx->stream, format, quality, alpha_mode, cw, d_rgba, w, h, d_out, bucket, &counts_zero);
}

// Surface::buildNextMipmap of one level enqueued on `stream` (the fallback
// two-pass form uses the context's temporary, so only one stream may run it).
static ntx_status downsample_on(ntx_hip_context *ctx, hipStream_t stream, int filter, float width, const float *params,
                                int wrap, int alpha_mode, const float *d_rgba, int W, int H, float *d_out)
{
    ntx_status s = NTX_OK;
    hipError_t e;
    if (filter == NTX_FILTER_BOX && width == 0.5f && alpha_mode != NTX_ALPHA_TRANSPARENCY) {
        e = timed_on(ctx, stream, NTX_KERNEL_BOX, [&] { return launch_box(d_rgba, W, H, d_out, stream); });
    } else {
        const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
        PolyDev *kx = nullptr, *ky = nullptr;
        if ((s = poly_tables(ctx, filter, width, params, W, w, &kx))) return s;
        if ((s = poly_tables(ctx, filter, width, params, H, h, &ky))) return s;
        const int win = kx->window;
        const int rows = poly_slide_rows(W, H, ctx->cus);
        if (kx->uniform && ky->uniform && ky->window == win && (win & 1) && win >= 3 && win <= 15 && rows > 0) {
            e = timed_on(ctx, stream, NTX_KERNEL_POLY_SLIDE, [&] {
                return launch_poly_slide(d_rgba, W, H, d_out, kx->w0, kx->left0, ky->w0, ky->left0, win, wrap, rows,
                                         stream);
            });
        } else if (kx->uniform && ky->uniform && ky->window == win && poly_fused_uniform_window(win) &&
                   poly_fused_fits(kx->span_x, ky->span_y, kx->window, ky->window)) {
            e = timed_on(ctx, stream, NTX_KERNEL_POLY_FUSED, [&] {
                return launch_poly_fused_uniform(d_rgba, W, H, d_out, kx->w0, kx->left0, ky->w0, ky->left0, win, wrap,
                                                 stream);
            });
        } else if (poly_fused_fits(kx->span_x, ky->span_y, kx->window, ky->window)) {
            e = timed_on(ctx, stream, NTX_KERNEL_POLY_FUSED, [&] {
                return launch_poly_fused(d_rgba, W, H, d_out, kx->weights, kx->left, kx->window, kx->stride2, ky->weights,
                                         ky->left, ky->window, wrap, stream);
            });
        } else {
            // Small images with wide windows: X pass into a temporary, then Y.
            if ((s = ensure(ctx, ctx->tmp, size_t(w) * size_t(H) * 4 * sizeof(float)))) return s;
            float *tmp = static_cast<float *>(ctx->tmp.ptr);
            e = timed_on(ctx, stream, NTX_KERNEL_POLY_X, [&] {
                return launch_poly_x(d_rgba, W, H, tmp, kx->weights, kx->left, kx->window, wrap, stream);
            });
            if (e == hipSuccess)
                e = timed_on(ctx, stream, NTX_KERNEL_POLY_Y, [&] {
                    return launch_poly_y(tmp, W, H, d_out, ky->weights, ky->left, ky->window, wrap, stream);
                });
        }
    }
    if (e != hipSuccess) return fail(ctx, e, "downsample launch");
    return NTX_OK;
}

// FloatImage.h:298-330 wrap helpers, as the kernels apply them.
static int host_wrap_index(int x, int n, int wm)
{
    if (x >= 0 && x < n) return x;
    if (wm == NTX_WRAP_CLAMP) return x < 0 ? 0 : n - 1;
    if (wm == NTX_WRAP_REPEAT) return x >= 0 ? x % n : (x + 1) % n + n - 1;
    if (n == 1) return 0;
    x = std::abs(x);
    while (x >= n) x = std::abs(n + n - x - 2);
    return x;
}

// The packed taps of a tail (MipTail::packed) from source W x H through
// `levels` halvings, built once per filter, size and wrap and kept on the
// device, so the tail kernel copies them into LDS with one coalesced pass
// instead of resolving each tap from the tables.
static ntx_status tail_taps(ntx_hip_context *ctx, int filter, float width, const float *params, int wrap, int W0, int H0,
                            int levels, TailTaps **out)
{
    const float alpha = (filter == NTX_FILTER_KAISER && params) ? params[0] : 4.0f;
    const float stretch = (filter == NTX_FILTER_KAISER && params) ? params[1] : 1.0f;
    const TailKey key(filter, fbits(width), fbits(alpha), fbits(stretch), W0, H0, levels, wrap);
    auto it = ctx->tail_taps.find(key);
    if (it != ctx->tail_taps.end()) { *out = &it->second; return NTX_OK; }
    std::vector<float> pk;
    int W = W0, H = H0;
    for (int k = 0; k < levels; k++) {
        const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
        PolyDev *kx = nullptr, *ky = nullptr;
        ntx_status s = poly_tables(ctx, filter, width, params, W, w, &kx);
        if (!s) s = poly_tables(ctx, filter, width, params, H, h, &ky);
        if (s) return s;
        const int wx = kx->window, wy = ky->window;
        if (int(pk.size() / 2) + w * wx + h * wy <= kTailTaps) {
            auto push = [&](float weight, int off) {
                float bits;
                std::memcpy(&bits, &off, sizeof bits);
                pk.push_back(weight);
                pk.push_back(bits);
            };
            for (int x = 0; x < w; x++)
                for (int j = 0; j < wx; j++)
                    push(kx->host_weights[size_t(x) * wx + j], host_wrap_index(kx->host_left[x] + j, W, wrap));
            for (int y = 0; y < h; y++)
                for (int j = 0; j < wy; j++)
                    push(ky->host_weights[size_t(y) * wy + j], host_wrap_index(ky->host_left[y] + j, H, wrap) * w);
        }
        W = w;
        H = h;
    }
    TailTaps t;
    t.count = int(pk.size() / 2);
    if (t.count > 0) {
        hipError_t e = hipMalloc(&t.packed, sizeof(float) * pk.size());
        if (e == hipSuccess) e = hipMemcpy(t.packed, pk.data(), sizeof(float) * pk.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (t.packed) (void)hipFree(t.packed);
            return fail(ctx, e, "tail tap upload");
        }
    }
    auto res = ctx->tail_taps.emplace(key, t);
    *out = &res.first->second;
    return NTX_OK;
}

// Levels 1 .. levels-1 of a chain on `stream`: a launch per level while the
// source is large, then every remaining level in one launch (launch_mip_tail)
// once the source fits a workgroup's LDS, when two or more levels remain.
static ntx_status downsample_chain_on(ntx_hip_context *ctx, hipStream_t stream, int filter, float width,
                                      const float *params, int wrap, int alpha_mode, const int *lw, const int *lh,
                                      int levels, float *const *d_levels)
{
    const bool box = filter == NTX_FILTER_BOX && width == 0.5f && alpha_mode != NTX_ALPHA_TRANSPARENCY;
    for (int k = 1; k < levels; k++) {
        const int rest = levels - k;
        if (rest >= 2 && rest <= kTailMaxLevels && size_t(lw[k - 1]) * size_t(lh[k - 1]) <= size_t(kTailSrc)) {
            MipTail tail{};
            tail.W = lw[k - 1];
            tail.H = lh[k - 1];
            tail.levels = rest;
            tail.box = box ? 1 : 0;
            for (int j = 0; j < rest; j++) {
                tail.dst[j] = d_levels[k + j];
                if (box) continue;
                PolyDev *kx = nullptr, *ky = nullptr;
                ntx_status s = poly_tables(ctx, filter, width, params, lw[k - 1 + j], lw[k + j], &kx);
                if (!s) s = poly_tables(ctx, filter, width, params, lh[k - 1 + j], lh[k + j], &ky);
                if (s) return s;
                tail.kx[j] = kx->weights;
                tail.lx[j] = kx->left;
                tail.wx[j] = kx->window;
                tail.ky[j] = ky->weights;
                tail.ly[j] = ky->left;
                tail.wy[j] = ky->window;
            }
            if (!box) {
                TailTaps *tt = nullptr;
                const ntx_status s = tail_taps(ctx, filter, width, params, wrap, lw[k - 1], lh[k - 1], rest, &tt);
                if (s) return s;
                tail.packed = tt->packed;
                tail.npacked = tt->count;
            }
            const hipError_t e = timed_on(ctx, stream, NTX_KERNEL_MIP_TAIL,
                                          [&] { return launch_mip_tail(d_levels[k - 1], tail, wrap, stream); });
            if (e != hipSuccess) return fail(ctx, e, "mip tail launch");
            return NTX_OK;
        }
        const ntx_status s = downsample_on(ctx, stream, filter, width, params, wrap, alpha_mode, d_levels[k - 1],
                                           lw[k - 1], lh[k - 1], d_levels[k]);
        if (s) return s;
    }
    return NTX_OK;
}

static bool valid_mip_args(int filter, int wrap)
{
    return filter >= NTX_FILTER_BOX && filter <= NTX_FILTER_KAISER && wrap >= NTX_WRAP_CLAMP && wrap <= NTX_WRAP_MIRROR;
}

// A timing-free event from the context's pool of fork/join events.
static hipEvent_t sync_event(ntx_hip_context *ctx, size_t i)
{
    while (ctx->sync_events.size() <= i) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        ctx->sync_events.push_back(e);
    }
    return ctx->sync_events[i];
}

// The context's bucket scratch for one image (BC3, two-pass BC1), or null.
static ntx_status single_bucket(ntx_hip_context *ctx, int format, int quality, int w, int h, int **bucket)
{
    *bucket = nullptr;
    const size_t ints = scratch_ints(ctx, format, quality, w, h);
    if (!ints) return NTX_OK;
    const size_t cap_before = ctx->bucket.cap;
    if (ntx_status s = ensure(ctx, ctx->bucket, sizeof(int) * ints)) return s;
    if (ctx->bucket.cap != cap_before) ctx->bc3_counts_zero = false;  // fresh memory
    *bucket = static_cast<int *>(ctx->bucket.ptr);
    return NTX_OK;
}

}  // namespace ntx

extern "C" {

NTX_API ntx_status ntx_hip_compress_device(ntx_hip_context *ctx, int format, int quality, int alpha_mode,
                                           const float *cw, const float *d_rgba, int w, int h, void *d_out)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 0 || h < 0 || (size_t(w) * size_t(h) > 0 && (!d_rgba || !d_out))) return NTX_ERR_INVALID_ARGUMENT;
    if (!supported(format, quality)) return NTX_ERR_UNSUPPORTED;
    if (size_t(w) * size_t(h) == 0) return NTX_OK;
    const float ones[4] = { 1.0f, 1.0f, 1.0f, 1.0f };
    if (!cw) cw = ones;
    int *bucket = nullptr;
    if ((s = ntx::single_bucket(ctx, format, quality, w, h, &bucket))) return s;
    hipError_t e = ntx::compress_on(ctx, ctx->stream, format, quality, alpha_mode, cw, d_rgba, w, h, d_out, bucket,
                                    &ctx->bc3_counts_zero);
    if (e != hipSuccess) return fail(ctx, e, "compress launch");
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_set_scratch_limit(ntx_hip_context *ctx, size_t bytes)
{
    if (!ctx) return NTX_ERR_INVALID_ARGUMENT;
    ctx->scratch_limit = bytes;
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_set_bc1_lanes_min(ntx_hip_context *ctx, int min_blocks)
{
    if (!ctx || min_blocks < 0) return NTX_ERR_INVALID_ARGUMENT;
    ctx->bc1_lanes_min = min_blocks ? min_blocks : ntx_hip_context::kBc1LanesMinDefault;
    return NTX_OK;
}

}  // extern "C"

namespace ntx {

// n independent images compressed concurrently, in the caller's order (the
// largest first): image 0 on the context's stream, images 1 and 2 on side
// stream 1, the rest (short, latency-bound) one after another on side stream
// 0 with raised wave priority, so their launches fill the gaps the large
// ones leave rather than being stretched by the large searches sharing every
// SIMD.  Images after the first aim their BC3 search at 2^16 lanes, since
// they run beside other work.  The side streams start once `ready` (recorded
// on the context's stream) has passed; the context's stream then waits for
// both.  Three streams: with the default four hardware queues per process,
// one of which the null stream takes, more streams would share queues and
// run one after another.  BC3 scratch: one region per image in ctx->chain,
// laid out by the image sizes.
static ntx_status compress_many(ntx_hip_context *ctx, int format, int quality, int alpha_mode, const float *cw, int n,
                                const float *const *d_in, const int *w, const int *h, void *const *d_out)
{
    ntx_status s = NTX_OK;
    std::vector<size_t> region(n, 0);
    // The offsets of the regions that hold tallies: they depend on the
    // format and quality as well as the sizes (scratch_ints).
    std::vector<size_t> layout;
    size_t total = 0;
    for (int k = 0; k < n; k++) {
        region[k] = total;
        const size_t ints = scratch_ints(ctx, format, quality, w[k], h[k]);
        if (ints) layout.push_back(total);
        total += (ints + 63) & ~size_t(63);
    }
    if (total) {
        const size_t cap_before = ctx->chain.cap;
        if ((s = ensure(ctx, ctx->chain, sizeof(int) * total))) return s;
        // The tallies are known to be zero only in the same memory at the
        // same region offsets as the last complete run.
        if (ctx->chain.cap != cap_before || ctx->chain_layout != layout) ctx->chain_counts_zero = false;
        ctx->chain_layout = layout;
    }
    for (int i = 0; i < ntx_hip_context::kSideStreams; i++)
        if (!ctx->side[i]) NTX_HIP(ctx, hipStreamCreateWithFlags(&ctx->side[i], hipStreamNonBlocking));
    hipStream_t small = ctx->side[0], second = ctx->side[1];
    hipEvent_t ready = sync_event(ctx, 0), ev_small = sync_event(ctx, 1), ev_second = sync_event(ctx, 2);
    if (!ready || !ev_small || !ev_second) return fail(ctx, hipErrorOutOfMemory, "schedule events");
    bool zero = ctx->chain_counts_zero;
    ctx->chain_counts_zero = false;
    bool all_zero = true;
    auto compress_one = [&](hipStream_t st, int k) -> ntx_status {
        int *bucket = scratch_ints(ctx, format, quality, w[k], h[k]) ? static_cast<int *>(ctx->chain.ptr) + region[k]
                                                                       : nullptr;
        bool z = zero;
        hipError_t e = compress_on(ctx, st, format, quality, alpha_mode, cw, d_in[k], w[k], h[k], d_out[k], bucket, &z,
                                   k ? 16 : 18, st == small);
        if (e != hipSuccess) return fail(ctx, e, "compress launch");
        all_zero = all_zero && z;
        return NTX_OK;
    };
    if (n > 1) NTX_HIP(ctx, hipEventRecord(ready, ctx->stream));
    // Image 0 is enqueued first: the host takes a while to enqueue the many
    // small launches, and the longest compression should not wait for that.
    if ((s = compress_one(ctx->stream, 0))) return s;
    if (n > 1) {
        NTX_HIP(ctx, hipStreamWaitEvent(second, ready, 0));
        for (int k = 1; k < n && k <= 2; k++)
            if ((s = compress_one(second, k))) return s;
        NTX_HIP(ctx, hipEventRecord(ev_second, second));
    }
    if (n > 3) {
        NTX_HIP(ctx, hipStreamWaitEvent(small, ready, 0));
        for (int k = 3; k < n; k++)
            if ((s = compress_one(small, k))) return s;
        NTX_HIP(ctx, hipEventRecord(ev_small, small));
    }
    if (n > 1) NTX_HIP(ctx, hipStreamWaitEvent(ctx->stream, ev_second, 0));
    if (n > 3) NTX_HIP(ctx, hipStreamWaitEvent(ctx->stream, ev_small, 0));
    // Every image's search cleared its own region.
    ctx->chain_counts_zero = all_zero;
    return NTX_OK;
}

// Side stream i, made to wait for the work already on the context's stream.
ntx_status side_stream_after(ntx_hip_context *ctx, int i, hipStream_t *out)
{
    for (int k = 0; k < ntx_hip_context::kSideStreams; k++)
        if (!ctx->side[k]) NTX_HIP(ctx, hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking));
    hipEvent_t ready = sync_event(ctx, 5 + 2 * size_t(i));
    if (!ready) return fail(ctx, hipErrorOutOfMemory, "schedule events");
    NTX_HIP(ctx, hipEventRecord(ready, ctx->stream));
    NTX_HIP(ctx, hipStreamWaitEvent(ctx->side[i], ready, 0));
    *out = ctx->side[i];
    return NTX_OK;
}

// An event recorded on side stream i after the work enqueued on it.
ntx_status side_stream_done(ntx_hip_context *ctx, int i, hipEvent_t *out)
{
    hipEvent_t done = sync_event(ctx, 6 + 2 * size_t(i));
    if (!done) return fail(ctx, hipErrorOutOfMemory, "schedule events");
    NTX_HIP(ctx, hipEventRecord(done, ctx->side[i]));
    *out = done;
    return NTX_OK;
}

// BC1 images compressed by the one-pass cluster fit on side stream 0 with
// raised wave priority, after the work already on the context's stream; the
// context's stream goes on at once, and the caller makes it wait for *join
// (null when n is 0).  Used for the small levels beside the learned form.
ntx_status side_plain_bc1(ntx_hip_context *ctx, int n, const float *const *d_in, const int *w, const int *h,
                          void *const *d_out, hipEvent_t *join)
{
    *join = nullptr;
    if (n == 0) return NTX_OK;
    for (int i = 0; i < ntx_hip_context::kSideStreams; i++)
        if (!ctx->side[i]) NTX_HIP(ctx, hipStreamCreateWithFlags(&ctx->side[i], hipStreamNonBlocking));
    hipEvent_t ready = sync_event(ctx, 3), done = sync_event(ctx, 4);
    if (!ready || !done) return fail(ctx, hipErrorOutOfMemory, "schedule events");
    NTX_HIP(ctx, hipEventRecord(ready, ctx->stream));
    NTX_HIP(ctx, hipStreamWaitEvent(ctx->side[0], ready, 0));
    const float ones[4] = { 1.0f, 1.0f, 1.0f, 1.0f };
    for (int k = 0; k < n; k++) {
        bool z = true;
        const hipError_t e = compress_on(ctx, ctx->side[0], NTX_FORMAT_BC1, NTX_QUALITY_NORMAL, NTX_ALPHA_NONE, ones,
                                         d_in[k], w[k], h[k], d_out[k], nullptr, &z, 16, true);
        if (e != hipSuccess) return fail(ctx, e, "compress launch");
    }
    NTX_HIP(ctx, hipEventRecord(done, ctx->side[0]));
    *join = done;
    return NTX_OK;
}

}  // namespace ntx

extern "C" {

NTX_API ntx_status ntx_hip_compress_chain_device(ntx_hip_context *ctx, int format, int quality, int alpha_mode,
                                                 const float *cw, int filter, float filter_width,
                                                 const float *filter_params, int wrap, int w, int h, int levels,
                                                 float *const *d_levels, void *const *d_out)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 1 || h < 1 || levels < 1 || !d_levels || !d_out || !valid_mip_args(filter, wrap))
        return NTX_ERR_INVALID_ARGUMENT;
    int full = 1;
    for (int a = w, b = h; a > 1 || b > 1; a = a / 2 > 1 ? a / 2 : 1, b = b / 2 > 1 ? b / 2 : 1) full++;
    if (levels > full) return NTX_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < levels; k++)
        if (!d_levels[k] || !d_out[k]) return NTX_ERR_INVALID_ARGUMENT;
    if (!supported(format, quality)) return NTX_ERR_UNSUPPORTED;
    const float ones[4] = { 1.0f, 1.0f, 1.0f, 1.0f };
    if (!cw) cw = ones;
    std::vector<int> lw(levels), lh(levels);
    for (int k = 0; k < levels; k++) {
        lw[k] = k ? (lw[k - 1] / 2 > 1 ? lw[k - 1] / 2 : 1) : w;
        lh[k] = k ? (lh[k - 1] / 2 > 1 ? lh[k - 1] / 2 : 1) : h;
    }
    // Every level is built first on the context's stream: the filter's
    // workgroups need 50 KB of LDS each and would wait behind any search
    // running beside them.  Then the levels compress concurrently, largest
    // first (compress_many).
    if ((s = ntx::downsample_chain_on(ctx, ctx->stream, filter, filter_width, filter_params, wrap, alpha_mode,
                                      lw.data(), lh.data(), levels, d_levels)))
        return s;
    return ntx::compress_many(ctx, format, quality, alpha_mode, cw, levels, d_levels, lw.data(), lh.data(), d_out);
}

NTX_API ntx_status ntx_hip_compress_images_device(ntx_hip_context *ctx, int format, int quality, int alpha_mode,
                                                  const float *cw, int n, const float *const *d_rgba, const int *w,
                                                  const int *h, void *const *d_out)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (n < 0 || (n > 0 && (!d_rgba || !w || !h || !d_out))) return NTX_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < n; k++)
        if (w[k] < 1 || h[k] < 1 || !d_rgba[k] || !d_out[k]) return NTX_ERR_INVALID_ARGUMENT;
    if (!supported(format, quality)) return NTX_ERR_UNSUPPORTED;
    if (n == 0) return NTX_OK;
    const float ones[4] = { 1.0f, 1.0f, 1.0f, 1.0f };
    if (!cw) cw = ones;
    // Largest first (stable), the order compress_many schedules by.
    std::vector<int> order(n);
    for (int k = 0; k < n; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return size_t(w[a]) * size_t(h[a]) > size_t(w[b]) * size_t(h[b]); });
    std::vector<const float *> in(n);
    std::vector<int> sw(n), sh(n);
    std::vector<void *> out(n);
    for (int k = 0; k < n; k++) {
        in[k] = d_rgba[order[k]];
        sw[k] = w[order[k]];
        sh[k] = h[order[k]];
        out[k] = d_out[order[k]];
    }
    return ntx::compress_many(ctx, format, quality, alpha_mode, cw, n, in.data(), sw.data(), sh.data(), out.data());
}

NTX_API ntx_status ntx_hip_bc1_sse_device(ntx_hip_context *ctx, const float *d_rgba, int w, int h, const void *d_blocks,
                                          double *sse)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 1 || h < 1 || !d_rgba || !d_blocks || !sse) return NTX_ERR_INVALID_ARGUMENT;
    if ((s = ensure(ctx, ctx->metric, sizeof(double) * (size_t(ntx::bc1_sse_partials(w, h)) + 1)))) return s;
    double *part = static_cast<double *>(ctx->metric.ptr);
    double *total = part + ntx::bc1_sse_partials(w, h);
    NTX_HIP(ctx, ntx::launch_bc1_sse(d_rgba, w, h, d_blocks, part, total, ctx->stream));
    NTX_HIP(ctx, hipMemcpyAsync(sse, total, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_decode_device(ntx_hip_context *ctx, int format, int decoder, const void *d_blocks, int w, int h,
                                         float *d_rgba)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (format != NTX_FORMAT_BC1 && format != NTX_FORMAT_BC3 && format != NTX_FORMAT_BC3N && format != NTX_FORMAT_BC4 &&
        format != NTX_FORMAT_BC5)
        return NTX_ERR_UNSUPPORTED;
    if (decoder != NTX_DECODER_D3D10 && decoder != NTX_DECODER_D3D9) return NTX_ERR_UNSUPPORTED;
    // Surface.cpp:1058-1072: BC3n blocks decode as BC3 (BlockDXT5), unswizzled.
    if (format == NTX_FORMAT_BC3N) format = NTX_FORMAT_BC3;
    if (w < 0 || h < 0 || (size_t(w) * size_t(h) > 0 && (!d_blocks || !d_rgba))) return NTX_ERR_INVALID_ARGUMENT;
    NTX_HIP(ctx, ntx::launch_decode(format, decoder == NTX_DECODER_D3D9 ? 1 : 0, d_blocks, w, h, d_rgba,
                                    ctx->d_tab->div255, ctx->stream));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_rms_error_device(ntx_hip_context *ctx, int metric, const float *d_reference,
                                            const float *d_image, int w, int h, float *result)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (metric < NTX_METRIC_RMS_COLOR || metric > NTX_METRIC_RMS_ALPHA || w < 1 || h < 1 || !d_reference || !d_image ||
        !result)
        return NTX_ERR_INVALID_ARGUMENT;
    if ((s = ensure(ctx, ctx->metric, sizeof(double) * (size_t(ntx::rms_partials()) + 1)))) return s;
    double *part = static_cast<double *>(ctx->metric.ptr);
    double *total = part + ntx::rms_partials();
    const size_t n = size_t(w) * size_t(h);
    NTX_HIP(ctx, ntx::launch_rms(d_reference, d_image, n, metric, part, total, ctx->stream));
    double mse = 0.0;
    NTX_HIP(ctx, hipMemcpyAsync(&mse, total, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *result = float(std::sqrt(mse / double(uint32_t(n))));  // ErrorMetric.cpp:45
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_compress_bc1_stats_device(ntx_hip_context *ctx, int alpha_mode, const float *cw,
                                                     const float *d_rgba, int w, int h, void *d_out, int16_t *d_nm)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 0 || h < 0 || !d_nm || (size_t(w) * size_t(h) > 0 && (!d_rgba || !d_out))) return NTX_ERR_INVALID_ARGUMENT;
    if (size_t(w) * size_t(h) == 0) return NTX_OK;
    const float ones[4] = { 1.0f, 1.0f, 1.0f, 1.0f };
    int *bucket = nullptr;
    if ((s = ntx::single_bucket(ctx, NTX_FORMAT_BC1, NTX_QUALITY_NORMAL, w, h, &bucket))) return s;
    hipError_t e = ntx::compress_on(ctx, ctx->stream, NTX_FORMAT_BC1, NTX_QUALITY_NORMAL, alpha_mode, cw ? cw : ones,
                                    d_rgba, w, h, d_out, bucket, &ctx->bc3_counts_zero, 18, false, d_nm);
    if (e != hipSuccess) return fail(ctx, e, "bc1 launch");
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_compress(ntx_hip_context *ctx, int format, int quality, int alpha_mode, const float *cw,
                                    const float *rgba, int w, int h, void *out, size_t out_bytes)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 0 || h < 0) return NTX_ERR_INVALID_ARGUMENT;
    if (!supported(format, quality)) return NTX_ERR_UNSUPPORTED;
    const size_t need = ntx_hip_compressed_size(format, w, h);
    if (need == 0) return NTX_OK;
    if (!rgba || !out || out_bytes < need) return NTX_ERR_INVALID_ARGUMENT;
    const size_t in_bytes = size_t(w) * size_t(h) * 4 * sizeof(float);
    if ((s = ensure(ctx, ctx->in, in_bytes)) || (s = ensure(ctx, ctx->out, need))) return s;
    NTX_HIP(ctx, hipMemcpyAsync(ctx->in.ptr, rgba, in_bytes, hipMemcpyHostToDevice, ctx->stream));
    s = ntx_hip_compress_device(ctx, format, quality, alpha_mode, cw, static_cast<const float *>(ctx->in.ptr), w, h,
                                ctx->out.ptr);
    if (s) return s;
    NTX_HIP(ctx, hipMemcpyAsync(out, ctx->out.ptr, need, hipMemcpyDeviceToHost, ctx->stream));
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_downsample_device(ntx_hip_context *ctx, int filter, float width, const float *params, int wrap,
                                             int alpha_mode, const float *d_rgba, int W, int H, float *d_out)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (W < 1 || H < 1 || (W == 1 && H == 1) || !d_rgba || !d_out) return NTX_ERR_INVALID_ARGUMENT;
    if (!ntx::valid_mip_args(filter, wrap)) return NTX_ERR_INVALID_ARGUMENT;
    return ntx::downsample_on(ctx, ctx->stream, filter, width, params, wrap, alpha_mode, d_rgba, W, H, d_out);
}

NTX_API ntx_status ntx_hip_downsample_chain_device(ntx_hip_context *ctx, int filter, float filter_width,
                                                   const float *filter_params, int wrap, int alpha_mode, int w, int h,
                                                   int levels, float *const *d_levels)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 1 || h < 1 || levels < 1 || !d_levels || !ntx::valid_mip_args(filter, wrap)) return NTX_ERR_INVALID_ARGUMENT;
    std::vector<int> lw(1, w), lh(1, h);
    while (lw.back() > 1 || lh.back() > 1) {
        lw.push_back(lw.back() / 2 > 1 ? lw.back() / 2 : 1);
        lh.push_back(lh.back() / 2 > 1 ? lh.back() / 2 : 1);
    }
    if (levels > int(lw.size())) return NTX_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < levels; k++)
        if (!d_levels[k]) return NTX_ERR_INVALID_ARGUMENT;
    return ntx::downsample_chain_on(ctx, ctx->stream, filter, filter_width, filter_params, wrap, alpha_mode, lw.data(),
                                    lh.data(), levels, d_levels);
}

NTX_API ntx_status ntx_hip_downsample(ntx_hip_context *ctx, int filter, float width, const float *params, int wrap,
                                      int alpha_mode, const float *rgba, int W, int H, float *out)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (W < 1 || H < 1 || (W == 1 && H == 1) || !rgba || !out) return NTX_ERR_INVALID_ARGUMENT;
    const int w = W / 2 > 1 ? W / 2 : 1, h = H / 2 > 1 ? H / 2 : 1;
    const size_t in_bytes = size_t(W) * size_t(H) * 4 * sizeof(float);
    const size_t out_bytes = size_t(w) * size_t(h) * 4 * sizeof(float);
    if ((s = ensure(ctx, ctx->in, in_bytes)) || (s = ensure(ctx, ctx->dst, out_bytes))) return s;
    NTX_HIP(ctx, hipMemcpyAsync(ctx->in.ptr, rgba, in_bytes, hipMemcpyHostToDevice, ctx->stream));
    s = ntx_hip_downsample_device(ctx, filter, width, params, wrap, alpha_mode, static_cast<const float *>(ctx->in.ptr), W,
                                  H, static_cast<float *>(ctx->dst.ptr));
    if (s) return s;
    NTX_HIP(ctx, hipMemcpyAsync(out, ctx->dst.ptr, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_srgb_to_linear_device(ntx_hip_context *ctx, float *d_rgba, int w, int h)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 0 || h < 0 || (size_t(w) * size_t(h) > 0 && !d_rgba)) return NTX_ERR_INVALID_ARGUMENT;
    hipError_t e = timed(ctx, NTX_KERNEL_SRGB, [&] { return ntx::launch_srgb(d_rgba, size_t(w) * size_t(h) * 3, ctx->d_srgb, ctx->stream); });
    if (e != hipSuccess) return fail(ctx, e, "srgb launch");
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_unpack_bgra8_device(ntx_hip_context *ctx, const uint8_t *d_bgra, int w, int h, float *d_rgba)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 0 || h < 0 || (size_t(w) * size_t(h) > 0 && (!d_bgra || !d_rgba))) return NTX_ERR_INVALID_ARGUMENT;
    hipError_t e = timed(ctx, NTX_KERNEL_UNPACK, [&] {
        return ntx::launch_unpack(d_bgra, size_t(w) * size_t(h), d_rgba, nullptr, ctx->stream);
    });
    if (e != hipSuccess) return fail(ctx, e, "unpack launch");
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_unpack_bgra8_srgb_device(ntx_hip_context *ctx, const uint8_t *d_bgra, int w, int h,
                                                    float *d_rgba)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 0 || h < 0 || (size_t(w) * size_t(h) > 0 && (!d_bgra || !d_rgba))) return NTX_ERR_INVALID_ARGUMENT;
    hipError_t e = timed(ctx, NTX_KERNEL_UNPACK, [&] {
        return ntx::launch_unpack(d_bgra, size_t(w) * size_t(h), d_rgba, ctx->d_srgb, ctx->stream);
    });
    if (e != hipSuccess) return fail(ctx, e, "unpack launch");
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_gamma_levels_device(ntx_hip_context *ctx, int to_gamma, float gamma, int n,
                                               const float *const *d_src, const int *w, const int *h,
                                               float *const *d_dst)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (n < 0 || (n > 0 && (!d_src || !w || !h || !d_dst)) || !(gamma > 0.0f)) return NTX_ERR_INVALID_ARGUMENT;
    for (int l = 0; l < n; l++)
        if (w[l] < 0 || h[l] < 0 || (size_t(w[l]) * size_t(h[l]) > 0 && (!d_src[l] || !d_dst[l])))
            return NTX_ERR_INVALID_ARGUMENT;
    // FloatImage.cpp:258-284: gamma 2.2f takes the table forms, anything
    // else powf(max(0, x), gamma) to linear or powf(max(0, x), 1 / gamma) back.
    const int mode = gamma == 2.2f ? (to_gamma ? 1 : 0) : 2;
    const float power = to_gamma ? 1.0f / gamma : gamma;
    for (int first = 0; first < n; first += ntx::kMaxGammaLevels) {
        ntx::GammaJob job{};
        job.levels = std::min(ntx::kMaxGammaLevels, n - first);
        for (int l = 0; l < job.levels; l++) {
            job.src[l] = d_src[first + l];
            job.dst[l] = d_dst[first + l];
            job.texels[l] = (unsigned long long)w[first + l] * (unsigned long long)h[first + l];
        }
        hipError_t e = timed(ctx, NTX_KERNEL_GAMMA, [&] { return ntx::launch_gamma(job, mode, power, ctx->d_tab, ctx->stream); });
        if (e != hipSuccess) return fail(ctx, e, "gamma launch");
    }
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_gamma_device(ntx_hip_context *ctx, int to_gamma, float gamma, const float *d_src, int w, int h,
                                        float *d_dst)
{
    return ntx_hip_gamma_levels_device(ctx, to_gamma, gamma, 1, &d_src, &w, &h, &d_dst);
}

NTX_API ntx_status ntx_hip_normalize_normals_device(ntx_hip_context *ctx, float *d_rgba, int w, int h)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 0 || h < 0 || (size_t(w) * size_t(h) > 0 && !d_rgba)) return NTX_ERR_INVALID_ARGUMENT;
    hipError_t e = timed(ctx, NTX_KERNEL_NORMALS, [&] { return ntx::launch_normals(d_rgba, size_t(w) * size_t(h), true, ctx->stream); });
    if (e != hipSuccess) return fail(ctx, e, "normals launch");
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_normalize_vectors_device(ntx_hip_context *ctx, float *d_rgba, int w, int h)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 0 || h < 0 || (size_t(w) * size_t(h) > 0 && !d_rgba)) return NTX_ERR_INVALID_ARGUMENT;
    hipError_t e = timed(ctx, NTX_KERNEL_NORMALS, [&] { return ntx::launch_normals(d_rgba, size_t(w) * size_t(h), false, ctx->stream); });
    if (e != hipSuccess) return fail(ctx, e, "normals launch");
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_unpack_bgra8_gamma_device(ntx_hip_context *ctx, const uint8_t *d_bgra, int w, int h,
                                                     float gamma, float *d_rgba)
{
    ntx_status s = bind(ctx);
    if (s) return s;
    if (w < 0 || h < 0 || (size_t(w) * size_t(h) > 0 && (!d_bgra || !d_rgba)) || !(gamma > 0.0f))
        return NTX_ERR_INVALID_ARGUMENT;
    const float *lut = nullptr;
    if (!ntx_gamma::is_unit_gamma(gamma)) {
        uint32_t key;
        std::memcpy(&key, &gamma, 4);
        if (!ctx->d_gamma_lut || ctx->gamma_lut_key != key) {
            // Surface::toLinear(gamma) of every 8-bit value v / 255
            // (Surface.cpp:1470-1477, FloatImage.cpp:258-270, 286-297), with
            // the same operations as the reference: the table form for 2.2,
            // the C library's powf otherwise.
            float table[512], h_lut[256];
            for (int k = 0; k < 512; k++) table[k] = ntx_gamma::table_entry(k, 11.0, 5.0);
            for (int k = 0; k < 256; k++) {
                const float v = float(k) / 255.0f;
                h_lut[k] = gamma == 2.2f ? ntx_gamma::pow_11_5(v, table) : ::powf(ntx_gamma::max0(v), gamma);
            }
            if (!ctx->d_gamma_lut) NTX_HIP(ctx, hipMalloc(&ctx->d_gamma_lut, sizeof(h_lut)));
            NTX_HIP(ctx, hipMemcpyAsync(ctx->d_gamma_lut, h_lut, sizeof(h_lut), hipMemcpyHostToDevice, ctx->stream));
            NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
            ctx->gamma_lut_key = key;
        }
        lut = ctx->d_gamma_lut;
    }
    hipError_t e = timed(ctx, NTX_KERNEL_UNPACK, [&] {
        return ntx::launch_unpack(d_bgra, size_t(w) * size_t(h), d_rgba, lut, ctx->stream);
    });
    if (e != hipSuccess) return fail(ctx, e, "unpack launch");
    return NTX_OK;
}

}  // extern "C"
