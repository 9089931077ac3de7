// noalloc.hip -- ntx_hip_compress_device_noalloc: ntx_hip_compress_device for
// callers that cannot allocate device memory (code under graph capture,
// serving loops, hosts that forbid hipMalloc).  The two-pass encoders' scratch
// (per-count tallies, block lists and, for BC3, the packed texels) lives in a
// file-scope device array instead of the context's allocation; the entry
// point calls no allocating routine.
//
// Each GPU holds its own copy of the array from module load until the process
// ends; every call on that GPU shares it (calls that overlap on one GPU race).
// The tallies are cleared by a fill kernel at the start of every call.
#include "bucket.hpp"
#include "ntx_context.h"

#include <mutex>
#include <string>

namespace {

// Largest image, in 4x4 blocks, whose scratch fits: 1024 x 1024 texels.
constexpr int kNoAllocMaxBlocks = 1 << 16;
// The BC3 layout (the larger of the two): tallies, 17 lists, then 16 words
// of packed texels per block from a 16-byte boundary.
constexpr size_t kScratchInts =
    ((size_t(ntx_bucket::kListsOffset) + size_t(ntx_bucket::kCountSlots) * kNoAllocMaxBlocks + 3) & ~size_t(3)) +
    size_t(16) * kNoAllocMaxBlocks;

static __device__ __align__(256) int scratch[kScratchInts];

// The array's device address per device ordinal, from hipGetSymbolAddress.
constexpr int kMaxDevices = 64;
int *g_scratch[kMaxDevices] = {};
std::mutex g_mutex;

}  // namespace

namespace ntx {

// Called by ntx_hip_create with the context's device current: looks up this
// device's copy of the scratch array once per device.
ntx_status noalloc_init(ntx_hip_context *ctx)
{
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return NTX_OK;  // the entry point reports it
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch[ctx->device]) return NTX_OK;
    int *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void **>(&p), HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail(ctx, e, "no-allocation scratch symbol");
    if (size_t(bc3_scratch_ints(kNoAllocMaxBlocks)) > kScratchInts ||
        size_t(bc1_scratch_ints(kNoAllocMaxBlocks)) > kScratchInts) {
        ctx->err = "no-allocation scratch smaller than the encoders' layout";
        return NTX_ERR_DEVICE;
    }
    g_scratch[ctx->device] = p;
    return NTX_OK;
}

}  // namespace ntx

extern "C" {

NTX_API ntx_status ntx_hip_compress_device_noalloc(ntx_hip_context *ctx, int format, int quality, int alpha_mode,
                                                   const float *cw, const float *d_rgba, int w, int h, void *d_out)
{
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    if (w < 0 || h < 0 || (size_t(w) * size_t(h) > 0 && (!d_rgba || !d_out))) return NTX_ERR_INVALID_ARGUMENT;
    if (!ntx::supported(format, quality)) return NTX_ERR_UNSUPPORTED;
    if (size_t(w) * size_t(h) == 0) return NTX_OK;
    const float ones[4] = { 1.0f, 1.0f, 1.0f, 1.0f };
    if (!cw) cw = ones;
    int *bucket = nullptr;
    if (ntx::scratch_ints(ctx, format, quality, w, h)) {
        const size_t nblocks = size_t((w + 3) / 4) * size_t((h + 3) / 4);
        if (nblocks <= size_t(kNoAllocMaxBlocks)) {
            if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch[ctx->device]) {
                ctx->err = "no-allocation scratch not set up for this device";
                return NTX_ERR_DEVICE;
            }
            bucket = g_scratch[ctx->device];
        } else if (format != NTX_FORMAT_BC1) {
            // BC1 falls back to its one-pass encoder, which needs no scratch.
            ctx->err = "image of " + std::to_string(nblocks) + " blocks exceeds the no-allocation limit of " +
                       std::to_string(kNoAllocMaxBlocks) + " blocks (1024 x 1024 texels)";
            return NTX_ERR_INVALID_ARGUMENT;
        }
    }
    const hipError_t e = ntx::compress_scratch(ctx, format, quality, alpha_mode, cw, d_rgba, w, h, d_out, bucket);
    if (e != hipSuccess) return ntx::fail(ctx, e, "compress launch");
    return NTX_OK;
}

}  // extern "C"
