// batch_api.cpp -- the batch entry points of include/ntx_hip.h: planning a
// batch of textures into level buffers, a work list and a captured graph
// (the mip stages, the persistent compression launch and the multi-image
// BC1 encoder; batch.hip, bc1.hip), and the autotuner that picks the work
// decomposition by timing it.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "ntx_batch.h"
#include "ntx_context.h"

namespace ntx {
hipError_t launch_batch(const BatchArgs &args, int variant, bool tiers, bool mips, int workgroups, hipStream_t stream);
hipError_t launch_batch_reset(int *p, int n, hipStream_t stream);
hipError_t launch_batch_mips(const BatchArgs &args, hipStream_t stream);
hipError_t batch_resident_workgroups(int device, int variant, bool tiers, bool mips, int *count);
int batch_kernel_variants();
}  // namespace ntx

struct ntx_hip_batch {
    ntx_hip_context *ctx = nullptr;
    std::vector<ntx_texture_desc> tex;
    std::vector<int> first;              // first level index of each texture, then the total
    std::vector<ntx::BatchLevel> levels;
    std::vector<size_t> out_bytes;       // compressed bytes per level
    void *d_images = nullptr;            // levels >= 1
    void *d_out = nullptr;               // compressed blocks of every level
    ntx::BatchLevel *d_levels = nullptr;
    ntx::BatchItem *d_items = nullptr;
    size_t items_cap = 0;
    int *d_sync = nullptr;               // ready counters, queue head, status
    int ncounters = 0;
    int nitems = 0;
    // The work list.  Staged schedule: the mip items of stage s at
    // [stage_first[s], stage_first[s + 1]), then the compression items from
    // stage_first.back().  One-launch schedule: stage_first = {0}, every item
    // in the persistent launch.
    std::vector<int> stage_first;
    bool bc1_tiers = false;              // some BC1 level is not at icbc level 8
    int workgroups = 0;
    ntx_batch_tuning tuning{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string signature;
    // BC1 levels at icbc Level8 go to the multi-image two-pass encoder
    // (classify, then search) after the persistent launch rather than to
    // its one-block-per-16-lanes items: the flag per level, the device image
    // table and 256-block chunk map, and the bucket scratch (tallies zero
    // between runs: the search clears them).
    std::vector<char> bc1_multi;
    ntx::Bc1Image *d_bc1_imgs = nullptr;
    void *d_bc1_map = nullptr;
    int *d_bc1_bucket = nullptr;
    int bc1_imgs = 0, bc1_chunks = 0, bc1_cap = 0;
};

namespace {

using ntx::BatchItem;
using ntx::BatchLevel;

constexpr int kSpinLimit = 1 << 22;  // polls (each a sleep plus an atomic read): seconds, far beyond any real wait


int count_mipmaps(int w, int h)
{
    // nv::countMipmaps (Surface.cpp:181-193).
    int n = 0;
    while (w != 1 || h != 1) {
        w = std::max(1, w / 2);
        h = std::max(1, h / 2);
        n++;
    }
    return n + 1;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

ntx_batch_tuning default_tuning()
{
    ntx_batch_tuning t;
    t.bc1_blocks_per_item = 64;
    t.channel_jobs_per_item = 1024;
    t.mip_tiles_per_item = 2;
    t.mip_tile_width = 64;
    t.kernel_variant = 1;
    t.workgroups = 0;
    t.schedule = 0;
    return t;
}

bool valid_tuning(const ntx_batch_tuning &t)
{
    const int tw = t.mip_tile_width;
    return t.bc1_blocks_per_item >= 16 && t.bc1_blocks_per_item % 16 == 0 && t.channel_jobs_per_item >= 256 &&
           t.channel_jobs_per_item % 256 == 0 && t.mip_tiles_per_item >= 1 &&
           (tw == 16 || tw == 32 || tw == 64 || tw == 128 || tw == 256) && t.kernel_variant >= 0 &&
           t.kernel_variant < ntx::batch_kernel_variants() && t.workgroups >= 0 && (t.schedule == 0 || t.schedule == 1);
}

int log2i(int v)
{
    int s = 0;
    while ((1 << s) < v) s++;
    return s;
}

void release_graph(ntx_hip_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// The work list for the current tuning.  Staged schedule: for each stage s
// the mip items that make level s + 1 of every texture (one launch per
// stage), then the compression items of every level (one launch).  One-launch
// schedule: stage s holds the mip items that make level s + 1 of every
// texture, then the compression items of level s, so every item depends only
// on items earlier in the list.  Compression items go stage by stage with BC1
// before channel jobs; each group in decreasing texture size.
ntx_status build_items(ntx_hip_batch *b)
{
    ntx_hip_context *ctx = b->ctx;
    const ntx_batch_tuning &tu = b->tuning;
    const bool one_launch = tu.schedule == 1;
    const int tw = tu.mip_tile_width, th = 256 / tu.mip_tile_width;
    const int n = int(b->tex.size());
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return size_t(b->tex[x].width) * b->tex[x].height > size_t(b->tex[y].width) * b->tex[y].height;
    });
    int stages = 0;
    for (int t = 0; t < n; t++) stages = std::max(stages, b->first[t + 1] - b->first[t]);
    std::vector<BatchItem> items;
    for (auto &lv : b->levels) lv.target = 0;
    auto mips = [&](int s) {
        for (int t : order) {
            if (s + 1 >= b->first[t + 1] - b->first[t]) continue;
            const int li = b->first[t] + s;
            const BatchLevel &dst = b->levels[li + 1];
            const int tiles = ((dst.w + tw - 1) / tw) * ((dst.h + th - 1) / th);
            for (int k = 0; k < tiles; k += tu.mip_tiles_per_item) {
                items.push_back({ ntx::kItemMip, li, k, std::min(tu.mip_tiles_per_item, tiles - k) });
                b->levels[li + 1].target++;
            }
        }
    };
    auto compression = [&](int s) {
        for (int pass = 0; pass < 2; pass++) {
            for (int t : order) {
                if (s >= b->first[t + 1] - b->first[t]) continue;
                const int li = b->first[t] + s;
                const BatchLevel &lv = b->levels[li];
                const int blocks = ((lv.w + 3) / 4) * ((lv.h + 3) / 4);
                if (pass == 0 && lv.format == NTX_FORMAT_BC1 && !b->bc1_multi[size_t(li)]) {
                    for (int k = 0; k < blocks; k += tu.bc1_blocks_per_item)
                        items.push_back({ ntx::kItemBc1, li, k, std::min(tu.bc1_blocks_per_item, blocks - k) });
                } else if (pass == 1 && lv.format != NTX_FORMAT_BC1) {
                    const int jobs = blocks * (lv.format == NTX_FORMAT_BC5 ? 2 : 1);
                    for (int k = 0; k < jobs; k += tu.channel_jobs_per_item)
                        items.push_back({ ntx::kItemChannel, li, k, std::min(tu.channel_jobs_per_item, jobs - k) });
                }
            }
        }
    };
    b->stage_first.clear();
    if (one_launch) {
        b->stage_first.push_back(0);
        for (int s = 0; s < stages; s++) {
            mips(s);
            compression(s);
        }
    } else {
        for (int s = 0; s < stages; s++) {
            b->stage_first.push_back(int(items.size()));
            mips(s);
        }
        b->stage_first.push_back(int(items.size()));
        for (int s = 0; s < stages; s++) compression(s);
    }
    b->nitems = int(items.size());
    if (items.size() > b->items_cap) {
        if (b->d_items) NTX_HIP(ctx, hipFree(b->d_items));
        b->d_items = nullptr;
        b->items_cap = 0;
        NTX_HIP(ctx, hipMalloc(&b->d_items, sizeof(BatchItem) * std::max<size_t>(items.size(), 1)));
        b->items_cap = items.size();
    }
    if (!items.empty())
        NTX_HIP(ctx, hipMemcpy(b->d_items, items.data(), sizeof(BatchItem) * items.size(), hipMemcpyHostToDevice));
    NTX_HIP(ctx, hipMemcpy(b->d_levels, b->levels.data(), sizeof(BatchLevel) * b->levels.size(), hipMemcpyHostToDevice));
    int resident = 0;
    NTX_HIP(ctx, ntx::batch_resident_workgroups(ctx->device, tu.kernel_variant, b->bc1_tiers, one_launch, &resident));
    int wgs = tu.workgroups > 0 ? std::min(tu.workgroups, resident) : resident;
    wgs = std::max(1, std::min(wgs, std::max(b->nitems - b->stage_first.back(), 1)));
    b->workgroups = wgs;
    return NTX_OK;
}

ntx::BatchArgs make_args(const ntx_hip_batch *b)
{
    ntx::BatchArgs a;
    a.items = b->d_items;
    a.levels = b->d_levels;
    a.tab = b->ctx->d_tab;
    a.counters = b->d_sync;
    a.head = b->d_sync + b->ncounters;
    a.status = b->d_sync + b->ncounters + 1;
    a.nitems = b->nitems;
    a.tile_shift = log2i(b->tuning.mip_tile_width);
    a.spin_limit = kSpinLimit;
    return a;
}

// The launches of one run, enqueued on the context's stream: the counter
// reset, the mip stages (staged schedule), the persistent launch, the
// multi-image BC1 passes.
hipError_t enqueue_run(ntx_hip_batch *b)
{
    hipStream_t s = b->ctx->stream;
    hipError_t e = ntx::launch_batch_reset(b->d_sync, b->ncounters + 2, s);
    const ntx::BatchArgs all = make_args(b);
    const int nstages = int(b->stage_first.size()) - 1;
    for (int k = 0; e == hipSuccess && k < nstages; k++) {
        ntx::BatchArgs a = all;
        a.items = b->d_items + b->stage_first[k];
        a.nitems = b->stage_first[k + 1] - b->stage_first[k];
        e = ntx::launch_batch_mips(a, s);
    }
    const int c0 = b->stage_first.back();
    if (e == hipSuccess && b->nitems > c0) {
        ntx::BatchArgs a = all;
        a.items = b->d_items + c0;
        a.nitems = b->nitems - c0;
        e = ntx::launch_batch(a, b->tuning.kernel_variant, b->bc1_tiers, b->tuning.schedule == 1, b->workgroups, s);
    }
    if (e == hipSuccess && b->bc1_imgs > 0) {
        int *counts = b->d_bc1_bucket, *lists = counts + ntx::bc3_lists_offset();
        e = ntx::launch_bc1_classify_multi(b->d_bc1_imgs, b->d_bc1_map, b->bc1_chunks, b->ctx->d_tab, counts, lists,
                                           b->bc1_cap, s);
        if (e == hipSuccess)
            e = ntx::launch_bc1_search_multi(b->d_bc1_imgs, b->bc1_imgs, b->ctx->d_tab, counts, lists, b->bc1_cap,
                                             ntx::kBc1SearchGroupsPerCU * b->ctx->cus, s);
    }
    return e;
}

ntx_status build_graph(ntx_hip_batch *b)
{
    ntx_hip_context *ctx = b->ctx;
    release_graph(b);
    NTX_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_run(b);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return ntx::fail(ctx, e, "batch capture");
    }
    if (e2 != hipSuccess) return ntx::fail(ctx, e2, "hipStreamEndCapture");
    b->graph = g;
    NTX_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return NTX_OK;
}

ntx_status apply_tuning(ntx_hip_batch *b, const ntx_batch_tuning &t)
{
    if (!valid_tuning(t)) return NTX_ERR_INVALID_ARGUMENT;
    ntx_hip_context *ctx = b->ctx;
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b->tuning = t;
    ntx_status s = build_items(b);
    if (s) return s;
    return build_graph(b);
}

std::string signature_of(const std::vector<ntx_texture_desc> &tex)
{
    std::string sig;
    char buf[160];
    for (const auto &d : tex) {
        uint32_t fw, p0, p1;
        std::memcpy(&fw, &d.filter_width, 4);
        std::memcpy(&p0, &d.filter_params[0], 4);
        std::memcpy(&p1, &d.filter_params[1], 4);
        std::snprintf(buf, sizeof(buf), "%d,%d,%d,%d,%d,%d,%d,%08x,%08x,%08x,%d;", d.width, d.height, d.format, d.quality,
                      d.alpha_mode, d.mip_levels, d.filter, fw, p0, p1, d.wrap);
        sig += buf;
    }
    return sig;
}

ntx_status time_runs(ntx_hip_batch *b, int reps, float *ms)
{
    ntx_hip_context *ctx = b->ctx;
    NTX_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));  // warm-up
    float best = 0.0f;
    for (int r = 0; r < reps; r++) {
        NTX_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        NTX_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
        NTX_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        NTX_HIP(ctx, hipEventSynchronize(ctx->ev1));
        float t = 0.0f;
        NTX_HIP(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
        if (r == 0 || t < best) best = t;
    }
    int status = 0;
    NTX_HIP(ctx, hipMemcpy(&status, b->d_sync + b->ncounters + 1, sizeof(int), hipMemcpyDeviceToHost));
    if (status) {
        ctx->err = "batch: a dependency wait did not complete";
        return NTX_ERR_DEVICE;
    }
    *ms = best;
    return NTX_OK;
}

}  // namespace

extern "C" {

NTX_API ntx_status ntx_hip_batch_create(ntx_hip_context *ctx, const ntx_texture_desc *descs, int n, ntx_hip_batch **out)
{
    if (!out) return NTX_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    if (n < 1 || !descs) return NTX_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < n; i++) {
        const ntx_texture_desc &d = descs[i];
        if (!d.d_rgba || d.width < 1 || d.height < 1) return NTX_ERR_INVALID_ARGUMENT;
        if (d.filter < NTX_FILTER_BOX || d.filter > NTX_FILTER_KAISER || d.wrap < NTX_WRAP_CLAMP || d.wrap > NTX_WRAP_MIRROR)
            return NTX_ERR_INVALID_ARGUMENT;
        if (d.mip_levels < 0 || d.mip_levels > count_mipmaps(d.width, d.height)) return NTX_ERR_INVALID_ARGUMENT;
        if (d.format != NTX_FORMAT_BC1 && d.format != NTX_FORMAT_BC4 && d.format != NTX_FORMAT_BC5) return NTX_ERR_UNSUPPORTED;
        if (!ntx::supported(d.format, d.quality)) return NTX_ERR_UNSUPPORTED;
        // The batch kernel's channel encoder is QuickCompress::compressDXT5A;
        // BC4 / BC5 at Production and Highest (OptimalCompress) go through
        // the per-image and chain entry points.
        if (d.format != NTX_FORMAT_BC1 && d.quality >= NTX_QUALITY_PRODUCTION) return NTX_ERR_UNSUPPORTED;
    }
    ntx_hip_batch *b = new ntx_hip_batch;
    b->ctx = ctx;
    b->tex.assign(descs, descs + n);
    auto cleanup = [&](ntx_status st) { ntx_hip_batch_destroy(b); return st; };

    // Levels, level buffers and output offsets.
    size_t img_bytes = 0, out_total = 0;
    std::vector<size_t> img_off, out_off;
    b->first.push_back(0);
    for (int t = 0; t < n; t++) {
        const ntx_texture_desc &d = b->tex[t];
        const int L = d.mip_levels ? d.mip_levels : count_mipmaps(d.width, d.height);
        int w = d.width, h = d.height;
        for (int l = 0; l < L; l++) {
            BatchLevel lv{};
            lv.w = w;
            lv.h = h;
            lv.format = d.format;
            lv.alpha_mode = d.alpha_mode;
            lv.bc1_level = ntx::bc1_level(d.quality);
            lv.ready = l == 0 ? -1 : b->ncounters++;
            if (d.format == NTX_FORMAT_BC1 && lv.bc1_level != 8) b->bc1_tiers = true;
            const bool zero_w = d.color_weights[0] == 0 && d.color_weights[1] == 0 && d.color_weights[2] == 0 &&
                                d.color_weights[3] == 0;
            for (int k = 0; k < 3; k++) lv.cw[k] = zero_w ? 1.0f : d.color_weights[k];
            lv.next = l + 1 < L ? int(b->levels.size()) + 1 : -1;
            img_off.push_back(img_bytes);
            if (l > 0) img_bytes += align256(size_t(w) * size_t(h) * 4 * sizeof(float));
            const size_t ob = ntx_hip_compressed_size(d.format, w, h);
            out_off.push_back(out_total);
            b->out_bytes.push_back(ob);
            out_total += align256(ob);
            b->levels.push_back(lv);
            w = std::max(1, w / 2);
            h = std::max(1, h / 2);
        }
        b->first.push_back(int(b->levels.size()));
    }
    hipError_t e = hipMalloc(&b->d_images, std::max<size_t>(img_bytes, 256));
    if (e == hipSuccess) e = hipMalloc(&b->d_out, std::max<size_t>(out_total, 256));
    if (e == hipSuccess) e = hipMalloc(&b->d_levels, sizeof(BatchLevel) * b->levels.size());
    if (e == hipSuccess) e = hipMalloc(&b->d_sync, sizeof(int) * size_t(b->ncounters + 2));
    if (e != hipSuccess) return cleanup(ntx::fail(ctx, e, "batch buffers"));
    for (int t = 0; t < n; t++) {
        const ntx_texture_desc &d = b->tex[t];
        float width = d.filter_width;
        float params[2] = { d.filter_params[0], d.filter_params[1] };
        if (!(width > 0.0f)) {
            // Surface.cpp:1121-1141 getDefaultFilterWidthAndParams.
            width = d.filter == NTX_FILTER_BOX ? 0.5f : d.filter == NTX_FILTER_TRIANGLE ? 1.0f : 3.0f;
            params[0] = 4.0f;
            params[1] = 1.0f;
        }
        const bool fast = d.filter == NTX_FILTER_BOX && width == 0.5f && d.alpha_mode != NTX_ALPHA_TRANSPARENCY;
        for (int li = b->first[t]; li < b->first[t + 1]; li++) {
            BatchLevel &lv = b->levels[li];
            lv.data = li == b->first[t] ? const_cast<float *>(d.d_rgba)
                                        : reinterpret_cast<float *>(static_cast<char *>(b->d_images) + img_off[li]);
            lv.out = static_cast<uint8_t *>(b->d_out) + out_off[li];
            lv.wrap = d.wrap;
            if (lv.next < 0) continue;
            const int w = lv.w, h = lv.h, nw = std::max(1, w / 2), nh = std::max(1, h / 2);
            if (fast) {
                lv.mip_kind = 0;
            } else {
                ntx::PolyDev *kx = nullptr, *ky = nullptr;
                if ((s = ntx::poly_tables(ctx, d.filter, width, params, w, nw, &kx))) return cleanup(s);
                if ((s = ntx::poly_tables(ctx, d.filter, width, params, h, nh, &ky))) return cleanup(s);
                lv.mip_kind = 1;
                lv.kx = kx->weights;
                lv.lx = kx->left;
                lv.wx = kx->window;
                lv.ky = ky->weights;
                lv.ly = ky->left;
                lv.wy = ky->window;
            }
        }
    }
    // The Level8 BC1 levels for the multi-image encoder (within its image
    // and block limits; any others stay persistent-kernel items).
    {
        std::vector<ntx::Bc1Image> imgs;
        std::vector<int> map;
        long long total = 0;
        b->bc1_multi.assign(b->levels.size(), 0);
        for (size_t li = 0; li < b->levels.size(); li++) {
            const BatchLevel &lv = b->levels[li];
            if (lv.format != NTX_FORMAT_BC1 || lv.bc1_level != 8) continue;
            const int blocks = ((lv.w + 3) / 4) * ((lv.h + 3) / 4);
            if (int(imgs.size()) >= ntx::bc1_multi_max_images() || blocks > ntx::bc1_multi_max_blocks() ||
                total + blocks >= (1ll << 28))
                continue;
            b->bc1_multi[li] = 1;
            ntx::Bc1Image im{};
            im.src = lv.data;
            im.out = lv.out;
            im.w = lv.w;
            im.h = lv.h;
            im.alpha_mode = lv.alpha_mode;
            im.rows16 = (reinterpret_cast<uintptr_t>(lv.data) & 15) == 0 && (lv.w & 3) == 0;
            for (int k = 0; k < 3; k++) im.cw[k] = lv.cw[k];
            for (int k = 0; k < blocks; k += 256) {
                map.push_back(int(imgs.size()));
                map.push_back(k);
            }
            imgs.push_back(im);
            total += blocks;
        }
        if (!imgs.empty()) {
            b->bc1_imgs = int(imgs.size());
            b->bc1_chunks = int(map.size() / 2);
            b->bc1_cap = int(total);
            const size_t bucket = sizeof(int) * size_t(ntx::bc1_scratch_ints(b->bc1_cap));
            e = hipMalloc(&b->d_bc1_imgs, sizeof(ntx::Bc1Image) * imgs.size());
            if (e == hipSuccess) e = hipMalloc(&b->d_bc1_map, sizeof(int) * map.size());
            if (e == hipSuccess) e = hipMalloc(&b->d_bc1_bucket, bucket);
            if (e == hipSuccess)
                e = hipMemcpy(b->d_bc1_imgs, imgs.data(), sizeof(ntx::Bc1Image) * imgs.size(), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(b->d_bc1_map, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice);
            // Tallies and tickets zero before the first run (the lists need no clearing).
            if (e == hipSuccess) e = hipMemset(b->d_bc1_bucket, 0, sizeof(int) * size_t(ntx::bc3_lists_offset()));
            if (e != hipSuccess) return cleanup(ntx::fail(ctx, e, "batch BC1 tables"));
        }
    }
    b->signature = signature_of(b->tex);
    auto it = ctx->batch_tuning.find(b->signature);
    if ((s = apply_tuning(b, it != ctx->batch_tuning.end() ? it->second : default_tuning()))) return cleanup(s);
    *out = b;
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_batch_destroy(ntx_hip_batch *b)
{
    if (!b) return NTX_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    release_graph(b);
    for (void *p : { b->d_images, b->d_out, static_cast<void *>(b->d_levels), static_cast<void *>(b->d_items),
                     static_cast<void *>(b->d_sync), static_cast<void *>(b->d_bc1_imgs), b->d_bc1_map,
                     static_cast<void *>(b->d_bc1_bucket) })
        if (p) (void)hipFree(p);
    delete b;
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_batch_submit(ntx_hip_batch *b)
{
    if (!b) return NTX_ERR_INVALID_ARGUMENT;
    ntx_hip_context *ctx = b->ctx;
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    if (!b->exec && (s = build_graph(b))) return s;
    hipError_t e = ntx::timed(ctx, NTX_KERNEL_BATCH, [&] { return hipGraphLaunch(b->exec, ctx->stream); });
    if (e != hipSuccess) return ntx::fail(ctx, e, "batch graph launch");
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_batch_wait(ntx_hip_batch *b)
{
    if (!b) return NTX_ERR_INVALID_ARGUMENT;
    ntx_hip_context *ctx = b->ctx;
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int status = 0;
    NTX_HIP(ctx, hipMemcpy(&status, b->d_sync + b->ncounters + 1, sizeof(int), hipMemcpyDeviceToHost));
    if (status) {
        ctx->err = "batch: a dependency wait did not complete";
        return NTX_ERR_DEVICE;
    }
    return NTX_OK;
}

NTX_API int ntx_hip_batch_levels(const ntx_hip_batch *b, int tex)
{
    if (!b || tex < 0 || tex >= int(b->tex.size())) return 0;
    return b->first[tex + 1] - b->first[tex];
}

NTX_API ntx_status ntx_hip_batch_level(const ntx_hip_batch *b, int tex, int level, const float **d_rgba, int *w, int *h,
                                       const void **d_blocks, size_t *bytes)
{
    if (!b || tex < 0 || tex >= int(b->tex.size()) || level < 0 || level >= ntx_hip_batch_levels(b, tex))
        return NTX_ERR_INVALID_ARGUMENT;
    const int li = b->first[tex] + level;
    const BatchLevel &lv = b->levels[li];
    if (d_rgba) *d_rgba = lv.data;
    if (w) *w = lv.w;
    if (h) *h = lv.h;
    if (d_blocks) *d_blocks = lv.out;
    if (bytes) *bytes = b->out_bytes[li];
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_batch_get_tuning(const ntx_hip_batch *b, ntx_batch_tuning *t)
{
    if (!b || !t) return NTX_ERR_INVALID_ARGUMENT;
    *t = b->tuning;
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_batch_set_tuning(ntx_hip_batch *b, const ntx_batch_tuning *t)
{
    if (!b || !t) return NTX_ERR_INVALID_ARGUMENT;
    ntx_status s = ntx::bind(b->ctx);
    if (s) return s;
    return apply_tuning(b, *t);
}

NTX_API ntx_status ntx_hip_batch_autotune(ntx_hip_batch *b, int reps, ntx_batch_tuning *best_out, float *best_ms_out)
{
    if (!b) return NTX_ERR_INVALID_ARGUMENT;
    ntx_hip_context *ctx = b->ctx;
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    reps = std::max(1, reps);
    ntx_batch_tuning best = b->tuning;
    if ((s = apply_tuning(b, best))) return s;
    float best_ms = 0.0f;
    if ((s = time_runs(b, reps, &best_ms))) return s;
    int resident = 0;
    // Coordinate descent: each field in turn over its candidates, others
    // fixed; the schedule first, the decomposition under it after.
    for (int field : { 6, 0, 1, 2, 3, 4, 5 }) {
        std::vector<int> cand;
        switch (field) {
        case 6: cand = { 0, 1 }; break;                             // schedule
        case 0: cand = { 0, 1, 2 }; break;                          // kernel_variant
        case 1: cand = { 16, 32, 64, 128, 256 }; break;             // mip_tile_width
        case 2: cand = { 1, 2, 4, 8 }; break;                       // mip_tiles_per_item
        case 3: cand = { 16, 32, 64, 128, 256 }; break;             // bc1_blocks_per_item
        case 4: cand = { 256, 512, 1024, 2048, 4096 }; break;       // channel_jobs_per_item
        default:                                                    // workgroups
            NTX_HIP(ctx, ntx::batch_resident_workgroups(ctx->device, best.kernel_variant, b->bc1_tiers,
                                                        best.schedule == 1, &resident));
            cand = { 0, resident * 3 / 4, resident / 2, resident / 4 };
            break;
        }
        for (int v : cand) {
            ntx_batch_tuning t = best;
            int *slot[] = { &t.kernel_variant, &t.mip_tile_width, &t.mip_tiles_per_item, &t.bc1_blocks_per_item,
                            &t.channel_jobs_per_item, &t.workgroups, &t.schedule };
            if (*slot[field] == v) continue;
            *slot[field] = v;
            if (field == 5 && v <= 0 && best.workgroups == 0) continue;
            if ((s = apply_tuning(b, t))) return s;
            float ms = 0.0f;
            if ((s = time_runs(b, reps, &ms))) return s;
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
    }
    if ((s = apply_tuning(b, best))) return s;
    ctx->batch_tuning[b->signature] = best;
    if (best_out) *best_out = best;
    if (best_ms_out) *best_ms_out = best_ms;
    return NTX_OK;
}

}  // extern "C"
