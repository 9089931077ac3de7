// mlp_api.cpp -- the learned BC1 endpoint entry points of include/ntx_hip.h:
// the network object (parameters and Adam state in device memory), feature
// and label preparation, training steps and the proposal compressor with its
// cluster-fit fallback (mlp.hip).
#include <cmath>
#include <cstring>
#include <vector>

#include "ntx_context.h"
#include "ntx_mlp.h"

namespace ntx {
hipError_t launch_mlp_features(const float *d_src, int w, int h, void *d_feat, float *d_aux, hipStream_t stream);
hipError_t launch_mlp_labels(const float *d_src, int w, int h, const void *d_bc1, const float *d_aux, float *d_labels,
                             hipStream_t stream);
hipError_t launch_mlp_forward(const float *d_params, const void *d_feat, int n, float *d_out, hipStream_t stream);
int mlp_train_rows(int n, int tiles_per_wave);
hipError_t launch_mlp_train(const float *d_params, const void *d_feat, const float *d_labels, int n, int tiles_per_wave,
                            float *d_partials, float *d_loss, hipStream_t stream);
hipError_t launch_mlp_adam(float *d_params, float *d_m1, float *d_m2, const float *d_partials, int rows, int n, float lr,
                           int step, float *d_grad, hipStream_t stream);
hipError_t launch_bc1_mlp(const float *d_params, const float *d_src, int w, int h, const Tables *d_tab, float tau_pred,
                          float tau_gap, int refine_iters, int greedy_rounds, int box_candidates, void *d_out,
                          int *d_fb_count, int *d_fb_list, float *d_err, int cus, hipStream_t stream);
}  // namespace ntx

static_assert(ntx_mlp::kParams == NTX_MLP_PARAMS, "parameter count");

struct ntx_hip_mlp {
    ntx_hip_context *ctx = nullptr;
    float *d_params = nullptr, *d_m1 = nullptr, *d_m2 = nullptr;
    int step = 0;
    ntx::DevBuf partials, loss, grad, fb;
};

namespace {

constexpr int kTilesPerWave = 8;  // training: 128 blocks per wave

// He-uniform hidden weights from a xorshift stream; zero biases and a zero
// output layer, so that the untrained network proposes the bounding-box
// diagonals unchanged.
void init_params(uint32_t seed, std::vector<float> &p)
{
    using namespace ntx_mlp;
    p.assign(kParams, 0.0f);
    uint64_t s = 0x9e3779b97f4a7c15ull ^ (uint64_t(seed) * 0xbf58476d1ce4e5b9ull);
    auto uni = [&]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return float((s >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
    };
    auto fill = [&](int off, int rows, int cols) {
        const float lim = std::sqrt(6.0f / float(cols));
        for (int i = 0; i < rows * cols; i++) p[off + i] = lim * uni();
    };
    fill(kW1, kHidden, kIn);
    fill(kW2, kHidden, kHidden);
}

ntx_status upload_params(ntx_hip_mlp *m, const float *params)
{
    ntx_hip_context *ctx = m->ctx;
    const size_t bytes = sizeof(float) * ntx_mlp::kParams;
    NTX_HIP(ctx, hipMemcpy(m->d_params, params, bytes, hipMemcpyHostToDevice));
    NTX_HIP(ctx, hipMemset(m->d_m1, 0, bytes));
    NTX_HIP(ctx, hipMemset(m->d_m2, 0, bytes));
    m->step = 0;
    return NTX_OK;
}

int resident_cus(ntx_hip_context *ctx)
{
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus <= 0) cus = 256;
    return cus;
}

}  // namespace

extern "C" {

NTX_API ntx_status ntx_hip_mlp_create(ntx_hip_context *ctx, const float *params, uint32_t seed, ntx_hip_mlp **out)
{
    if (!out) return NTX_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    ntx_hip_mlp *m = new ntx_hip_mlp;
    m->ctx = ctx;
    const size_t bytes = sizeof(float) * ntx_mlp::kParams;
    hipError_t e = hipMalloc(&m->d_params, bytes);
    if (e == hipSuccess) e = hipMalloc(&m->d_m1, bytes);
    if (e == hipSuccess) e = hipMalloc(&m->d_m2, bytes);
    if (e != hipSuccess) {
        ntx_hip_mlp_destroy(m);
        return ntx::fail(ctx, e, "network buffers");
    }
    std::vector<float> init;
    if (!params) {
        init_params(seed, init);
        params = init.data();
    }
    if ((s = upload_params(m, params))) {
        ntx_hip_mlp_destroy(m);
        return s;
    }
    *out = m;
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_mlp_destroy(ntx_hip_mlp *m)
{
    if (!m) return NTX_OK;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    for (void *p : { static_cast<void *>(m->d_params), static_cast<void *>(m->d_m1), static_cast<void *>(m->d_m2) })
        if (p) (void)hipFree(p);
    for (ntx::DevBuf *b : { &m->partials, &m->loss, &m->grad, &m->fb }) ntx::release(m->ctx, *b);
    delete m;
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_mlp_get_params(ntx_hip_mlp *m, float *params)
{
    if (!m || !params) return NTX_ERR_INVALID_ARGUMENT;
    ntx_status s = ntx::bind(m->ctx);
    if (s) return s;
    NTX_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
    NTX_HIP(m->ctx, hipMemcpy(params, m->d_params, sizeof(float) * ntx_mlp::kParams, hipMemcpyDeviceToHost));
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_mlp_set_params(ntx_hip_mlp *m, const float *params)
{
    if (!m || !params) return NTX_ERR_INVALID_ARGUMENT;
    ntx_status s = ntx::bind(m->ctx);
    if (s) return s;
    NTX_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
    return upload_params(m, params);
}

NTX_API ntx_status ntx_hip_mlp_features_device(ntx_hip_context *ctx, const float *d_rgba, int w, int h, void *d_features,
                                               float *d_aux)
{
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    if (w < 1 || h < 1 || !d_rgba || !d_features || !d_aux) return NTX_ERR_INVALID_ARGUMENT;
    hipError_t e = ntx::launch_mlp_features(d_rgba, w, h, d_features, d_aux, ctx->stream);
    if (e != hipSuccess) return ntx::fail(ctx, e, "features launch");
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_mlp_labels_device(ntx_hip_context *ctx, const float *d_rgba, int w, int h, const void *d_bc1,
                                             const float *d_aux, float *d_labels)
{
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    if (w < 1 || h < 1 || !d_rgba || !d_bc1 || !d_aux || !d_labels) return NTX_ERR_INVALID_ARGUMENT;
    hipError_t e = ntx::launch_mlp_labels(d_rgba, w, h, d_bc1, d_aux, d_labels, ctx->stream);
    if (e != hipSuccess) return ntx::fail(ctx, e, "labels launch");
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_mlp_forward_device(ntx_hip_mlp *m, const void *d_features, int n, float *d_out)
{
    if (!m || n < 0 || (n > 0 && (!d_features || !d_out))) return NTX_ERR_INVALID_ARGUMENT;
    ntx_hip_context *ctx = m->ctx;
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    hipError_t e = ntx::launch_mlp_forward(m->d_params, d_features, n, d_out, ctx->stream);
    if (e != hipSuccess) return ntx::fail(ctx, e, "forward launch");
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_mlp_train_step(ntx_hip_mlp *m, const void *d_features, const float *d_labels, int n, float lr,
                                          float *loss, float *grad)
{
    if (!m || n < 1 || !d_features || !d_labels || !(lr >= 0.0f)) return NTX_ERR_INVALID_ARGUMENT;
    ntx_hip_context *ctx = m->ctx;
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    const int rows = ntx::mlp_train_rows(n, kTilesPerWave);
    if ((s = ntx::ensure(ctx, m->partials, sizeof(float) * size_t(rows) * ntx_mlp::kParams))) return s;
    if ((s = ntx::ensure(ctx, m->loss, sizeof(float) * size_t(rows)))) return s;
    if (grad && (s = ntx::ensure(ctx, m->grad, sizeof(float) * ntx_mlp::kParams))) return s;
    float *d_part = static_cast<float *>(m->partials.ptr), *d_loss = static_cast<float *>(m->loss.ptr);
    if (lr > 0.0f) m->step++;
    hipError_t e = ntx::timed(ctx, NTX_KERNEL_MLP_TRAIN, [&] {
        hipError_t r = ntx::launch_mlp_train(m->d_params, d_features, d_labels, n, kTilesPerWave, d_part, d_loss, ctx->stream);
        if (r == hipSuccess)
            r = ntx::launch_mlp_adam(m->d_params, m->d_m1, m->d_m2, d_part, rows, n, lr, m->step > 0 ? m->step : 1,
                                     grad ? static_cast<float *>(m->grad.ptr) : nullptr, ctx->stream);
        return r;
    });
    if (e != hipSuccess) return ntx::fail(ctx, e, "training launch");
    if (loss) {
        std::vector<float> l(rows);
        NTX_HIP(ctx, hipMemcpyAsync(l.data(), d_loss, sizeof(float) * rows, hipMemcpyDeviceToHost, ctx->stream));
        NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double t = 0.0;
        for (float v : l) t += v;
        *loss = float(t / n);
    }
    if (grad) {
        NTX_HIP(ctx, hipMemcpyAsync(grad, m->grad.ptr, sizeof(float) * ntx_mlp::kParams, hipMemcpyDeviceToHost, ctx->stream));
        NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return NTX_OK;
}

}  // extern "C"

namespace {

// One image through the learned form on `stream`: the network's proposals,
// then the blocks it sends to the fallback through the cluster fit.  d_count
// (one int, then padding to 64) and d_list (the image's block count) are the
// fallback count and list.
ntx_status mlp_compress_on(ntx_hip_mlp *m, hipStream_t stream, int *d_count, int *d_list, const float *d_rgba, int w,
                           int h, float tau_pred, float tau_gap, int refine_iters, int greedy_rounds, int flags,
                           void *d_out)
{
    ntx_hip_context *ctx = m->ctx;
    const int n = ((w + 3) / 4) * ((h + 3) / 4);
    NTX_HIP(ctx, hipMemsetAsync(d_count, 0, sizeof(int), stream));
    const int cus = resident_cus(ctx);
    hipError_t e = ntx::timed_on(ctx, stream, NTX_KERNEL_BC1_MLP, [&] {
        return ntx::launch_bc1_mlp(m->d_params, d_rgba, w, h, ctx->d_tab, tau_pred, tau_gap, refine_iters, greedy_rounds,
                                   (flags & NTX_MLP_BOX_CANDIDATES) ? 1 : 0, d_out, d_count, d_list, nullptr, cus,
                                   stream);
    });
    // The fallback blocks through the cluster fit, one block per 16-lane
    // group over the list (the bytes of every other BC1 Level8 encoder).
    const float ones[3] = { 1.0f, 1.0f, 1.0f };
    if (e == hipSuccess)
        e = ntx::timed_on(ctx, stream, NTX_KERNEL_BC1_LIST, [&] {
            return ntx::launch_bc1_list(d_rgba, w, h, NTX_ALPHA_NONE, ones, ctx->d_tab, d_out, d_list, d_count, n,
                                        8 * ctx->cus, stream);
        });
    if (e != hipSuccess) return ntx::fail(ctx, e, "bc1 proposal launch");
    return NTX_OK;
}

}  // namespace

extern "C" {

NTX_API ntx_status ntx_hip_compress_bc1_mlp_device(ntx_hip_mlp *m, const float *d_rgba, int w, int h, float tau_pred,
                                                   float tau_gap, int refine_iters, int greedy_rounds, int flags,
                                                   void *d_out, int *fallback_blocks)
{
    if (!m || w < 0 || h < 0 || refine_iters < 0 || greedy_rounds < 0) return NTX_ERR_INVALID_ARGUMENT;
    ntx_hip_context *ctx = m->ctx;
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    const int n = ((w + 3) / 4) * ((h + 3) / 4);
    if (fallback_blocks) *fallback_blocks = 0;
    if (n == 0) return NTX_OK;
    if (!d_rgba || !d_out) return NTX_ERR_INVALID_ARGUMENT;
    // fallback count, then the list of up to n block indices
    if ((s = ntx::ensure(ctx, m->fb, sizeof(int) * (size_t(n) + 64)))) return s;
    int *d_count = static_cast<int *>(m->fb.ptr);
    if ((s = mlp_compress_on(m, ctx->stream, d_count, d_count + 64, d_rgba, w, h, tau_pred, tau_gap, refine_iters,
                             greedy_rounds, flags, d_out)))
        return s;
    if (fallback_blocks) {
        NTX_HIP(ctx, hipMemcpyAsync(fallback_blocks, m->fb.ptr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        NTX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return NTX_OK;
}

NTX_API ntx_status ntx_hip_compress_bc1_mlp_images_device(ntx_hip_mlp *m, int n, const float *const *d_rgba,
                                                          const int *w, const int *h,
                                                          const ntx_mlp_settings *settings, void *const *d_out)
{
    if (!m || n < 0) return NTX_ERR_INVALID_ARGUMENT;
    if (n > 0 && (!d_rgba || !w || !h || !settings || !d_out)) return NTX_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < n; k++) {
        if (w[k] < 1 || h[k] < 1 || !d_rgba[k] || !d_out[k]) return NTX_ERR_INVALID_ARGUMENT;
        if (!settings[k].plain && (settings[k].refine_iters < 0 || settings[k].greedy_rounds < 0))
            return NTX_ERR_INVALID_ARGUMENT;
    }
    ntx_hip_context *ctx = m->ctx;
    ntx_status s = ntx::bind(ctx);
    if (s) return s;
    std::vector<const float *> plain_in;
    std::vector<void *> plain_out;
    std::vector<int> plain_w, plain_h, learned;
    for (int k = 0; k < n; k++) {
        if (settings[k].plain) {
            plain_in.push_back(d_rgba[k]);
            plain_out.push_back(d_out[k]);
            plain_w.push_back(w[k]);
            plain_h.push_back(h[k]);
        } else {
            learned.push_back(k);
        }
    }
    // Each learned image's fallback count and list in its own region of
    // m->fb, so that the images can run on two streams at once.
    std::vector<size_t> region(learned.size());
    size_t total = 0;
    for (size_t i = 0; i < learned.size(); i++) {
        const int k = learned[i];
        region[i] = total;
        total += 64 + ((size_t((w[k] + 3) / 4) * size_t((h[k] + 3) / 4) + 63) & ~size_t(63));
    }
    if (total && (s = ntx::ensure(ctx, m->fb, sizeof(int) * total))) return s;
    hipEvent_t join = nullptr, side_done = nullptr;
    if ((s = ntx::side_plain_bc1(ctx, int(plain_in.size()), plain_in.data(), plain_w.data(), plain_h.data(),
                                 plain_out.data(), &join)))
        return s;
    // Images of 2^17 blocks or more fill the device on the context's stream;
    // the smaller learned ones go one after another on side stream 1 beside
    // them, starting once the work already on the context's stream is done.
    hipStream_t side = nullptr;
    for (size_t i = 0; i < learned.size(); i++) {
        const int k = learned[i];
        const bool large = size_t((w[k] + 3) / 4) * size_t((h[k] + 3) / 4) >= (size_t(1) << 17);
        hipStream_t st = ctx->stream;
        if (!large) {
            if (!side) {
                if ((s = ntx::side_stream_after(ctx, 1, &side))) return s;
            }
            st = side;
        }
        int *d_count = static_cast<int *>(m->fb.ptr) + region[i];
        const ntx_mlp_settings &c = settings[k];
        if ((s = mlp_compress_on(m, st, d_count, d_count + 64, d_rgba[k], w[k], h[k], c.tau_pred, c.tau_gap,
                                 c.refine_iters, c.greedy_rounds, c.flags, d_out[k])))
            return s;
    }
    if (side) {
        if ((s = ntx::side_stream_done(ctx, 1, &side_done))) return s;
        NTX_HIP(ctx, hipStreamWaitEvent(ctx->stream, side_done, 0));
    }
    if (join) NTX_HIP(ctx, hipStreamWaitEvent(ctx->stream, join, 0));
    return NTX_OK;
}

}  // extern "C"
