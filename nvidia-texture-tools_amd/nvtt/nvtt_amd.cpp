// nvtt_amd.cpp -- nvtt::Compressor / Surface / options on the gfx950 backend.
//
// Mirrors the reference's Surface-API flow (Context.cpp:146-190, 486-516;
// Surface.cpp:155-218, 728-815, 1335-1406, 1562-1592) with every per-texel
// step on the device through include/ntx_hip.h.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "nvtt.h"
#include "ntx_hip.h"

namespace nvtt {

// ---------------------------------------------------------------- backend --

namespace {

// One backend context per process for Surface storage and compression (device
// 0), created on first use.  Calls on it are serialised by backend_lock().
std::mutex &backend_lock()
{
    static std::mutex m;
    return m;
}

ntx_hip_context *backend()
{
    static ntx_hip_context *ctx = nullptr;
    static bool tried = false;
    if (!tried) {
        tried = true;
        if (ntx_hip_is_hardware_present() && ntx_hip_create(0, &ctx) != NTX_OK) ctx = nullptr;
    }
    return ctx;
}

void report(const char *what, ntx_status s)
{
    std::fprintf(stderr, "nvtt (gfx950 backend): %s failed: %s\n", what, ntx_hip_status_string(s));
}

// Surface.cpp:155-207 (nv::canMakeNextMipmap / countMipmaps*).
bool can_make_next(unsigned w, unsigned h, unsigned d, unsigned min_size)
{
    if (min_size == 1u) {
        if (w == 1u && h == 1u && d == 1u) return false;
    } else if ((w <= min_size || h <= min_size) && d == 1u) {
        return false;
    }
    return true;
}

unsigned half_step(unsigned v) { return v / 2 > 1 ? v / 2 : 1; }

// IEEE binary16 -> binary32 (exact), for InputFormat_RGBA_16F.
float half_to_float(uint16_t h)
{
    const uint32_t sign = uint32_t(h & 0x8000) << 16;
    uint32_t exp = (h >> 10) & 0x1f, man = h & 0x3ff, bits;
    if (exp == 0) {
        if (man == 0) {
            bits = sign;
        } else {
            exp = 113;
            while (!(man & 0x400)) { man <<= 1; exp--; }
            bits = sign | (exp << 23) | ((man & 0x3ff) << 13);
        }
    } else if (exp == 31) {
        bits = sign | 0x7f800000u | (man << 13);
    } else {
        bits = sign | ((exp + 112) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

}  // namespace

// ---------------------------------------------------------------- options --

struct CompressionOptions::Private {
    Format format;
    Quality quality;
    float colorWeight[4];
};

CompressionOptions::CompressionOptions() : m(*new Private) { reset(); }
CompressionOptions::~CompressionOptions() { delete &m; }

// CompressionOptions.cpp:48-75 defaults.
void CompressionOptions::reset()
{
    m.format = Format_DXT1;
    m.quality = Quality_Normal;
    m.colorWeight[0] = m.colorWeight[1] = m.colorWeight[2] = m.colorWeight[3] = 1.0f;
}
void CompressionOptions::setFormat(Format format) { m.format = format; }
void CompressionOptions::setQuality(Quality quality) { m.quality = quality; }
void CompressionOptions::setColorWeights(float r, float g, float b, float a)
{
    m.colorWeight[0] = r; m.colorWeight[1] = g; m.colorWeight[2] = b; m.colorWeight[3] = a;
}
Format CompressionOptions::format() const { return m.format; }

struct OutputOptions::Private {
    OutputHandler *outputHandler = nullptr;
    ErrorHandler *errorHandler = nullptr;
    // OutputOptions.cpp:159-177.
    void beginImage(int size, int w, int h, int d, int face, int mip) const
    {
        if (outputHandler) outputHandler->beginImage(size, w, h, d, face, mip);
    }
    bool writeData(const void *data, int size) const { return outputHandler == nullptr || outputHandler->writeData(data, size); }
    void endImage() const
    {
        if (outputHandler) outputHandler->endImage();
    }
    void error(Error e) const
    {
        if (errorHandler) errorHandler->error(e);
    }
};

OutputOptions::OutputOptions() : m(*new Private) {}
OutputOptions::~OutputOptions() { delete &m; }
void OutputOptions::reset() { m.outputHandler = nullptr; m.errorHandler = nullptr; }
void OutputOptions::setOutputHandler(OutputHandler *h) { m.outputHandler = h; }
void OutputOptions::setErrorHandler(ErrorHandler *h) { m.errorHandler = h; }

// ---------------------------------------------------------------- surface --

struct Surface::Private {
    int refs = 1;
    int w = 0, h = 0, d = 0;
    TextureType type = TextureType_2D;
    WrapMode wrapMode = WrapMode_Mirror;   // Surface.h:47-50 defaults
    AlphaMode alphaMode = AlphaMode_None;
    bool isNormalMap = false;
    float *dev = nullptr;                  // 4 planes on the device
    mutable std::vector<float> host;       // lazily downloaded copy
    mutable bool hostValid = false;

    size_t count() const { return size_t(w) * size_t(h) * size_t(d); }
    size_t bytes() const { return 4 * count() * sizeof(float); }

    ~Private()
    {
        if (dev) {
            std::lock_guard<std::mutex> g(backend_lock());
            ntx_hip_free(backend(), dev);
        }
    }
};

Surface::Surface() : m(new Private) {}
Surface::Surface(const Surface &s) : m(s.m) { m->refs++; }
Surface::~Surface()
{
    if (--m->refs == 0) delete m;
}
void Surface::operator=(const Surface &s)
{
    if (s.m == m) return;
    s.m->refs++;
    if (--m->refs == 0) delete m;
    m = s.m;
}

namespace {

// Surface.cpp:354-363 detach(): give this surface its own storage before a write.
bool detach(Surface::Private *&m)
{
    if (m->refs == 1) return true;
    Surface::Private *c = new Surface::Private;
    c->w = m->w; c->h = m->h; c->d = m->d;
    c->type = m->type; c->wrapMode = m->wrapMode; c->alphaMode = m->alphaMode; c->isNormalMap = m->isNormalMap;
    if (m->dev) {
        std::lock_guard<std::mutex> g(backend_lock());
        ntx_hip_context *ctx = backend();
        void *p = nullptr;
        ntx_status s = ntx_hip_malloc(ctx, m->bytes(), &p);
        if (s == NTX_OK) s = ntx_hip_copy_device(ctx, p, m->dev, m->bytes());
        if (s != NTX_OK) { report("surface copy", s); delete c; return false; }
        c->dev = static_cast<float *>(p);
    }
    m->refs--;
    m = c;
    return true;
}

}  // namespace

void Surface::setWrapMode(WrapMode mode) { if (m->wrapMode != mode && detach(m)) m->wrapMode = mode; }
void Surface::setAlphaMode(AlphaMode mode) { if (m->alphaMode != mode && detach(m)) m->alphaMode = mode; }
void Surface::setNormalMap(bool b) { if (m->isNormalMap != b && detach(m)) m->isNormalMap = b; }
bool Surface::isNull() const { return m->dev == nullptr; }
int Surface::width() const { return m->w; }
int Surface::height() const { return m->h; }
int Surface::depth() const { return m->d; }
TextureType Surface::type() const { return m->type; }
WrapMode Surface::wrapMode() const { return m->wrapMode; }
AlphaMode Surface::alphaMode() const { return m->alphaMode; }
bool Surface::isNormalMap() const { return m->isNormalMap; }

int Surface::countMipmaps() const
{
    unsigned w = m->w, h = m->h, d = m->d, n = 0;
    while (w != 1 || h != 1 || d != 1) { w = half_step(w); h = half_step(h); d = half_step(d); n++; }
    return int(n + 1);
}

int Surface::countMipmaps(int min_size) const
{
    unsigned w = m->w, h = m->h, d = m->d, n = 0;
    while (can_make_next(w, h, d, unsigned(min_size))) { w = half_step(w); h = half_step(h); d = half_step(d); n++; }
    return int(n + 1);
}

const float *Surface::deviceData() const { return m->dev; }

const float *Surface::data() const
{
    if (!m->dev) return nullptr;
    if (!m->hostValid) {
        std::lock_guard<std::mutex> g(backend_lock());
        m->host.resize(4 * m->count());
        ntx_status s = ntx_hip_download(backend(), m->host.data(), m->dev, m->bytes());
        if (s != NTX_OK) { report("surface download", s); return nullptr; }
        m->hostValid = true;
    }
    return m->host.data();
}

const float *Surface::channel(int i) const
{
    if (i < 0 || i > 3) return nullptr;
    const float *p = data();
    return p ? p + size_t(i) * m->count() : nullptr;
}

// Surface.cpp:728-815 (2D only on this path).
bool Surface::setImage(int w, int h, int d)
{
    if (w <= 0 || h <= 0 || d != 1) return false;
    if (!detach(m)) return false;
    std::lock_guard<std::mutex> g(backend_lock());
    ntx_hip_context *ctx = backend();
    if (!ctx) { report("setImage", NTX_ERR_NO_DEVICE); return false; }
    void *p = nullptr;
    const size_t bytes = 4 * size_t(w) * size_t(h) * sizeof(float);
    ntx_status s = ntx_hip_malloc(ctx, bytes, &p);
    if (s == NTX_OK) s = ntx_hip_memset_device(ctx, p, 0, bytes);
    if (s != NTX_OK) { report("setImage", s); return false; }
    if (m->dev) ntx_hip_free(ctx, m->dev);
    m->dev = static_cast<float *>(p);
    m->w = w; m->h = h; m->d = d;
    m->type = TextureType_2D;
    m->hostValid = false;
    return true;
}

bool Surface::setImage(InputFormat format, int w, int h, int d, const void *data)
{
    if (w <= 0 || h <= 0 || d != 1 || !data) return false;
    if (!detach(m)) return false;
    std::lock_guard<std::mutex> g(backend_lock());
    ntx_hip_context *ctx = backend();
    if (!ctx) { report("setImage", NTX_ERR_NO_DEVICE); return false; }
    const size_t n = size_t(w) * size_t(h);
    void *p = nullptr;
    ntx_status s = ntx_hip_malloc(ctx, 4 * n * sizeof(float), &p);
    if (s != NTX_OK) { report("setImage", s); return false; }
    float *dst = static_cast<float *>(p);
    if (format == InputFormat_BGRA_8UB) {
        // Surface.cpp:745-761: float(byte) / 255 per channel, done on the device.
        void *staging = nullptr;
        s = ntx_hip_malloc(ctx, 4 * n, &staging);
        if (s == NTX_OK) s = ntx_hip_upload(ctx, staging, data, 4 * n);
        if (s == NTX_OK) s = ntx_hip_unpack_bgra8_device(ctx, static_cast<const uint8_t *>(staging), w, h, dst);
        if (s == NTX_OK) s = ntx_hip_synchronize(ctx);
        if (staging) ntx_hip_free(ctx, staging);
    } else {
        std::vector<float> planar(4 * n);
        if (format == InputFormat_RGBA_16F) {
            const uint16_t *src = static_cast<const uint16_t *>(data);
            for (size_t i = 0; i < n; i++)
                for (int c = 0; c < 4; c++) planar[c * n + i] = half_to_float(src[4 * i + c]);
        } else if (format == InputFormat_RGBA_32F) {
            const float *src = static_cast<const float *>(data);
            for (size_t i = 0; i < n; i++)
                for (int c = 0; c < 4; c++) planar[c * n + i] = src[4 * i + c];
        } else {  // InputFormat_R_32F
            const float *src = static_cast<const float *>(data);
            for (size_t i = 0; i < n; i++) planar[i] = src[i];
        }
        s = ntx_hip_upload(ctx, dst, planar.data(), planar.size() * sizeof(float));
    }
    if (s != NTX_OK) { ntx_hip_free(ctx, dst); report("setImage", s); return false; }
    if (m->dev) ntx_hip_free(ctx, m->dev);
    m->dev = dst;
    m->w = w; m->h = h; m->d = d;
    m->type = TextureType_2D;
    m->hostValid = false;
    return true;
}

bool Surface::canMakeNextMipmap(int min_size)
{
    if (isNull()) return false;
    return can_make_next(unsigned(m->w), unsigned(m->h), unsigned(m->d), unsigned(min_size));
}

// Surface.cpp:1121-1141 getDefaultFilterWidthAndParams.
bool Surface::buildNextMipmap(MipmapFilter filter, int min_size)
{
    float width = 0.5f;
    float params[2] = { 0.0f, 0.0f };
    if (filter == MipmapFilter_Triangle) width = 1.0f;
    else if (filter == MipmapFilter_Kaiser) { width = 3.0f; params[0] = 4.0f; params[1] = 1.0f; }
    return buildNextMipmap(filter, width, params, min_size);
}

// Surface.cpp:1344-1406: one device downsample into a fresh buffer.
bool Surface::buildNextMipmap(MipmapFilter filter, float filterWidth, const float *params, int min_size)
{
    if (!canMakeNextMipmap(min_size)) return false;
    if (!detach(m)) return false;
    std::lock_guard<std::mutex> g(backend_lock());
    ntx_hip_context *ctx = backend();
    const int w = int(half_step(unsigned(m->w))), h = int(half_step(unsigned(m->h)));
    void *p = nullptr;
    ntx_status s = ntx_hip_malloc(ctx, 4 * size_t(w) * size_t(h) * sizeof(float), &p);
    if (s == NTX_OK)
        s = ntx_hip_downsample_device(ctx, int(filter), filterWidth, params, int(m->wrapMode), int(m->alphaMode), m->dev,
                                      m->w, m->h, static_cast<float *>(p));
    if (s == NTX_OK) s = ntx_hip_synchronize(ctx);
    if (s != NTX_OK) {
        if (p) ntx_hip_free(ctx, p);
        report("buildNextMipmap", s);
        return false;
    }
    ntx_hip_free(ctx, m->dev);
    m->dev = static_cast<float *>(p);
    m->w = w; m->h = h;
    m->hostValid = false;
    return true;
}

void Surface::toLinearFromSrgb()
{
    if (isNull() || !detach(m)) return;
    std::lock_guard<std::mutex> g(backend_lock());
    ntx_hip_context *ctx = backend();
    ntx_status s = ntx_hip_srgb_to_linear_device(ctx, m->dev, m->w, m->h);
    if (s == NTX_OK) s = ntx_hip_synchronize(ctx);
    if (s != NTX_OK) report("toLinearFromSrgb", s);
    m->hostValid = false;
}

// ------------------------------------------------------------- compressor --

// The backend seam: same virtual as nv::CompressorInterface (Compressor.h:34-38),
// with the input known to be host or device memory.
struct CompressorInterface {
    virtual ~CompressorInterface() {}
    virtual void compress(AlphaMode alphaMode, unsigned w, unsigned h, unsigned d, const float *rgba, TaskDispatcher *dispatcher,
                          const CompressionOptions::Private &co, const OutputOptions::Private &oo) = 0;
};

// The CudaCompressor replacement (CudaCompressorDXT.cpp:128-206): compresses
// on the device and hands the whole level to writeData in one call, on the
// calling thread, in block-raster order.
struct HipCompressorBCn : CompressorInterface {
    explicit HipCompressorBCn(bool device_input) : device_input(device_input) {}
    bool device_input;
    void compress(AlphaMode alphaMode, unsigned w, unsigned h, unsigned d, const float *rgba, TaskDispatcher *,
                  const CompressionOptions::Private &co, const OutputOptions::Private &oo) override
    {
        if (d != 1) { oo.error(Error_UnsupportedFeature); return; }
        const size_t size = ntx_hip_compressed_size(int(co.format), int(w), int(h));
        std::vector<uint8_t> out(size);
        ntx_status s;
        {
            std::lock_guard<std::mutex> g(backend_lock());
            ntx_hip_context *ctx = backend();
            if (!device_input) {
                s = ntx_hip_compress(ctx, int(co.format), int(co.quality), int(alphaMode), co.colorWeight, rgba, int(w),
                                     int(h), out.data(), size);
            } else {
                void *dout = nullptr;
                s = ntx_hip_malloc(ctx, size, &dout);
                if (s == NTX_OK)
                    s = ntx_hip_compress_device(ctx, int(co.format), int(co.quality), int(alphaMode), co.colorWeight, rgba,
                                                int(w), int(h), dout);
                if (s == NTX_OK) s = ntx_hip_download(ctx, out.data(), dout, size);
                if (dout) ntx_hip_free(ctx, dout);
            }
        }
        if (s == NTX_ERR_UNSUPPORTED) { oo.error(Error_UnsupportedFeature); return; }
        if (s == NTX_ERR_INVALID_ARGUMENT) { oo.error(Error_InvalidInput); return; }
        if (s != NTX_OK) { oo.error(Error_CudaError); return; }
        oo.writeData(out.data(), int(size));
    }
};

struct Compressor::Private {
    bool hipEnabled = false;
    TaskDispatcher *dispatcher = nullptr;

    // Context.cpp:1166-1224 chooseGpuCompressor: the formats and qualities the
    // device implements exactly; anything else gets no compressor.
    CompressorInterface *chooseCompressor(const CompressionOptions::Private &co, bool device_input) const
    {
        switch (co.format) {
        case Format_BC1:
            if (co.quality == Quality_Normal || co.quality == Quality_Highest) return new HipCompressorBCn(device_input);
            break;
        case Format_BC3:
            if (co.quality == Quality_Normal || co.quality == Quality_Production) return new HipCompressorBCn(device_input);
            break;
        case Format_BC4:
        case Format_BC5:
            if (co.quality == Quality_Fastest || co.quality == Quality_Normal) return new HipCompressorBCn(device_input);
            break;
        default:
            break;
        }
        return nullptr;
    }

    // Context.cpp:486-516.
    bool compress(AlphaMode alphaMode, int w, int h, int d, int face, int mipmap, const float *rgba, bool device_input,
                  const CompressionOptions::Private &co, const OutputOptions::Private &oo) const
    {
        const int size = int(ntx_hip_compressed_size(int(co.format), w, h)) * (d > 0 ? d : 1);
        oo.beginImage(size, w, h, d, face, mipmap);
        if (!hipEnabled || !backend()) {
            oo.error(Error_CudaError);
        } else {
            CompressorInterface *c = chooseCompressor(co, device_input);
            if (!c) {
                oo.error(Error_UnsupportedFeature);
            } else {
                c->compress(alphaMode, unsigned(w), unsigned(h), unsigned(d), rgba, dispatcher, co, oo);
                delete c;
            }
        }
        oo.endImage();
        return true;
    }
};

struct BatchList::Private {
    struct Entry {
        const Surface *img;
        int face, mipmap;
        bool chain;
        MipmapFilter filter;
        int levels;
        const OutputOptions *oo;
    };
    std::vector<Entry> items;
};

BatchList::BatchList() : m(new Private) {}
BatchList::~BatchList() { delete m; }
void BatchList::Clear() { m->items.clear(); }
void BatchList::Append(const Surface *img, int face, int mipmap, const OutputOptions *oo)
{
    m->items.push_back({ img, face, mipmap, false, MipmapFilter_Box, 1, oo });
}
void BatchList::AppendChain(const Surface *img, int face, int mipmap, MipmapFilter filter, int mipmapCount,
                            const OutputOptions *oo)
{
    m->items.push_back({ img, face, mipmap, true, filter, mipmapCount, oo });
}
unsigned BatchList::GetSize() const { return unsigned(m->items.size()); }

namespace {

// Surface::buildNextMipmap(filter)'s defaults (Surface.cpp:1121-1141).
void filter_defaults(MipmapFilter f, float *width, float params[2])
{
    *width = f == MipmapFilter_Kaiser ? 3.0f : f == MipmapFilter_Triangle ? 1.0f : 0.5f;
    params[0] = f == MipmapFilter_Kaiser ? 4.0f : 0.0f;
    params[1] = f == MipmapFilter_Kaiser ? 1.0f : 0.0f;
}

// Level count of an entry: the chain length asked for, bounded by the
// surface's own mip count; one for a single image.
int entry_levels(const BatchList::Private::Entry &e)
{
    if (!e.chain) return 1;
    const int full = e.img->countMipmaps();
    return e.levels <= 0 || e.levels > full ? full : e.levels;
}

}  // namespace

// Context.cpp:61-73: acceleration is enabled when a device is present.
Compressor::Compressor() : m(*new Private) { enableCudaAcceleration(true); }
Compressor::~Compressor() { delete &m; }

void Compressor::enableCudaAcceleration(bool enable)
{
    if (!enable) { m.hipEnabled = false; return; }
    std::lock_guard<std::mutex> g(backend_lock());
    m.hipEnabled = backend() != nullptr;
}

bool Compressor::isCudaAccelerationEnabled() const { return m.hipEnabled; }
void Compressor::setTaskDispatcher(TaskDispatcher *disp) { m.dispatcher = disp; }

bool Compressor::compress(const Surface &img, int face, int mipmap, const CompressionOptions &co,
                          const OutputOptions &oo) const
{
    return m.compress(img.alphaMode(), img.width(), img.height(), img.depth(), face, mipmap, img.deviceData(), true, co.m,
                      oo.m);
}

bool Compressor::compress(int w, int h, int d, int face, int mipmap, const float *rgba, const CompressionOptions &co,
                          const OutputOptions &oo) const
{
    return m.compress(AlphaMode_None, w, h, d, face, mipmap, rgba, false, co.m, oo.m);
}

bool Compressor::compress(const BatchList &list, const CompressionOptions &co) const
{
    typedef BatchList::Private::Entry Entry;
    const std::vector<Entry> &items = list.m->items;
    // Every image of every entry, in delivery order: (entry, level, width, height).
    struct Image { int entry, level, w, h; };
    std::vector<Image> images;
    for (int i = 0; i < int(items.size()); i++) {
        int w = items[i].img->width(), h = items[i].img->height();
        for (int k = 0; k < entry_levels(items[i]); k++) {
            images.push_back({ i, k, w, h });
            w = w / 2 > 1 ? w / 2 : 1;
            h = h / 2 > 1 ? h / 2 : 1;
        }
    }
    const int fmt = int(co.m.format);
    auto fail_all = [&](Error e) {
        for (const Image &im : images) {
            const OutputOptions::Private &oo = items[im.entry].oo->m;
            oo.beginImage(int(ntx_hip_compressed_size(fmt, im.w, im.h)), im.w, im.h, 1, items[im.entry].face,
                          items[im.entry].mipmap + im.level);
            oo.error(e);
            oo.endImage();
        }
        return true;
    };
    if (!m.hipEnabled || !backend()) return fail_all(Error_CudaError);
    CompressorInterface *probe = m.chooseCompressor(co.m, true);
    if (!probe) return fail_all(Error_UnsupportedFeature);
    delete probe;
    for (const Entry &e : items)
        if (!e.img || e.img->isNull() || e.img->depth() != 1 || !e.oo) return fail_all(Error_InvalidInput);

    // The compressed levels, gathered on the host before delivery.
    std::vector<std::vector<uint8_t>> bytes(images.size());
    ntx_status s = NTX_OK;
    {
        std::lock_guard<std::mutex> g(backend_lock());
        ntx_hip_context *ctx = backend();
        if (co.m.format == Format_BC1 || co.m.format == Format_BC4 || co.m.format == Format_BC5) {
            // One persistent launch for the whole list (ntx_hip_batch_*).
            std::vector<ntx_texture_desc> descs(items.size());
            for (size_t i = 0; i < items.size(); i++) {
                const Entry &e = items[i];
                ntx_texture_desc &d = descs[i];
                std::memset(&d, 0, sizeof d);
                d.d_rgba = e.img->deviceData();
                d.width = e.img->width();
                d.height = e.img->height();
                d.format = fmt;
                d.quality = int(co.m.quality);
                d.alpha_mode = int(e.img->alphaMode());
                for (int c = 0; c < 4; c++) d.color_weights[c] = co.m.colorWeight[c];
                d.mip_levels = entry_levels(e);
                d.filter = int(e.filter);
                d.filter_width = 0.0f;
                d.wrap = int(e.img->wrapMode());
            }
            ntx_hip_batch *b = nullptr;
            s = ntx_hip_batch_create(ctx, descs.data(), int(descs.size()), &b);
            if (s == NTX_OK) s = ntx_hip_batch_submit(b);
            if (s == NTX_OK) s = ntx_hip_batch_wait(b);
            for (size_t j = 0; s == NTX_OK && j < images.size(); j++) {
                const float *lv = nullptr;
                const void *blocks = nullptr;
                int lw = 0, lh = 0;
                size_t n = 0;
                s = ntx_hip_batch_level(b, images[j].entry, images[j].level, &lv, &lw, &lh, &blocks, &n);
                if (s == NTX_OK) {
                    bytes[j].resize(n);
                    s = ntx_hip_download(ctx, bytes[j].data(), blocks, n);
                }
            }
            if (b) ntx_hip_batch_destroy(b);
        } else {
            // BC3: each entry's chain through the chain scheduler.
            size_t j = 0;
            for (size_t i = 0; s == NTX_OK && i < items.size(); i++) {
                const Entry &e = items[i];
                const int levels = entry_levels(e);
                std::vector<float *> lv(levels, nullptr);
                std::vector<void *> out(levels, nullptr);
                lv[0] = const_cast<float *>(e.img->deviceData());
                for (int k = 0; s == NTX_OK && k < levels; k++) {
                    const Image &im = images[j + k];
                    if (k > 0) {
                        void *p = nullptr;
                        s = ntx_hip_malloc(ctx, 4 * size_t(im.w) * size_t(im.h) * sizeof(float), &p);
                        lv[k] = static_cast<float *>(p);
                    }
                    if (s == NTX_OK) s = ntx_hip_malloc(ctx, ntx_hip_compressed_size(fmt, im.w, im.h), &out[k]);
                }
                float width, params[2];
                filter_defaults(e.filter, &width, params);
                if (s == NTX_OK)
                    s = ntx_hip_compress_chain_device(ctx, fmt, int(co.m.quality), int(e.img->alphaMode()), co.m.colorWeight,
                                                      int(e.filter), width, params, int(e.img->wrapMode()),
                                                      e.img->width(), e.img->height(), levels, lv.data(), out.data());
                if (s == NTX_OK) s = ntx_hip_synchronize(ctx);
                for (int k = 0; s == NTX_OK && k < levels; k++) {
                    const Image &im = images[j + k];
                    bytes[j + k].resize(ntx_hip_compressed_size(fmt, im.w, im.h));
                    s = ntx_hip_download(ctx, bytes[j + k].data(), out[k], bytes[j + k].size());
                }
                for (int k = 0; k < levels; k++) {
                    if (k > 0 && lv[k]) ntx_hip_free(ctx, lv[k]);
                    if (out[k]) ntx_hip_free(ctx, out[k]);
                }
                j += size_t(levels);
            }
        }
    }
    if (s != NTX_OK) {
        report("compress(BatchList)", s);
        return fail_all(s == NTX_ERR_INVALID_ARGUMENT ? Error_InvalidInput
                        : s == NTX_ERR_UNSUPPORTED   ? Error_UnsupportedFeature
                                                     : Error_CudaError);
    }
    // Context.cpp:486-516, per image, on this thread.
    for (size_t j = 0; j < images.size(); j++) {
        const Image &im = images[j];
        const Entry &e = items[im.entry];
        const OutputOptions::Private &oo = e.oo->m;
        oo.beginImage(int(bytes[j].size()), im.w, im.h, 1, e.face, e.mipmap + im.level);
        oo.writeData(bytes[j].data(), int(bytes[j].size()));
        oo.endImage();
    }
    return true;
}

// Context.cpp:192-213 estimateSize for block formats.
int Compressor::estimateSize(int w, int h, int d, int mipmapCount, const CompressionOptions &co) const
{
    int size = 0;
    for (int i = 0; i < mipmapCount; i++) {
        size += int(ntx_hip_compressed_size(int(co.m.format), w, h)) * d;
        w = w / 2 > 1 ? w / 2 : 1;
        h = h / 2 > 1 ? h / 2 : 1;
        d = d / 2 > 1 ? d / 2 : 1;
    }
    return size;
}

int Compressor::estimateSize(const Surface &img, int mipmapCount, const CompressionOptions &co) const
{
    return estimateSize(img.width(), img.height(), img.depth(), mipmapCount, co);
}

}  // namespace nvtt
