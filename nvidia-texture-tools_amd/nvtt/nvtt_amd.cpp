// nvtt_amd.cpp -- nvtt::Compressor / Surface / options on the gfx950 backend.
//
// Mirrors the reference's Surface-API flow (Context.cpp:146-190, 486-516;
// Surface.cpp:155-218, 728-815, 1335-1406, 1562-1592) with every per-texel
// step on the device through include/ntx_hip.h.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <mutex>
#include <string>
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
        // NTX_SCRATCH_LIMIT_BYTES: cap on the backend's device scratch
        // (ntx_hip_set_scratch_limit); images that need more report
        // Error_CudaError.
        if (ctx) {
            if (const char *lim = std::getenv("NTX_SCRATCH_LIMIT_BYTES"))
                ntx_hip_set_scratch_limit(ctx, size_t(std::strtoull(lim, nullptr, 10)));
        }
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
    unsigned int bitcount, rmask, gmask, bmask, amask;
    bool enableColorDithering, enableAlphaDithering, binaryAlpha;
    int alphaThreshold;
    std::vector<float> learned;  // NTX_MLP_PARAMS floats, or empty
};

CompressionOptions::CompressionOptions() : m(*new Private) { reset(); }
CompressionOptions::~CompressionOptions() { delete &m; }

// CompressionOptions.cpp:48-75 defaults.
void CompressionOptions::reset()
{
    m.format = Format_DXT1;
    m.quality = Quality_Normal;
    m.colorWeight[0] = m.colorWeight[1] = m.colorWeight[2] = m.colorWeight[3] = 1.0f;
    m.bitcount = 32;
    m.rmask = 0x00FF0000u; m.gmask = 0x0000FF00u; m.bmask = 0x000000FFu; m.amask = 0xFF000000u;
    m.enableColorDithering = m.enableAlphaDithering = m.binaryAlpha = false;
    m.alphaThreshold = 127;
    m.learned.clear();
}
void CompressionOptions::setPixelFormat(unsigned int bitcount, unsigned int rmask, unsigned int gmask, unsigned int bmask,
                                        unsigned int amask)
{
    m.bitcount = bitcount; m.rmask = rmask; m.gmask = gmask; m.bmask = bmask; m.amask = amask;
}
void CompressionOptions::setQuantization(bool colorDithering, bool alphaDithering, bool binaryAlpha, int alphaThreshold)
{
    m.enableColorDithering = colorDithering;
    m.enableAlphaDithering = alphaDithering;
    m.binaryAlpha = binaryAlpha;
    m.alphaThreshold = alphaThreshold;
}

bool CompressionOptions::setLearnedBC1(const float *params, int count)
{
    if (!params || count == 0) {
        m.learned.clear();
        return true;
    }
    if (count != NTX_MLP_PARAMS) return false;
    m.learned.assign(params, params + count);
    return true;
}

// nvtt.cpp:31-59 (worded here).
const char *errorString(Error e)
{
    switch (e) {
    case Error_Unknown: return "unknown error";
    case Error_InvalidInput: return "invalid input";
    case Error_UnsupportedFeature: return "unsupported feature";
    case Error_CudaError: return "device (HIP) error";
    case Error_FileOpen: return "cannot open the output file";
    case Error_FileWrite: return "cannot write through the output handler";
    case Error_UnsupportedOutputFormat: return "the container does not support the output format";
    default: return "invalid error code";
    }
}

unsigned int version() { return NVTT_VERSION; }
void CompressionOptions::setFormat(Format format) { m.format = format; }
void CompressionOptions::setQuality(Quality quality) { m.quality = quality; }
void CompressionOptions::setColorWeights(float r, float g, float b, float a)
{
    m.colorWeight[0] = r; m.colorWeight[1] = g; m.colorWeight[2] = b; m.colorWeight[3] = a;
}
Format CompressionOptions::format() const { return m.format; }

namespace {

// The handler setFileName installs (OutputOptions.cpp:60-90): every byte to
// one file, in order.
struct FileOutputHandler : OutputHandler {
    std::FILE *f = nullptr;
    explicit FileOutputHandler(const char *name) : f(std::fopen(name, "wb")) {}
    ~FileOutputHandler() override
    {
        if (f) std::fclose(f);
    }
    void beginImage(int, int, int, int, int, int) override {}
    bool writeData(const void *data, int size) override
    {
        return f && (size == 0 || std::fwrite(data, size_t(size), 1, f) == 1);
    }
    void endImage() override {}
};

}  // namespace

struct OutputOptions::Private {
    OutputHandler *outputHandler = nullptr;
    ErrorHandler *errorHandler = nullptr;
    std::string fileName;
    FileOutputHandler *fileHandler = nullptr;  // owned; the outputHandler when a file name is set
    bool outputHeader = true;                  // OutputOptions.cpp:44-57 defaults
    Container container = Container_DDS;
    int version = 0;
    bool srgb = false;

    void dropFile()
    {
        delete fileHandler;
        fileHandler = nullptr;
        fileName.clear();
    }
    // OutputOptions.cpp:149-157: with a file name the file must have opened.
    bool hasValidOutputHandler() const { return fileName.empty() || outputHandler != nullptr; }
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
OutputOptions::~OutputOptions()
{
    m.dropFile();
    delete &m;
}
void OutputOptions::reset()
{
    m.dropFile();
    m.outputHandler = nullptr;
    m.errorHandler = nullptr;
    m.outputHeader = true;
    m.container = Container_DDS;
    m.version = 0;
    m.srgb = false;
}
void OutputOptions::setFileName(const char *fileName)
{
    m.dropFile();
    m.outputHandler = nullptr;
    if (!fileName) return;
    m.fileName = fileName;
    m.fileHandler = new FileOutputHandler(fileName);
    if (m.fileHandler->f) {
        m.outputHandler = m.fileHandler;
    } else {
        delete m.fileHandler;
        m.fileHandler = nullptr;  // process() then reports Error_FileOpen
    }
}
void OutputOptions::setOutputHandler(OutputHandler *h)
{
    m.dropFile();
    m.outputHandler = h;
}
void OutputOptions::setErrorHandler(ErrorHandler *h) { m.errorHandler = h; }
void OutputOptions::setOutputHeader(bool b) { m.outputHeader = b; }
void OutputOptions::setContainer(Container c) { m.container = c; }
void OutputOptions::setUserVersion(int version) { m.version = version; }
void OutputOptions::setSrgbFlag(bool b) { m.srgb = b; }

// ---------------------------------------------------------- input options --

struct InputOptions::Private {
    // InputOptions.cpp:80-110 defaults.
    WrapMode wrapMode;
    TextureType textureType;
    InputFormat inputFormat;
    AlphaMode alphaMode;
    float inputGamma, outputGamma;
    bool generateMipmaps;
    int maxLevel;
    MipmapFilter mipmapFilter;
    float kaiserWidth, kaiserAlpha, kaiserStretch;
    bool isNormalMap, normalizeMipmaps, convertToNormalMap;
    float heightFactors[4], bumpFrequencyScale[4];
    int maxExtent;
    RoundMode roundMode;
    // Layout (InputOptions.cpp:113-185): faceCount faces x mipmapCount levels,
    // image (mip, face) at mip * faceCount + face; empty = not supplied.
    int width = 0, height = 0, depth = 0, faceCount = 0, mipmapCount = 0;
    std::vector<std::vector<uint8_t>> images;
};

InputOptions::InputOptions() : m(*new Private) { reset(); }
InputOptions::~InputOptions() { delete &m; }

void InputOptions::reset()
{
    m.wrapMode = WrapMode_Mirror;
    m.textureType = TextureType_2D;
    m.inputFormat = InputFormat_BGRA_8UB;
    m.alphaMode = AlphaMode_None;
    m.inputGamma = 2.2f;
    m.outputGamma = 2.2f;
    m.generateMipmaps = true;
    m.maxLevel = -1;
    m.mipmapFilter = MipmapFilter_Box;
    m.kaiserWidth = 3.0f;
    m.kaiserAlpha = 4.0f;
    m.kaiserStretch = 1.0f;
    m.isNormalMap = false;
    m.normalizeMipmaps = true;
    m.convertToNormalMap = false;
    const float hf[4] = { 0.0f, 0.0f, 0.0f, 1.0f };
    const float bs = 1.0f + 0.5f + 0.25f + 0.125f;
    const float bf[4] = { 1.0f / bs, 0.5f / bs, 0.25f / bs, 0.125f / bs };
    std::memcpy(m.heightFactors, hf, sizeof hf);
    std::memcpy(m.bumpFrequencyScale, bf, sizeof bf);
    m.maxExtent = 0;
    m.roundMode = RoundMode_None;
}

namespace {

int count_mipmaps3(int w, int h, int d)
{
    int n = 0;
    while (w != 1 || h != 1 || d != 1) {
        w = w / 2 > 1 ? w / 2 : 1;
        h = h / 2 > 1 ? h / 2 : 1;
        d = d / 2 > 1 ? d / 2 : 1;
        n++;
    }
    return n + 1;
}

size_t input_texel_bytes(InputFormat f)
{
    switch (f) {
    case InputFormat_BGRA_8UB: return 4;
    case InputFormat_RGBA_16F: return 8;
    case InputFormat_RGBA_32F: return 16;
    case InputFormat_R_32F: return 4;
    }
    return 0;
}

}  // namespace

// InputOptions.cpp:113-162.
void InputOptions::setTextureLayout(TextureType type, int w, int h, int d, int arraySize)
{
    if (w <= 0) w = 1;
    if (h <= 0) h = 1;
    if (d <= 0) d = 1;
    if (arraySize <= 0) arraySize = 1;
    resetTextureLayout();
    m.textureType = type;
    m.width = w;
    m.height = h;
    m.depth = d;
    m.faceCount = type == TextureType_Cube ? 6 : type == TextureType_Array ? arraySize : 1;
    m.mipmapCount = count_mipmaps3(w, h, d);
    m.images.assign(size_t(m.faceCount) * size_t(m.mipmapCount), std::vector<uint8_t>());
}

void InputOptions::resetTextureLayout()
{
    m.images.clear();
    m.faceCount = 0;
    m.mipmapCount = 0;
}

// InputOptions.cpp:188-245: the extents must be those of that level.
bool InputOptions::setMipmapData(const void *data, int w, int h, int d, int face, int mipmap)
{
    if (face < 0 || face >= m.faceCount || mipmap < 0 || mipmap >= m.mipmapCount || !data) return false;
    int ew = m.width, eh = m.height, ed = m.depth;
    for (int i = 0; i < mipmap; i++) {
        ew = ew / 2 > 1 ? ew / 2 : 1;
        eh = eh / 2 > 1 ? eh / 2 : 1;
        ed = ed / 2 > 1 ? ed / 2 : 1;
    }
    if (ew != w || eh != h || ed != d) return false;
    const size_t bytes = size_t(w) * size_t(h) * size_t(d) * input_texel_bytes(m.inputFormat);
    if (bytes == 0) return false;
    std::vector<uint8_t> &img = m.images[size_t(mipmap) * size_t(m.faceCount) + size_t(face)];
    img.assign(static_cast<const uint8_t *>(data), static_cast<const uint8_t *>(data) + bytes);
    return true;
}

void InputOptions::setFormat(InputFormat format) { m.inputFormat = format; }
void InputOptions::setAlphaMode(AlphaMode alphaMode) { m.alphaMode = alphaMode; }
void InputOptions::setGamma(float inputGamma, float outputGamma)
{
    m.inputGamma = inputGamma;
    m.outputGamma = outputGamma;
}
void InputOptions::setWrapMode(WrapMode mode) { m.wrapMode = mode; }
void InputOptions::setMipmapFilter(MipmapFilter filter) { m.mipmapFilter = filter; }
void InputOptions::setMipmapGeneration(bool enabled, int maxLevel)
{
    m.generateMipmaps = enabled;
    m.maxLevel = maxLevel;
}
void InputOptions::setKaiserParameters(float width, float alpha, float stretch)
{
    m.kaiserWidth = width;
    m.kaiserAlpha = alpha;
    m.kaiserStretch = stretch;
}
void InputOptions::setNormalMap(bool b) { m.isNormalMap = b; }
void InputOptions::setConvertToNormalMap(bool convert) { m.convertToNormalMap = convert; }
void InputOptions::setHeightEvaluation(float r, float g, float b, float a)
{
    // InputOptions.cpp:313-318: kept as given.
    m.heightFactors[0] = r; m.heightFactors[1] = g; m.heightFactors[2] = b; m.heightFactors[3] = a;
}
// InputOptions.cpp:321-325: normalised by their sum.
void InputOptions::setNormalFilter(float sm, float medium, float big, float large)
{
    const float total = sm + medium + big + large;
    m.bumpFrequencyScale[0] = sm / total; m.bumpFrequencyScale[1] = medium / total;
    m.bumpFrequencyScale[2] = big / total; m.bumpFrequencyScale[3] = large / total;
}
void InputOptions::setNormalizeMipmaps(bool b) { m.normalizeMipmaps = b; }
void InputOptions::setMaxExtents(int d) { m.maxExtent = d; }
void InputOptions::setRoundMode(RoundMode mode) { m.roundMode = mode; }

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

bool Surface::setImage2D(Format format, Decoder decoder, int w, int h, const void *data)
{
    if (w <= 0 || h <= 0 || !data) return false;
    if (format != Format_BC1 && format != Format_BC3 && format != Format_BC3n && format != Format_BC4 &&
        format != Format_BC5)
        return false;
    if (decoder != Decoder_D3D10 && decoder != Decoder_D3D9) return false;
    if (!detach(m)) return false;
    std::lock_guard<std::mutex> g(backend_lock());
    ntx_hip_context *ctx = backend();
    if (!ctx) { report("setImage2D", NTX_ERR_NO_DEVICE); return false; }
    const size_t bytes = ntx_hip_compressed_size(int(format), w, h);
    void *p = nullptr, *blocks = nullptr;
    ntx_status s = ntx_hip_malloc(ctx, 4 * size_t(w) * size_t(h) * sizeof(float), &p);
    if (s == NTX_OK) s = ntx_hip_malloc(ctx, bytes, &blocks);
    if (s == NTX_OK) s = ntx_hip_upload(ctx, blocks, data, bytes);
    if (s == NTX_OK) s = ntx_hip_decode_device(ctx, int(format), int(decoder), blocks, w, h, static_cast<float *>(p));
    if (s == NTX_OK) s = ntx_hip_synchronize(ctx);
    if (blocks) ntx_hip_free(ctx, blocks);
    if (s != NTX_OK) {
        if (p) ntx_hip_free(ctx, p);
        report("setImage2D", s);
        return false;
    }
    if (m->dev) ntx_hip_free(ctx, m->dev);
    m->dev = static_cast<float *>(p);
    m->w = w; m->h = h; m->d = 1;
    m->type = TextureType_2D;
    m->hostValid = false;
    return true;
}

namespace {

float rms_metric(const Surface &reference, const Surface &img, int metric)
{
    if (reference.isNull() || img.isNull() || reference.width() != img.width() || reference.height() != img.height() ||
        reference.depth() != img.depth())
        return FLT_MAX;  // ErrorMetric.cpp:15-17 sameLayout
    std::lock_guard<std::mutex> g(backend_lock());
    float r = FLT_MAX;
    ntx_status s = ntx_hip_rms_error_device(backend(), metric, reference.deviceData(), img.deviceData(),
                                            reference.width(), reference.height(), &r);
    if (s != NTX_OK) report("rmsError", s);
    return s == NTX_OK ? r : FLT_MAX;
}

}  // namespace

// Surface.cpp:3270-3274.
float rmsError(const Surface &reference, const Surface &img)
{
    return rms_metric(reference, img,
                      reference.alphaMode() == AlphaMode_Transparency ? NTX_METRIC_RMS_COLOR_ALPHA_WEIGHTED
                                                                      : NTX_METRIC_RMS_COLOR);
}

// Surface.cpp:3276-3280.
float rmsAlphaError(const Surface &reference, const Surface &img) { return rms_metric(reference, img, NTX_METRIC_RMS_ALPHA); }

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

namespace {

// nv::equal(gamma, 1.0f) (nvmath/nvmath.h:52, 139-143).
bool unit_gamma(float gamma)
{
    const float ag = std::fabs(gamma);
    return std::fabs(gamma - 1.0f) <= 0.0001f * (ag > 1.0f ? ag : 1.0f);
}

}  // namespace

void Surface::toLinear(float gamma)
{
    if (isNull() || unit_gamma(gamma) || !detach(m)) return;
    std::lock_guard<std::mutex> g(backend_lock());
    ntx_hip_context *ctx = backend();
    ntx_status s = ntx_hip_gamma_device(ctx, 0, gamma, m->dev, m->w, m->h, m->dev);
    if (s == NTX_OK) s = ntx_hip_synchronize(ctx);
    if (s != NTX_OK) report("toLinear", s);
    m->hostValid = false;
}

void Surface::toGamma(float gamma)
{
    if (isNull() || unit_gamma(gamma) || !detach(m)) return;
    std::lock_guard<std::mutex> g(backend_lock());
    ntx_hip_context *ctx = backend();
    ntx_status s = ntx_hip_gamma_device(ctx, 1, gamma, m->dev, m->w, m->h, m->dev);
    if (s == NTX_OK) s = ntx_hip_synchronize(ctx);
    if (s != NTX_OK) report("toGamma", s);
    m->hostValid = false;
}

// ------------------------------------------------------------- compressor --

namespace {

// Whether the learned BC1 form applies (CompressionOptions::setLearnedBC1):
// BC1 at Normal quality without transparency weighting or colour weights.
bool learned_bc1(const CompressionOptions::Private &co, AlphaMode alphaMode)
{
    return !co.learned.empty() && co.format == Format_BC1 && co.quality == Quality_Normal &&
           alphaMode != AlphaMode_Transparency && co.colorWeight[0] == 1.0f && co.colorWeight[1] == 1.0f &&
           co.colorWeight[2] == 1.0f;
}

// The learned form's settings for a w x h image at mip level `level`: the
// plain cluster fit under 4096 blocks; otherwise tau_pred 8 with no
// least-squares pass, greedy round or gap test at level 0, and one
// least-squares pass, one greedy round and a gap of 1.0 below it (the
// SHIPPED and REFINED settings of ntx_amd/learned.py).
ntx_mlp_settings learned_settings(int w, int h, int level)
{
    ntx_mlp_settings st{};
    if (long((w + 3) / 4) * long((h + 3) / 4) < 4096) {
        st.plain = 1;
        return st;
    }
    st.tau_pred = 8.0f;
    st.tau_gap = level == 0 ? -1.0f : 1.0f;
    st.refine_iters = level == 0 ? 0 : 1;
    st.greedy_rounds = level == 0 ? 0 : 1;
    return st;
}

// n device images (level[i]: the image's mip level) through the learned form
// with the network in co; waits for the device before the network is freed.
ntx_status compress_learned(ntx_hip_context *ctx, const CompressionOptions::Private &co, int n,
                            const float *const *d_in, const int *w, const int *h, const int *level, void *const *d_out)
{
    ntx_hip_mlp *mlp = nullptr;
    ntx_status s = ntx_hip_mlp_create(ctx, co.learned.data(), 0, &mlp);
    std::vector<ntx_mlp_settings> st(size_t(n > 0 ? n : 0));
    for (int i = 0; i < n; i++) st[size_t(i)] = learned_settings(w[i], h[i], level[i]);
    if (s == NTX_OK) s = ntx_hip_compress_bc1_mlp_images_device(mlp, n, d_in, w, h, st.data(), d_out);
    if (s == NTX_OK) s = ntx_hip_synchronize(ctx);
    if (mlp) ntx_hip_mlp_destroy(mlp);
    return s;
}

}  // namespace

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
    HipCompressorBCn(bool device_input, int level) : device_input(device_input), level(level) {}
    bool device_input;
    int level;  // the image's mip level (the learned form's settings depend on it)
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
            if (learned_bc1(co, alphaMode)) {
                void *din = nullptr, *dout = nullptr;
                const float *src = rgba;
                s = NTX_OK;
                if (!device_input) {
                    s = ntx_hip_malloc(ctx, 16 * size_t(w) * size_t(h), &din);
                    if (s == NTX_OK) s = ntx_hip_upload(ctx, din, rgba, 16 * size_t(w) * size_t(h));
                    src = static_cast<const float *>(din);
                }
                if (s == NTX_OK) s = ntx_hip_malloc(ctx, size, &dout);
                const int wi = int(w), hi = int(h);
                if (s == NTX_OK) s = compress_learned(ctx, co, 1, &src, &wi, &hi, &level, &dout);
                if (s == NTX_OK) s = ntx_hip_download(ctx, out.data(), dout, size);
                if (dout) ntx_hip_free(ctx, dout);
                if (din) ntx_hip_free(ctx, din);
            } else if (!device_input) {
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
    // Device memory process() keeps between calls (level planes, their gamma
    // copies, staging, compressed output), grown on demand, freed with the
    // Compressor.
    mutable void *arena = nullptr;
    mutable size_t arenaBytes = 0;
    ~Private()
    {
        if (arena) {
            std::lock_guard<std::mutex> g(backend_lock());
            ntx_hip_free(backend(), arena);
        }
    }

    // Context.cpp:1166-1224 chooseGpuCompressor: the formats and qualities the
    // device implements exactly; anything else gets no compressor.
    CompressorInterface *chooseCompressor(const CompressionOptions::Private &co, bool device_input, int level = 0) const
    {
        switch (co.format) {
        case Format_BC1:  // BlockCompressor.cpp:211-217: icbc Level1 / 8 / 9 by quality
            return new HipCompressorBCn(device_input, level);
        case Format_BC1a:  // CompressorDXT1a (CompressorDX9.cpp:83-111); not FastCompressorDXT1a
            if (co.quality != Quality_Fastest) return new HipCompressorBCn(device_input, level);
            break;
        // Context.cpp:1080-1116: Fast*/Production*/Optimal variants by quality,
        // every one reproduced by the device.
        case Format_BC3:
        case Format_BC3n:  // CompressorDXT5n (CompressorDX9.cpp:179-222)
        case Format_BC4:
        case Format_BC5:
            return new HipCompressorBCn(device_input, level);
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
            CompressorInterface *c = chooseCompressor(co, device_input, mipmap);
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

    bool outputHeader(TextureType type, int w, int h, int d, int arraySize, int mipmapCount, bool isNormalMap,
                      const CompressionOptions::Private &co, const OutputOptions::Private &oo) const;
    bool outputKtxHeader(TextureType type, int w, int h, int d, int arraySize, int mipmapCount,
                         const CompressionOptions::Private &co, const OutputOptions::Private &oo) const;
    bool process(const InputOptions::Private &io, const CompressionOptions::Private &co,
                 const OutputOptions::Private &oo) const;
};

namespace {

// DDS header constants (DirectDrawSurface.h/.cpp, the DDS_HEADER layout of
// the DirectDraw Surface format).
constexpr uint32_t kDDSD_CAPS = 0x1, kDDSD_HEIGHT = 0x2, kDDSD_WIDTH = 0x4, kDDSD_PITCH = 0x8,
                   kDDSD_PIXELFORMAT = 0x1000, kDDSD_MIPMAPCOUNT = 0x20000, kDDSD_LINEARSIZE = 0x80000,
                   kDDSD_DEPTH = 0x800000;
constexpr uint32_t kDDSCAPS_COMPLEX = 0x8, kDDSCAPS_TEXTURE = 0x1000, kDDSCAPS_MIPMAP = 0x400000;
constexpr uint32_t kDDSCAPS2_CUBEMAP = 0x200, kDDSCAPS2_CUBEMAP_ALL_FACES = 0xFC00, kDDSCAPS2_VOLUME = 0x200000;
constexpr uint32_t kDDPF_ALPHAPIXELS = 0x1, kDDPF_FOURCC = 0x4, kDDPF_NORMAL = 0x80000000u, kDDPF_SRGB = 0x40000000u;
constexpr uint32_t kDimTexture2D = 3, kDimTexture3D = 4, kMiscTextureCube = 0x4;
constexpr uint32_t kDXGI_BC1 = 71, kDXGI_BC1_SRGB = 72, kDXGI_BC2 = 74, kDXGI_BC2_SRGB = 75, kDXGI_BC3 = 77,
                   kDXGI_BC3_SRGB = 78, kDXGI_BC4 = 80, kDXGI_BC5 = 83;

constexpr uint32_t fourcc(char a, char b, char c, char d)
{
    return uint32_t(uint8_t(a)) | uint32_t(uint8_t(b)) << 8 | uint32_t(uint8_t(c)) << 16 | uint32_t(uint8_t(d)) << 24;
}

// The 148-byte header (128 + the DX10 extension), as 37 little-endian words.
struct DdsHeader {
    uint32_t magic = fourcc('D', 'D', 'S', ' ');
    uint32_t size = 124;
    uint32_t flags = kDDSD_CAPS | kDDSD_PIXELFORMAT;
    uint32_t height = 0, width = 0, pitch = 0, depth = 0, mipmapcount = 0;
    uint32_t reserved[11] = {};
    uint32_t pf_size = 32, pf_flags = 0, pf_fourcc = 0, pf_bitcount = 0, pf_rmask = 0, pf_gmask = 0, pf_bmask = 0,
             pf_amask = 0;
    uint32_t caps1 = kDDSCAPS_TEXTURE, caps2 = 0, caps3 = 0, caps4 = 0;
    uint32_t notused = 0;
    uint32_t dxgiFormat = 0, resourceDimension = 0, miscFlag = 0, arraySize = 0, reserved10 = 0;

    // DirectDrawSurface.cpp:574-609: the writer's version in reserved[9..10].
    DdsHeader()
    {
        reserved[9] = fourcc('N', 'V', 'T', 'T');
        reserved[10] = (2u << 16) | (1u << 8) | 2u;
    }
    // DirectDrawSurface.cpp:629-650.
    void setMipmapCount(uint32_t count)
    {
        if (count == 0 || count == 1) {
            flags &= ~kDDSD_MIPMAPCOUNT;
            mipmapcount = 1;
            caps1 = caps2 == 0 ? kDDSCAPS_TEXTURE : (kDDSCAPS_TEXTURE | kDDSCAPS_COMPLEX);
        } else {
            flags |= kDDSD_MIPMAPCOUNT;
            mipmapcount = count;
            caps1 |= kDDSCAPS_COMPLEX | kDDSCAPS_MIPMAP;
        }
    }
    void setFourCC(uint32_t code)
    {
        pf_flags = kDDPF_FOURCC;
        pf_fourcc = code;
        pf_bitcount = pf_rmask = pf_gmask = pf_bmask = pf_amask = 0;
    }
    void setDX10Format(uint32_t format)
    {
        pf_flags = kDDPF_FOURCC;
        pf_fourcc = fourcc('D', 'X', '1', '0');
        dxgiFormat = format;
    }
    bool hasDX10() const { return pf_fourcc == fourcc('D', 'X', '1', '0'); }
};
static_assert(sizeof(DdsHeader) == 148, "DDS header layout");

// nv::getTargetExtent (Surface.cpp:220-305), restated with its helpers
// (Surface.cpp:60-100; nvcore/Utils.h nextPowerOfTwo).
uint32_t next_pow2(uint32_t v)
{
    v--;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1;
}
uint32_t prev_pow2(uint32_t v) { return next_pow2(v + 1) / 2; }
uint32_t nearest_pow2(uint32_t v)
{
    const uint32_t np2 = next_pow2(v), pp2 = prev_pow2(v);
    return (np2 - v <= v - pp2) ? np2 : pp2;
}
uint32_t prev_mul4(uint32_t v) { return v & ~3u; }
uint32_t next_mul4(uint32_t v) { return (v + 3) & ~3u; }

void target_extent(int *width, int *height, int *depth, int maxExtent, RoundMode roundMode, TextureType type)
{
    int w = *width, h = *height, d = *depth;
    if (roundMode != RoundMode_None && maxExtent > 0) maxExtent = int(prev_pow2(uint32_t(maxExtent)));
    const int mx = std::max(std::max(w, h), d);
    if (maxExtent > 0 && mx > maxExtent) {
        w = std::max((w * maxExtent) / mx, 1);
        h = std::max((h * maxExtent) / mx, 1);
        d = std::max((d * maxExtent) / mx, 1);
    }
    if (type == TextureType_2D) {
        d = 1;
    } else if (type == TextureType_Cube) {
        w = h = (w + h) / 2;
        d = 1;
    }
    switch (roundMode) {
    case RoundMode_ToNextPowerOfTwo: w = int(next_pow2(w)); h = int(next_pow2(h)); d = int(next_pow2(d)); break;
    case RoundMode_ToNearestPowerOfTwo: w = int(nearest_pow2(w)); h = int(nearest_pow2(h)); d = int(nearest_pow2(d)); break;
    case RoundMode_ToPreviousPowerOfTwo: w = int(prev_pow2(w)); h = int(prev_pow2(h)); d = int(prev_pow2(d)); break;
    case RoundMode_ToNextMultipleOfFour: w = int(next_mul4(w)); h = int(next_mul4(h)); d = int(next_mul4(d)); break;
    // RoundMode_ToNearestMultipleOfFour: the reference tests ToNextMultipleOfFour
    // a second time in its place (Surface.cpp:273-284), so it leaves the extents.
    case RoundMode_ToPreviousMultipleOfFour: w = int(prev_mul4(w)); h = int(prev_mul4(h)); d = int(prev_mul4(d)); break;
    default: break;
    }
    *width = w; *height = h; *depth = d;
}

// Context.cpp:181-213 computeImageSize for the block formats.
int block_bytes(Format f)
{
    switch (f) {
    case Format_DXT1: case Format_DXT1a: case Format_DXT1n: case Format_BC4: case Format_CTX1: return 8;
    case Format_DXT3: case Format_DXT5: case Format_DXT5n: case Format_BC5: case Format_BC3_RGBM: return 16;
    default: return 0;
    }
}

// Filter width and parameters of a mip step of process() (Context.cpp:297-304;
// the KTX branch passes stretch and alpha the other way round, :427-429).
void process_filter(const InputOptions::Private &io, bool ktx, float *width, float params[2])
{
    if (io.mipmapFilter == MipmapFilter_Kaiser) {
        *width = io.kaiserWidth;
        params[0] = ktx ? io.kaiserStretch : io.kaiserAlpha;
        params[1] = ktx ? io.kaiserAlpha : io.kaiserStretch;
    } else {
        *width = io.mipmapFilter == MipmapFilter_Triangle ? 1.0f : 0.5f;
        params[0] = params[1] = 0.0f;
    }
}

}  // namespace

// Context.cpp:870-1030: the 64-byte KTX 1.1 header (KtxFile.h:113-131,
// KtxFile.cpp:17-34) for the block formats.
bool Compressor::Private::outputKtxHeader(TextureType type, int w, int h, int d, int arraySize, int mipmapCount,
                                          const CompressionOptions::Private &co, const OutputOptions::Private &oo) const
{
    struct KtxHeader {
        uint8_t identifier[12] = { 0xAB, 'K', 'T', 'X', ' ', '1', '1', 0xBB, '\r', '\n', 0x1A, '\n' };
        uint32_t endianness = 0x04030201, glType = 0, glTypeSize = 1, glFormat = 0, glInternalFormat = 0,
                 glBaseInternalFormat = 0, pixelWidth = 0, pixelHeight = 0, pixelDepth = 0, numberOfArrayElements = 0,
                 numberOfFaces = 1, numberOfMipmapLevels = 0, bytesOfKeyValueData = 0;
    } hd;
    static_assert(sizeof(KtxHeader) == 64, "KTX header layout");
    // GL enumerants of the compressed formats (KtxFile.h:58-109).
    constexpr uint32_t kRGB_DXT1 = 0x83F0, kSRGB_DXT1 = 0x8C4C, kRGBA_DXT1 = 0x83F1, kSRGB_ALPHA_DXT1 = 0x8C4D,
                       kRGBA_DXT3 = 0x83F2, kSRGB_ALPHA_DXT3 = 0x8C4E, kRGBA_DXT5 = 0x83F3, kSRGB_ALPHA_DXT5 = 0x8C4F,
                       kRED_RGTC1 = 0x8DBB, kRG_RGTC2 = 0x8DBD, kBaseRED = 0x1903, kBaseRG = 0x8227, kBaseRGB = 0x1907,
                       kBaseRGBA = 0x1908;
    if (type == TextureType_Cube) hd.numberOfFaces = 6;
    else if (type == TextureType_3D) hd.pixelDepth = uint32_t(d);
    else if (type == TextureType_Array) hd.numberOfArrayElements = uint32_t(arraySize);
    hd.pixelWidth = uint32_t(w);
    hd.pixelHeight = uint32_t(h);
    hd.numberOfMipmapLevels = uint32_t(mipmapCount);
    const Format f = co.format;
    if (f == Format_DXT1 || f == Format_DXT1n) {
        hd.glInternalFormat = oo.srgb ? kSRGB_DXT1 : kRGB_DXT1; hd.glBaseInternalFormat = kBaseRGB;
    } else if (f == Format_DXT1a) {
        hd.glInternalFormat = oo.srgb ? kSRGB_ALPHA_DXT1 : kRGBA_DXT1; hd.glBaseInternalFormat = kBaseRGBA;
    } else if (f == Format_DXT3) {
        hd.glInternalFormat = oo.srgb ? kSRGB_ALPHA_DXT3 : kRGBA_DXT3; hd.glBaseInternalFormat = kBaseRGBA;
    } else if (f == Format_DXT5 || f == Format_DXT5n || f == Format_BC3_RGBM) {
        hd.glInternalFormat = oo.srgb ? kSRGB_ALPHA_DXT5 : kRGBA_DXT5; hd.glBaseInternalFormat = kBaseRGBA;
    } else if (f == Format_BC4) {
        hd.glInternalFormat = kRED_RGTC1; hd.glBaseInternalFormat = kBaseRED;
    } else if (f == Format_BC5) {
        hd.glInternalFormat = kRG_RGTC2; hd.glBaseInternalFormat = kBaseRG;
    } else {
        oo.error(Error_UnsupportedOutputFormat);  // uncompressed, BC6/BC7, ETC: not on this path
        return false;
    }
    const bool ok = oo.writeData(&hd, 64);
    if (!ok) oo.error(Error_FileWrite);
    return ok;
}

// Context.cpp:604-868 for the DDS containers.
bool Compressor::Private::outputHeader(TextureType type, int w, int h, int d, int arraySize, int mipmapCount,
                                       bool isNormalMap, const CompressionOptions::Private &co,
                                       const OutputOptions::Private &oo) const
{
    if (w <= 0 || h <= 0 || d <= 0 || arraySize <= 0 || mipmapCount <= 0) {
        oo.error(Error_InvalidInput);
        return false;
    }
    if (!oo.outputHeader) return true;
    if (oo.container == Container_KTX) return outputKtxHeader(type, w, h, d, arraySize, mipmapCount, co, oo);
    if (oo.container != Container_DDS && oo.container != Container_DDS10) {
        oo.error(Error_UnsupportedOutputFormat);
        return false;
    }
    DdsHeader hd;
    hd.reserved[7] = fourcc('U', 'V', 'E', 'R');  // setUserVersion (DirectDrawSurface.cpp:809-813)
    hd.reserved[8] = uint32_t(oo.version);
    // setTexture2D / setTextureCube / setTexture3D / setTextureArray (DirectDrawSurface.cpp:652-682).
    if (type == TextureType_2D) {
        hd.resourceDimension = kDimTexture2D; hd.miscFlag = 0; hd.arraySize = 1;
    } else if (type == TextureType_Cube) {
        hd.caps1 |= kDDSCAPS_COMPLEX;
        hd.caps2 = kDDSCAPS2_CUBEMAP | kDDSCAPS2_CUBEMAP_ALL_FACES;
        hd.resourceDimension = kDimTexture2D; hd.miscFlag = kMiscTextureCube; hd.arraySize = 1;
    } else if (type == TextureType_3D) {
        hd.caps2 = kDDSCAPS2_VOLUME;
        hd.resourceDimension = kDimTexture3D; hd.miscFlag = 0; hd.arraySize = 1;
        hd.flags |= kDDSD_DEPTH;
        hd.depth = uint32_t(d);
    } else {
        hd.resourceDimension = kDimTexture2D; hd.arraySize = uint32_t(arraySize);
    }
    hd.flags |= kDDSD_WIDTH;
    hd.width = uint32_t(w);
    hd.flags |= kDDSD_HEIGHT;
    hd.height = uint32_t(h);
    hd.setMipmapCount(uint32_t(mipmapCount));

    bool supported = true;
    const Format f = co.format;
    if (oo.container == Container_DDS10) {
        if (f == Format_DXT1 || f == Format_DXT1a || f == Format_DXT1n) {
            hd.setDX10Format(oo.srgb ? kDXGI_BC1_SRGB : kDXGI_BC1);
            if (f == Format_DXT1a) hd.pf_flags |= kDDPF_ALPHAPIXELS;
            if (isNormalMap) hd.pf_flags |= kDDPF_NORMAL;
        } else if (f == Format_DXT3) {
            hd.setDX10Format(oo.srgb ? kDXGI_BC2_SRGB : kDXGI_BC2);
        } else if (f == Format_DXT5 || f == Format_BC3_RGBM) {
            hd.setDX10Format(oo.srgb ? kDXGI_BC3_SRGB : kDXGI_BC3);
        } else if (f == Format_DXT5n) {
            hd.setDX10Format(kDXGI_BC3);
            if (isNormalMap) hd.pf_flags |= kDDPF_NORMAL;
        } else if (f == Format_BC4) {
            hd.setDX10Format(kDXGI_BC4);
        } else if (f == Format_BC5) {
            hd.setDX10Format(kDXGI_BC5);
            if (isNormalMap) hd.pf_flags |= kDDPF_NORMAL;
        } else {
            supported = false;
        }
    } else {
        const int bb = block_bytes(f);
        // setLinearSize (DirectDrawSurface.cpp:684-689): the top level's bytes.
        hd.flags &= ~kDDSD_PITCH;
        hd.flags |= kDDSD_LINEARSIZE;
        hd.pitch = uint32_t(((w + 3) / 4) * ((h + 3) / 4) * bb * d);
        if (f == Format_DXT1 || f == Format_DXT1a || f == Format_DXT1n) {
            hd.setFourCC(fourcc('D', 'X', 'T', '1'));
            if (isNormalMap) hd.pf_flags |= kDDPF_NORMAL;
        } else if (f == Format_DXT3) {
            hd.setFourCC(fourcc('D', 'X', 'T', '3'));
        } else if (f == Format_DXT5 || f == Format_BC3_RGBM) {
            hd.setFourCC(fourcc('D', 'X', 'T', '5'));
        } else if (f == Format_DXT5n) {
            hd.setFourCC(fourcc('D', 'X', 'T', '5'));
            if (isNormalMap) {
                hd.pf_flags |= kDDPF_NORMAL;
                hd.pf_bitcount = fourcc('A', '2', 'D', '5');  // setSwizzleCode (DirectDrawSurface.cpp:724-728)
            }
        } else if (f == Format_BC4) {
            hd.setFourCC(fourcc('A', 'T', 'I', '1'));
        } else if (f == Format_BC5) {
            hd.setFourCC(fourcc('A', 'T', 'I', '2'));
            if (isNormalMap) {
                hd.pf_flags |= kDDPF_NORMAL;
                hd.pf_bitcount = fourcc('A', '2', 'X', 'Y');
            }
        } else if (f == Format_CTX1) {
            hd.setFourCC(fourcc('C', 'T', 'X', '1'));
            if (isNormalMap) hd.pf_flags |= kDDPF_NORMAL;
        } else {
            supported = false;  // uncompressed, BC6, BC7 and the rest: not on this path
        }
        if (oo.srgb) hd.pf_flags |= kDDPF_SRGB;
    }
    if (!supported) {
        oo.error(Error_UnsupportedOutputFormat);
        return false;
    }
    const int size = hd.hasDX10() ? 148 : 128;
    const bool ok = oo.writeData(&hd, size);
    if (!ok) oo.error(Error_FileWrite);
    return ok;
}

namespace {

// Surface::setImage(format, w, h, 1, data) into the device planes dst
// (Surface.cpp:728-815), followed by Surface::toLinear(gamma) when linearize
// (Surface.cpp:1470-1477).  staging: 4 w h bytes of device memory for 8-bit
// input.
ntx_status load_level(ntx_hip_context *ctx, InputFormat format, int w, int h, const std::vector<uint8_t> &data,
                      bool linearize, float gamma, float *dst, void *staging)
{
    const size_t n = size_t(w) * size_t(h);
    const float g = linearize ? gamma : 1.0f;
    if (format == InputFormat_BGRA_8UB) {
        void *p = staging;  // 4 n bytes
        ntx_status s = ntx_hip_upload(ctx, p, data.data(), 4 * n);
        if (s == NTX_OK) s = ntx_hip_unpack_bgra8_gamma_device(ctx, static_cast<const uint8_t *>(p), w, h, g, dst);
        return s;
    }
    std::vector<float> planar(4 * n, 0.0f);
    if (format == InputFormat_RGBA_16F) {
        const uint16_t *src = reinterpret_cast<const uint16_t *>(data.data());
        for (size_t i = 0; i < n; i++)
            for (int c = 0; c < 4; c++) planar[c * n + i] = half_to_float(src[4 * i + c]);
    } else if (format == InputFormat_RGBA_32F) {
        const float *src = reinterpret_cast<const float *>(data.data());
        for (size_t i = 0; i < n; i++)
            for (int c = 0; c < 4; c++) planar[c * n + i] = src[4 * i + c];
    } else {  // InputFormat_R_32F
        const float *src = reinterpret_cast<const float *>(data.data());
        for (size_t i = 0; i < n; i++) planar[i] = src[i];
    }
    ntx_status s = ntx_hip_upload(ctx, dst, planar.data(), planar.size() * sizeof(float));
    if (s == NTX_OK && !unit_gamma(g)) s = ntx_hip_gamma_device(ctx, 0, g, dst, w, h, dst);
    return s;
}

}  // namespace

// Context.cpp:217-344 (the DDS branch).  Per face, level 0 comes from the
// input (to linear unless a normal map) and every further level is either the
// supplied image (to linear) or buildNextMipmap of the one before, a normal
// map's renormalised; each level is compressed from its toGamma copy.  Here
// every face's levels are built on the device first, the gamma copies of all
// of them made in one launch, and all images compressed together by the
// image scheduler (ntx_hip_compress_images_device); they are then delivered
// face by face, level by level, with the reference's begin/write/end order.
bool Compressor::Private::process(const InputOptions::Private &io, const CompressionOptions::Private &co,
                                  const OutputOptions::Private &oo) const
{
    if (!oo.hasValidOutputHandler()) {
        oo.error(Error_FileOpen);
        return false;
    }
    const int faceCount = io.faceCount;
    if (faceCount <= 0) {
        oo.error(Error_InvalidInput);
        return false;
    }
    int width = io.width, height = io.height, depth = io.depth;
    const int arraySize = io.textureType == TextureType_Array ? faceCount : 1;
    target_extent(&width, &height, &depth, io.maxExtent, io.roundMode, io.textureType);
    const bool canUseSourceImages = io.width == width && io.height == height && io.depth == depth;
    int mipmapCount = 1;
    if (io.generateMipmaps) {
        mipmapCount = count_mipmaps3(width, height, depth);
        if (io.maxLevel > 0) mipmapCount = std::min(mipmapCount, io.maxLevel);
    }
    // Not on this path: resizing, volume textures, height-map conversion, and
    // the quantize step (Context.cpp:519-541: colour dithering for BC1-BC3,
    // binary alpha for any format).
    const bool quantize = (co.enableColorDithering && co.format >= Format_BC1 && co.format <= Format_BC3) ||
                          (!co.enableAlphaDithering && co.binaryAlpha);
    if (!canUseSourceImages || io.textureType == TextureType_3D || depth != 1 || io.convertToNormalMap || quantize) {
        oo.error(Error_UnsupportedFeature);
        return false;
    }
    for (int f = 0; f < faceCount; f++)
        if (io.images[size_t(f)].empty()) {
            oo.error(Error_InvalidInput);
            return false;
        }
    if (!outputHeader(io.textureType, width, height, depth, arraySize, mipmapCount, io.isNormalMap, co, oo)) return false;

    // The images in delivery order: face-major, then level.
    struct Image { int face, mip, w, h; };
    std::vector<Image> images;
    for (int f = 0; f < faceCount; f++) {
        int w = width, h = height;
        for (int m = 0; m < mipmapCount; m++) {
            images.push_back({ f, m, w, h });
            w = w / 2 > 1 ? w / 2 : 1;
            h = h / 2 > 1 ? h / 2 : 1;
        }
    }
    const int fmt = int(co.format);
    auto deliver_error = [&](Error e) {
        for (const Image &im : images) {
            oo.beginImage(int(ntx_hip_compressed_size(fmt, im.w, im.h)), im.w, im.h, 1, im.face, im.mip);
            oo.error(e);
            oo.endImage();
        }
        return true;
    };
    if (!hipEnabled || !backend()) return deliver_error(Error_CudaError);
    CompressorInterface *probe = chooseCompressor(co, true);
    if (!probe) return deliver_error(Error_UnsupportedFeature);
    delete probe;

    const size_t n = images.size();
    const bool normal = io.isNormalMap;
    const bool regamma = !normal && !unit_gamma(io.outputGamma);
    const AlphaMode alphaMode = io.alphaMode;
    const bool ktx = oo.container == Container_KTX;
    float fw, fp[2];
    process_filter(io, ktx, &fw, fp);
    // One device arena: level planes, gamma copies, the largest 8-bit staging
    // image, compressed output (one block for a single download); offsets
    // rounded to 256 bytes.
    std::vector<float *> lin(n, nullptr), gam(n, nullptr);
    std::vector<void *> out(n, nullptr);
    std::vector<int> ws(n), hs(n);
    std::vector<size_t> outOff(n);
    auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t planeBytes = 0, outBytes = 0, stageBytes = 0;
    for (size_t i = 0; i < n; i++) {
        ws[i] = images[i].w;
        hs[i] = images[i].h;
        planeBytes += up(4 * size_t(ws[i]) * size_t(hs[i]) * sizeof(float));
        outOff[i] = outBytes;
        outBytes += ntx_hip_compressed_size(fmt, ws[i], hs[i]);
        stageBytes = std::max(stageBytes, 4 * size_t(ws[i]) * size_t(hs[i]));
    }
    const size_t total = planeBytes * (regamma ? 2 : 1) + up(stageBytes) + up(outBytes);
    std::vector<uint8_t> bytes(outBytes);
    ntx_status s = NTX_OK;
    {
        std::lock_guard<std::mutex> g(backend_lock());
        ntx_hip_context *ctx = backend();
        if (arenaBytes < total) {
            if (arena) ntx_hip_free(ctx, arena);
            arena = nullptr;
            arenaBytes = 0;
            s = ntx_hip_malloc(ctx, total, &arena);
            if (s == NTX_OK) arenaBytes = total;
        }
        uint8_t *base = static_cast<uint8_t *>(arena);
        size_t off = 0;
        for (size_t i = 0; s == NTX_OK && i < n; i++) {
            lin[i] = reinterpret_cast<float *>(base + off);
            off += up(4 * size_t(ws[i]) * size_t(hs[i]) * sizeof(float));
        }
        for (size_t i = 0; s == NTX_OK && i < n; i++) {
            if (regamma) {
                gam[i] = reinterpret_cast<float *>(base + off);
                off += up(4 * size_t(ws[i]) * size_t(hs[i]) * sizeof(float));
            } else {
                gam[i] = lin[i];
            }
        }
        void *staging = base + off;
        off += up(stageBytes);
        for (size_t i = 0; i < n; i++) out[i] = base + off + outOff[i];
        size_t i = 0;
        for (int f = 0; s == NTX_OK && f < faceCount; f++) {
            bool canUseFace = canUseSourceImages;
            for (int m = 0; s == NTX_OK && m < mipmapCount; m++, i++) {
                if (m == 0) {
                    s = load_level(ctx, io.inputFormat, ws[i], hs[i], io.images[size_t(f)], !normal, io.inputGamma, lin[i],
                                   staging);
                    continue;
                }
                const std::vector<uint8_t> &src = io.images[size_t(m) * size_t(faceCount) + size_t(f)];
                bool useSource = false;
                if (ktx) {
                    // Context.cpp:411-418: the chain is never marked broken, so
                    // every supplied level is used.
                    useSource = !src.empty();
                } else if (canUseFace) {
                    if (src.empty()) canUseFace = false;  // a missing level: generate this and the rest
                    else useSource = true;
                }
                if (useSource)
                    s = load_level(ctx, io.inputFormat, ws[i], hs[i], src, !normal, io.inputGamma, lin[i], staging);
                else
                    s = ntx_hip_downsample_device(ctx, int(io.mipmapFilter), fw, fp, int(io.wrapMode), int(alphaMode),
                                                  lin[i - 1], ws[i - 1], hs[i - 1], lin[i]);
                // DDS: expand, normalise, pack (Context.cpp:329-336); KTX:
                // normalise only (:446-449).
                if (s == NTX_OK && normal && io.normalizeMipmaps)
                    s = ktx ? ntx_hip_normalize_vectors_device(ctx, lin[i], ws[i], hs[i])
                            : ntx_hip_normalize_normals_device(ctx, lin[i], ws[i], hs[i]);
            }
        }
        if (s == NTX_OK && regamma)
            s = ntx_hip_gamma_levels_device(ctx, 1, io.outputGamma, int(n), const_cast<const float *const *>(lin.data()),
                                            ws.data(), hs.data(), gam.data());
        if (s == NTX_OK && learned_bc1(co, alphaMode)) {
            std::vector<int> lv(n);
            for (size_t k = 0; k < n; k++) lv[k] = int(k % size_t(mipmapCount));  // images are face-major
            s = compress_learned(ctx, co, int(n), const_cast<const float *const *>(gam.data()), ws.data(), hs.data(),
                                 lv.data(), out.data());
        } else if (s == NTX_OK) {
            s = ntx_hip_compress_images_device(ctx, fmt, int(co.quality), int(alphaMode), co.colorWeight, int(n),
                                               const_cast<const float *const *>(gam.data()), ws.data(), hs.data(),
                                               out.data());
        }
        if (s == NTX_OK) s = ntx_hip_download(ctx, bytes.data(), out[0], outBytes);
        (void)ntx_hip_synchronize(ctx);
    }
    if (s != NTX_OK) {
        report("process", s);
        return deliver_error(s == NTX_ERR_INVALID_ARGUMENT ? Error_InvalidInput
                             : s == NTX_ERR_UNSUPPORTED   ? Error_UnsupportedFeature
                                                          : Error_CudaError);
    }
    auto deliver = [&](size_t j) {
        const int size = int(ntx_hip_compressed_size(fmt, ws[j], hs[j]));
        oo.beginImage(size, ws[j], hs[j], 1, images[j].face, images[j].mip);
        oo.writeData(bytes.data() + outOff[j], size);
        oo.endImage();
    };
    if (!ktx) {
        for (size_t j = 0; j < n; j++) deliver(j);
        return true;
    }
    // Context.cpp:345-474: per level, the byte count of all its faces
    // (estimateSize x faceCount, as a uint32, outside any begin/endImage),
    // then that level of every face; after each level past the first, the
    // padding to a multiple of four bytes.
    static const uint8_t padding[3] = { 0, 0, 0 };
    for (int m = 0; m < mipmapCount; m++) {
        const uint32_t imageSize = uint32_t(ntx_hip_compressed_size(fmt, ws[size_t(m)], hs[size_t(m)])) * uint32_t(faceCount);
        oo.writeData(&imageSize, int(sizeof imageSize));
        for (int f = 0; f < faceCount; f++) deliver(size_t(f) * size_t(mipmapCount) + size_t(m));
        if (m > 0) {
            const int mipPadding = 3 - int((imageSize + 3) % 4);
            if (mipPadding != 0) oo.writeData(padding, mipPadding);
        }
    }
    return true;
}

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

bool Compressor::process(const InputOptions &io, const CompressionOptions &co, const OutputOptions &oo) const
{
    return m.process(io.m, co.m, oo.m);
}

// Context.cpp:122-137.
int Compressor::estimateSize(const InputOptions &io, const CompressionOptions &co) const
{
    int w = io.m.width, h = io.m.height, d = io.m.depth;
    if (w <= 0 || h <= 0 || d <= 0) return 0;
    target_extent(&w, &h, &d, io.m.maxExtent, io.m.roundMode, io.m.textureType);
    int mipmapCount = 1;
    if (io.m.generateMipmaps) {
        mipmapCount = count_mipmaps3(w, h, d);
        if (io.m.maxLevel > 0) mipmapCount = std::min(mipmapCount, io.m.maxLevel);
    }
    return io.m.faceCount * estimateSize(w, h, d, mipmapCount, co);
}

// Context.cpp:141-144.
bool Compressor::outputHeader(const Surface &img, int mipmapCount, const CompressionOptions &co,
                              const OutputOptions &oo) const
{
    return m.outputHeader(img.type(), img.width(), img.height(), img.depth(), 1, mipmapCount, img.isNormalMap(), co.m,
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
