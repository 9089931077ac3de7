// nvtt_wrapper.cpp -- the C interface (nvtt_wrapper.h) over the nvtt.h front
// end.  Each OutputOptions handle made here owns a proxy that forwards the
// handler interfaces to the C callbacks (nvtt_wrapper.cpp:32-90 keeps one in
// OutputOptions::Private; here a side table keyed by the handle).
#include "nvtt_wrapper.h"

#include <map>
#include <mutex>

#include "nvtt.h"

namespace {

struct CallbackProxy : nvtt::OutputHandler, nvtt::ErrorHandler {
    nvttBeginImageHandler begin = nullptr;
    nvttOutputHandler write = nullptr;
    nvttEndImageHandler end = nullptr;
    nvttErrorHandler err = nullptr;
    void beginImage(int size, int w, int h, int d, int face, int mip) override
    {
        if (begin) begin(size, w, h, d, face, mip);
    }
    bool writeData(const void *data, int size) override { return write ? write(data, size) : false; }
    void endImage() override
    {
        if (end) end();
    }
    void error(nvtt::Error e) override
    {
        if (err) err(NvttError(e));
    }
};

std::mutex &proxy_lock()
{
    static std::mutex m;
    return m;
}
std::map<const nvtt::OutputOptions *, CallbackProxy *> &proxies()
{
    static std::map<const nvtt::OutputOptions *, CallbackProxy *> p;
    return p;
}
CallbackProxy *proxy_of(const nvtt::OutputOptions *oo)
{
    std::lock_guard<std::mutex> g(proxy_lock());
    CallbackProxy *&p = proxies()[oo];
    if (!p) p = new CallbackProxy;
    return p;
}

}  // namespace

extern "C" {

NvttInputOptions *nvttCreateInputOptions(void) { return new nvtt::InputOptions(); }
void nvttDestroyInputOptions(NvttInputOptions *io) { delete io; }
void nvttSetInputOptionsTextureLayout(NvttInputOptions *io, NvttTextureType type, int w, int h, int d, int arraySize)
{
    io->setTextureLayout(nvtt::TextureType(type), w, h, d, arraySize);
}
void nvttResetInputOptionsTextureLayout(NvttInputOptions *io) { io->resetTextureLayout(); }
NvttBoolean nvttSetInputOptionsMipmapData(NvttInputOptions *io, const void *data, int w, int h, int d, int face, int mipmap)
{
    return io->setMipmapData(data, w, h, d, face, mipmap) ? NVTT_True : NVTT_False;
}
void nvttSetInputOptionsFormat(NvttInputOptions *io, NvttInputFormat format) { io->setFormat(nvtt::InputFormat(format)); }
void nvttSetInputOptionsAlphaMode(NvttInputOptions *io, NvttAlphaMode mode) { io->setAlphaMode(nvtt::AlphaMode(mode)); }
void nvttSetInputOptionsGamma(NvttInputOptions *io, float inputGamma, float outputGamma)
{
    io->setGamma(inputGamma, outputGamma);
}
void nvttSetInputOptionsWrapMode(NvttInputOptions *io, NvttWrapMode mode) { io->setWrapMode(nvtt::WrapMode(mode)); }
void nvttSetInputOptionsMipmapFilter(NvttInputOptions *io, NvttMipmapFilter filter)
{
    io->setMipmapFilter(nvtt::MipmapFilter(filter));
}
void nvttSetInputOptionsMipmapGeneration(NvttInputOptions *io, NvttBoolean enabled, int maxLevel)
{
    io->setMipmapGeneration(enabled != NVTT_False, maxLevel);
}
void nvttSetInputOptionsKaiserParameters(NvttInputOptions *io, float width, float alpha, float stretch)
{
    io->setKaiserParameters(width, alpha, stretch);
}
void nvttSetInputOptionsNormalMap(NvttInputOptions *io, NvttBoolean b) { io->setNormalMap(b != NVTT_False); }
void nvttSetInputOptionsConvertToNormalMap(NvttInputOptions *io, NvttBoolean convert)
{
    io->setConvertToNormalMap(convert != NVTT_False);
}
void nvttSetInputOptionsHeightEvaluation(NvttInputOptions *io, float r, float g, float b, float a)
{
    io->setHeightEvaluation(r, g, b, a);
}
void nvttSetInputOptionsNormalFilter(NvttInputOptions *io, float sm, float medium, float big, float large)
{
    io->setNormalFilter(sm, medium, big, large);
}
void nvttSetInputOptionsNormalizeMipmaps(NvttInputOptions *io, NvttBoolean b) { io->setNormalizeMipmaps(b != NVTT_False); }
void nvttSetInputOptionsMaxExtents(NvttInputOptions *io, int dim) { io->setMaxExtents(dim); }
void nvttSetInputOptionsRoundMode(NvttInputOptions *io, NvttRoundMode mode) { io->setRoundMode(nvtt::RoundMode(mode)); }

NvttCompressionOptions *nvttCreateCompressionOptions(void) { return new nvtt::CompressionOptions(); }
void nvttDestroyCompressionOptions(NvttCompressionOptions *co) { delete co; }
void nvttSetCompressionOptionsFormat(NvttCompressionOptions *co, NvttFormat format) { co->setFormat(nvtt::Format(format)); }
void nvttSetCompressionOptionsQuality(NvttCompressionOptions *co, NvttQuality quality)
{
    co->setQuality(nvtt::Quality(quality));
}
void nvttSetCompressionOptionsColorWeights(NvttCompressionOptions *co, float r, float g, float b, float a)
{
    co->setColorWeights(r, g, b, a);
}
void nvttSetCompressionOptionsPixelFormat(NvttCompressionOptions *co, unsigned int bitcount, unsigned int rmask,
                                          unsigned int gmask, unsigned int bmask, unsigned int amask)
{
    co->setPixelFormat(bitcount, rmask, gmask, bmask, amask);
}
void nvttSetCompressionOptionsQuantization(NvttCompressionOptions *co, NvttBoolean colorDithering, NvttBoolean alphaDithering,
                                           NvttBoolean binaryAlpha, int alphaThreshold)
{
    co->setQuantization(colorDithering != NVTT_False, alphaDithering != NVTT_False, binaryAlpha != NVTT_False,
                        alphaThreshold);
}

NvttOutputOptions *nvttCreateOutputOptions(void)
{
    nvtt::OutputOptions *oo = new nvtt::OutputOptions();
    oo->setErrorHandler(proxy_of(oo));
    return oo;
}
void nvttDestroyOutputOptions(NvttOutputOptions *oo)
{
    CallbackProxy *p = nullptr;
    {
        std::lock_guard<std::mutex> g(proxy_lock());
        auto it = proxies().find(oo);
        if (it != proxies().end()) {
            p = it->second;
            proxies().erase(it);
        }
    }
    delete oo;
    delete p;
}
void nvttSetOutputOptionsFileName(NvttOutputOptions *oo, const char *fileName) { oo->setFileName(fileName); }
void nvttSetOutputOptionsOutputHeader(NvttOutputOptions *oo, NvttBoolean b) { oo->setOutputHeader(b != NVTT_False); }
void nvttSetOutputOptionsContainer(NvttOutputOptions *oo, NvttContainer c) { oo->setContainer(nvtt::Container(c)); }
void nvttSetOutputOptionsSrgbFlag(NvttOutputOptions *oo, NvttBoolean b) { oo->setSrgbFlag(b != NVTT_False); }
void nvttSetOutputOptionsErrorHandler(NvttOutputOptions *oo, nvttErrorHandler errorHandler)
{
    proxy_of(oo)->err = errorHandler;
}
// nvtt_wrapper.cpp:273-289: no callbacks at all detaches the handler.
void nvttSetOutputOptionsOutputHandler(NvttOutputOptions *oo, nvttBeginImageHandler begin, nvttOutputHandler write,
                                       nvttEndImageHandler end)
{
    CallbackProxy *p = proxy_of(oo);
    p->begin = begin;
    p->write = write;
    p->end = end;
    oo->setOutputHandler((begin || write || end) ? p : nullptr);
}

NvttCompressor *nvttCreateCompressor(void) { return new nvtt::Compressor(); }
void nvttDestroyCompressor(NvttCompressor *c) { delete c; }
void nvttEnableCudaAcceleration(NvttCompressor *c, NvttBoolean b) { c->enableCudaAcceleration(b != NVTT_False); }
NvttBoolean nvttIsCudaAccelerationEnabled(const NvttCompressor *c)
{
    return c->isCudaAccelerationEnabled() ? NVTT_True : NVTT_False;
}
NvttBoolean nvttCompress(const NvttCompressor *c, const NvttInputOptions *io, const NvttCompressionOptions *co,
                         const NvttOutputOptions *oo)
{
    return c->process(*io, *co, *oo) ? NVTT_True : NVTT_False;
}
int nvttEstimateSize(const NvttCompressor *c, const NvttInputOptions *io, const NvttCompressionOptions *co)
{
    return c->estimateSize(*io, *co);
}

const char *nvttErrorString(NvttError e) { return nvtt::errorString(nvtt::Error(e)); }
unsigned int nvttVersion(void) { return nvtt::version(); }

}  // extern "C"
