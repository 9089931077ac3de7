// ref_harness_nv.cpp -- exposes the reference's own nvimage / nvtt CPU code on
// the hot path through a C ABI, so the tests can pin the oracle against it and
// bench.py can time it.  TEST INFRASTRUCTURE ONLY.
//
// Nothing from the reference is copied here.  oracle/Makefile compiles the
// reference sources where they lie under REF (nvtt block compressors,
// QuickCompress / OptimalCompress / SingleColorLookup, nvimage ColorBlock /
// BlockDXT / FloatImage / Filter, nvmath, nvcore, nvthread, nvsquish, icbc)
// with the reference's own platform header REF/project/vc2017/nvconfig.h and
// the pinned settings ICBC_SIMD=2, SQUISH_USE_SSE=0.  This file only builds
// the reference's objects, calls their methods and copies bytes in and out.
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <thread>
#include <vector>

#include "nvtt/BlockCompressor.h"
#include "nvtt/CompressorDX9.h"
#include "nvtt/CompressorDX10.h"
#include "nvtt/CompressionOptions.h"
#include "nvtt/OutputOptions.h"
#include "nvtt/TaskDispatcher.h"
#include "nvtt/QuickCompressDXT.h"
#include "nvtt/OptimalCompressDXT.h"
#include "nvimage/ColorBlock.h"
#include "nvimage/BlockDXT.h"
#include "nvimage/FloatImage.h"
#include "nvimage/Filter.h"
#include "nvmath/Gamma.h"
#include "nvmath/Vector.inl"
#include "nvthread/nvthread.h"

using namespace nv;

namespace {

// Collects what ColorBlockCompressor / FloatColorCompressor hand to
// OutputOptions::Private::writeData.
struct Sink : public nvtt::OutputHandler {
    uint8_t *dst = nullptr;
    size_t cap = 0, used = 0;
    void beginImage(int, int, int, int, int, int) override {}
    bool writeData(const void *data, int size) override
    {
        const size_t n = used + size_t(size) <= cap ? size_t(size) : cap - used;
        memcpy(dst + used, data, n);
        used += n;
        return true;
    }
    void endImage() override {}
};

// CompressionOptions defaults (CompressionOptions.cpp:48-75) for the fields
// the block compressors read: format, quality, colorWeight.
void options(nvtt::CompressionOptions::Private &o, nvtt::Format f, int quality, const float *cw)
{
    o.format = f;
    o.quality = nvtt::Quality(quality);
    o.colorWeight = cw ? Vector4(cw[0], cw[1], cw[2], cw[3]) : Vector4(1.0f);
    o.rgbmThreshold = 0.15f;
    o.bitcount = 32;
    o.rmask = 0xFF0000; o.gmask = 0xFF00; o.bmask = 0xFF; o.amask = 0xFF000000;
    o.rsize = o.gsize = o.bsize = o.asize = 8;
    o.pixelType = nvtt::PixelType_UnsignedNorm;
    o.pitchAlignment = 1;
    o.enableColorDithering = o.enableAlphaDithering = o.binaryAlpha = false;
    o.alphaThreshold = 127;
    o.decoder = nvtt::Decoder_D3D10;
}

// Compressor::Private::chooseCpuCompressor (Context.cpp:1038-1115) for the
// formats on the path: kind 0 BC1, 1 BC1a, 2 BC3, 3 BC3n, 4 BC4, 5 BC5, with
// the quality picking the Fast*/Production* variants as it does there.
CompressorInterface *make_compressor(int kind, int quality)
{
    const bool fastest = quality == nvtt::Quality_Fastest;
    const bool production = quality >= nvtt::Quality_Production;
    switch (kind) {
    case 0: return new CompressorDXT1;
    case 1: return fastest ? (CompressorInterface *)new FastCompressorDXT1a : new CompressorDXT1a;
    case 2: return fastest ? (CompressorInterface *)new FastCompressorDXT5 : new CompressorDXT5;
    case 3: return fastest ? (CompressorInterface *)new FastCompressorDXT5n : new CompressorDXT5n;
    case 4: return production ? (CompressorInterface *)new ProductionCompressorBC4 : new FastCompressorBC4;
    case 5: return production ? (CompressorInterface *)new ProductionCompressorBC5 : new FastCompressorBC5;
    }
    return nullptr;
}

nvtt::Format kind_format(int kind)
{
    static const nvtt::Format f[] = { nvtt::Format_BC1, nvtt::Format_BC1a, nvtt::Format_BC3, nvtt::Format_BC3n,
                                      nvtt::Format_BC4, nvtt::Format_BC5 };
    return f[kind];
}

void to_colorblock(const uint8_t *bgra64, ColorBlock &cb)
{
    for (int i = 0; i < 16; i++) {
        Color32 &c = cb.color(i);
        c.b = bgra64[4 * i]; c.g = bgra64[4 * i + 1]; c.r = bgra64[4 * i + 2]; c.a = bgra64[4 * i + 3];
    }
}

FloatImage *planar_image(int W, int H, const float *src)
{
    FloatImage *img = new FloatImage();
    img->allocate(4, W, H);
    for (int c = 0; c < 4; c++) memcpy(img->channel(c), src + size_t(c) * W * H, sizeof(float) * size_t(W) * H);
    return img;
}

void copy_out(const FloatImage *img, float *dst)
{
    const size_t n = size_t(img->width()) * img->height();
    for (int c = 0; c < 4; c++) memcpy(dst + c * n, img->channel(c), sizeof(float) * n);
}

Filter *make_filter(int filter, float width, const float *params)
{
    if (filter == 0) return new BoxFilter(width);
    if (filter == 1) return new TriangleFilter(width);
    KaiserFilter *k = new KaiserFilter(width);
    if (params) k->setParameters(params[0], params[1]);
    return k;
}

}  // namespace

extern "C" {

// ColorBlock::init(w, h, const float *, x, y) (ColorBlock.cpp:80-110): the
// 8-bit gather of the BC3/BC4/BC5 compressors, out as 16 BGRA texels.
void ntxref_colorblock_init(const float *planar, int w, int h, int x, int y, uint8_t *bgra64)
{
    ColorBlock cb;
    cb.init(uint(w), uint(h), planar, uint(x), uint(y));
    for (int i = 0; i < 16; i++) {
        const Color32 c = cb.color(i);
        bgra64[4 * i] = c.b; bgra64[4 * i + 1] = c.g; bgra64[4 * i + 2] = c.r; bgra64[4 * i + 3] = c.a;
    }
}

// QuickCompress::compressDXT5A(AlphaBlock4x4) (QuickCompressDXT.cpp:772-821):
// one BC4 channel block of 16 values.
void ntxref_bc4_channel(const uint8_t *v16, uint8_t *out8)
{
    AlphaBlock4x4 a;
    for (int i = 0; i < 16; i++) { a.alpha[i] = v16[i]; a.weights[i] = 1.0f; }
    AlphaBlockDXT5 blk;
    blk.u = 0;
    QuickCompress::compressDXT5A(a, &blk);
    memcpy(out8, &blk.u, 8);
}

// OptimalCompress::compressDXT5A(AlphaBlock4x4) (OptimalCompressDXT.cpp:512-607)
// with its first error seed (line 546, read from the output block before it
// is written) taken from the endpoints (max, min) that the search falls back
// to; the reference leaves that seed to whatever the output memory held.
void ntxref_bc4_channel_optimal(const uint8_t *v16, uint8_t *out8)
{
    AlphaBlock4x4 a;
    uint8_t lo = 255, hi = 0;
    for (int i = 0; i < 16; i++) {
        a.alpha[i] = v16[i]; a.weights[i] = 1.0f;
        lo = v16[i] < lo ? v16[i] : lo; hi = v16[i] > hi ? v16[i] : hi;
    }
    AlphaBlockDXT5 blk;
    blk.u = 0;
    blk.alpha0 = hi;
    blk.alpha1 = lo;
    OptimalCompress::compressDXT5A(a, &blk);
    memcpy(out8, &blk.u, 8);
}

// n channel blocks (16 values each) through ntxref_bc4_channel (optimal = 0)
// or ntxref_bc4_channel_optimal (optimal = 1).
void ntxref_bc4_channels(const uint8_t *v, size_t n, int optimal, uint8_t *out)
{
    for (size_t i = 0; i < n; i++) {
        if (optimal) ntxref_bc4_channel_optimal(v + 16 * i, out + 8 * i);
        else ntxref_bc4_channel(v + 16 * i, out + 8 * i);
    }
}

// One block through a reference block compressor's compressBlock (kind as in
// make_compressor, 1..5: the 8-bit ColorBlock compressors), on 16 BGRA texels.  Production BC4/BC5 and Highest
// BC3/BC3n use OptimalCompress::compressDXT5A, whose first error seed reads
// the output block before writing it: the output is pre-filled the way
// ntxref_bc4_channel_optimal fills it, channel by channel.
void ntxref_block(int kind, int quality, int alpha_mode, const float *cw, const uint8_t *bgra64, uint8_t *out)
{
    if (kind < 1 || kind > 5) return;
    ColorBlock cb;
    to_colorblock(bgra64, cb);
    nvtt::CompressionOptions::Private o;
    options(o, kind_format(kind), quality, cw);
    ColorBlockCompressor *c = static_cast<ColorBlockCompressor *>(make_compressor(kind, quality));
    const int bs = c->blockSize();
    memset(out, 0, bs);
    auto seed = [&](int channel, uint8_t *dst) {  // channel of the BGRA bytes
        uint8_t lo = 255, hi = 0;
        for (int i = 0; i < 16; i++) {
            const uint8_t v = bgra64[4 * i + channel];
            lo = v < lo ? v : lo; hi = v > hi ? v : hi;
        }
        dst[0] = hi; dst[1] = lo;
    };
    if ((kind == 4 || kind == 5) && quality >= nvtt::Quality_Production) {
        seed(2, out);
        if (kind == 5) seed(1, out + 8);
    } else if (kind == 2 && quality == nvtt::Quality_Highest) {
        seed(3, out);
    } else if (kind == 3 && quality == nvtt::Quality_Highest) {
        seed(2, out);  // CompressorDXT5n moves red into the alpha block
    }
    c->compressBlock(cb, nvtt::AlphaMode(alpha_mode), o, out);
    delete c;
}

// A whole planar RGBA float image through the reference's own image loop
// (ColorBlockCompressor::compress / FloatColorCompressor::compress,
// BlockCompressor.cpp:60-205) with its ParallelTaskDispatcher over
// nv::ThreadPool (TaskDispatcher.h:44-50, ParallelFor.cpp:14-60), as
// Compressor::compress runs it.  kind as in make_compressor; the output
// buffer receives the single writeData of the compressor.  Production
// BC4/BC5 are not deterministic this way (their seed reads fresh heap
// memory); use ntxref_block for those.
size_t ntxref_compress_image(int kind, int quality, int alpha_mode, const float *cw, const float *planar, int w, int h,
                             uint8_t *out, size_t out_bytes)
{
    nvtt::CompressionOptions::Private o;
    options(o, kind_format(kind), quality, cw);
    Sink sink;
    sink.dst = out; sink.cap = out_bytes;
    nvtt::OutputOptions::Private oo;
    oo.fileHandle = nullptr;
    oo.outputHandler = &sink;
    oo.errorHandler = nullptr;
    oo.outputHeader = false;
    oo.container = nvtt::Container_DDS;
    oo.version = 0;
    oo.srgb = false;
    oo.deleteOutputHandler = false;
    oo.wrapperProxy = nullptr;
    CompressorInterface *c = make_compressor(kind, quality);
    nvtt::ParallelTaskDispatcher dispatcher;
    c->compress(nvtt::AlphaMode(alpha_mode), uint(w), uint(h), 1, planar, &dispatcher, o, oo);
    delete c;
    return sink.used;
}

// Worker threads nv::ParallelFor uses (nvthread's processor count).
int ntxref_thread_count(void) { return int(nv::processorCount()); }

// ntxref_compress_image with the reference's task loop on exactly `threads`
// host threads: nv::ParallelFor's scheme (ParallelFor.cpp:14-60: every
// worker takes the next task index from one atomic counter, one block per
// task) on a pool of the given size, instead of nv::ThreadPool's
// processorCount() workers (sysconf, which ignores the CPU share a process
// is given).  The compressor and its per-block code are the reference's.
size_t ntxref_compress_image_threads(int kind, int quality, int alpha_mode, const float *cw, const float *planar,
                                     int w, int h, uint8_t *out, size_t out_bytes, int threads)
{
    struct Dispatcher : public nvtt::TaskDispatcher {
        int threads = 1;
        void dispatch(nvtt::Task *task, void *context, int count) override
        {
            std::atomic<int> next(0);
            auto work = [&]() {
                for (int i = next.fetch_add(1); i < count; i = next.fetch_add(1)) task(context, i);
            };
            std::vector<std::thread> pool;
            for (int t = 1; t < threads; t++) pool.emplace_back(work);
            work();
            for (auto &t : pool) t.join();
        }
    } dispatcher;
    dispatcher.threads = threads < 1 ? 1 : threads;
    nvtt::CompressionOptions::Private o;
    options(o, kind_format(kind), quality, cw);
    Sink sink;
    sink.dst = out; sink.cap = out_bytes;
    nvtt::OutputOptions::Private oo;
    oo.fileHandle = nullptr;
    oo.outputHandler = &sink;
    oo.errorHandler = nullptr;
    oo.outputHeader = false;
    oo.container = nvtt::Container_DDS;
    oo.version = 0;
    oo.srgb = false;
    oo.deleteOutputHandler = false;
    oo.wrapperProxy = nullptr;
    CompressorInterface *c = make_compressor(kind, quality);
    c->compress(nvtt::AlphaMode(alpha_mode), uint(w), uint(h), 1, planar, &dispatcher, o, oo);
    delete c;
    return sink.used;
}

// FloatImage::fastDownSample (FloatImage.cpp:559-737) on planar RGBA.
void ntxref_fast_downsample(int W, int H, const float *src, float *dst)
{
    FloatImage *img = planar_image(W, H, src);
    FloatImage *out = img->fastDownSample();
    copy_out(out, dst);
    delete out;
    delete img;
}

// Surface::buildNextMipmap (Surface.cpp:1344-1406) on a depth-1 planar RGBA
// image: fastDownSample for a Box of width 0.5 without transparency,
// otherwise FloatImage::downSample with the filter (the alpha-first order
// for AlphaMode_Transparency).  filter: 0 Box, 1 Triangle, 2 Kaiser; params
// = (alpha, stretch) for Kaiser or NULL; wrap: FloatImage::WrapMode.
void ntxref_build_next_mipmap(int filter, float width, const float *params, int wrap, int alpha_mode, int W, int H,
                              const float *src, float *dst)
{
    FloatImage *img = planar_image(W, H, src);
    FloatImage *out;
    if (alpha_mode == nvtt::AlphaMode_Transparency) {
        Filter *f = make_filter(filter, width, params);
        out = img->downSample(*f, FloatImage::WrapMode(wrap), 3);
        delete f;
    } else if (filter == 0 && width == 0.5f) {
        out = img->fastDownSample();
    } else {
        Filter *f = make_filter(filter, width, params);
        out = img->downSample(*f, FloatImage::WrapMode(wrap));
        delete f;
    }
    copy_out(out, dst);
    delete out;
    delete img;
}

// PolyphaseKernel(filter, src, dst, 32) (Filter.cpp:563-608): window size,
// kernel width and the dst x window weights.
int ntxref_polyphase_kernel(int filter, float width, const float *params, int src_len, int dst_len, float *weights,
                            float *kwidth)
{
    Filter *f = make_filter(filter, width, params);
    PolyphaseKernel k(*f, uint(src_len), uint(dst_len), 32);
    const int ws = k.windowSize();
    if (weights)
        for (int i = 0; i < dst_len; i++)
            for (int j = 0; j < ws; j++) weights[i * ws + j] = k.valueAt(uint(i), uint(j));
    if (kwidth) *kwidth = k.width();
    delete f;
    return ws;
}

// nvmath/Gamma.cpp powf_5_11 / powf_11_5 over an array (which: 0 = 5/11,
// 1 = 11/5), in place.
void ntxref_powf_array(int which, float *x, int count)
{
    if (which) powf_11_5(x, x, count);
    else powf_5_11(x, x, count);
}

// FloatImage::toLinear / toGamma (FloatImage.cpp:258-297) on the R, G, B
// planes of a planar RGBA image of n texels, in place.
void ntxref_gamma(int to_gamma, size_t n, float *planar, float gamma)
{
    FloatImage *img = planar_image(int(n), 1, planar);
    if (to_gamma) img->toGamma(0, 3, gamma);
    else img->toLinear(0, 3, gamma);
    copy_out(img, planar);
    delete img;
}

// FloatImage::normalize(0) (FloatImage.cpp:201-219): the first three planes
// as vectors, normalised in place.
void ntxref_normalize(size_t n, float *planar)
{
    FloatImage *img = planar_image(int(n), 1, planar);
    img->normalize(0);
    copy_out(img, planar);
    delete img;
}

// Block decoders (BlockDXT.cpp): format 1 BC1 (BlockDXT1::decodeBlock),
// 4 BC3 (BlockDXT5), 6 BC4 (BlockATI1), 7 BC5 (BlockATI2); 16 BGRA texels out.
void ntxref_decode_block(int format, int d3d9, const uint8_t *blk, uint8_t *bgra64)
{
    ColorBlock cb;
    for (int i = 0; i < 16; i++) cb.color(i) = Color32(0, 0, 0, 0);
    switch (format) {
    case 1: { BlockDXT1 b; memcpy(&b, blk, 8); b.decodeBlock(&cb, d3d9 != 0); break; }
    case 4: { BlockDXT5 b; memcpy(&b, blk, 16); b.decodeBlock(&cb, d3d9 != 0); break; }
    case 6: { BlockATI1 b; memcpy(&b, blk, 8); b.decodeBlock(&cb, d3d9 != 0); break; }
    case 7: { BlockATI2 b; memcpy(&b, blk, 16); b.decodeBlock(&cb, d3d9 != 0); break; }
    }
    for (int i = 0; i < 16; i++) {
        const Color32 c = cb.color(i);
        bgra64[4 * i] = c.b; bgra64[4 * i + 1] = c.g; bgra64[4 * i + 2] = c.r; bgra64[4 * i + 3] = c.a;
    }
}

}  // extern "C"
