// nvtt_api_test.cpp -- drives the nvtt.h front end the way an NVTT user does
// (compare src/nvtt/tests/imperativeapi.cpp): Surface::setImage, optional
// toLinearFromSrgb, a buildNextMipmap loop, Compressor::compress per level with
// an OutputHandler that captures the bytes.  Used by tests/test_gpu_nvtt_api.py.
//
//   nvtt_api_test chain <in.raw> <out.bin> <format> <filter> <srgb> <alphaMode> <wrapMode>
//       in.raw:  int32 w, int32 h, then w*h BGRA8 texels
//       out.bin: per level int32 w, h, size, then size bytes; then int32 event count
//                and the handler events (int32 code, six int32 arguments)
//   nvtt_api_test api      checks error reporting, handler order and Surface semantics
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nvtt.h"

namespace {

struct Capture : nvtt::OutputHandler {
    std::vector<uint8_t> bytes;
    std::vector<int> events;  // code, a0..a5
    void beginImage(int size, int w, int h, int d, int face, int mip) override
    {
        int e[7] = { 1, size, w, h, d, face, mip };
        events.insert(events.end(), e, e + 7);
    }
    bool writeData(const void *data, int size) override
    {
        const uint8_t *p = static_cast<const uint8_t *>(data);
        bytes.insert(bytes.end(), p, p + size);
        int e[7] = { 2, size, 0, 0, 0, 0, 0 };
        events.insert(events.end(), e, e + 7);
        return true;
    }
    void endImage() override
    {
        int e[7] = { 3, 0, 0, 0, 0, 0, 0 };
        events.insert(events.end(), e, e + 7);
    }
};

struct Errors : nvtt::ErrorHandler {
    std::vector<nvtt::Error> seen;
    void error(nvtt::Error e) override { seen.push_back(e); }
};

int fail(const char *what)
{
    std::fprintf(stderr, "FAILED: %s\n", what);
    return 1;
}

#define CHECK(cond)                      \
    do {                                 \
        if (!(cond)) return fail(#cond); \
    } while (0)

int run_chain(int argc, char **argv)
{
    if (argc < 9) return fail("usage");
    FILE *f = std::fopen(argv[2], "rb");
    if (!f) return fail("open input");
    int32_t wh[2];
    if (std::fread(wh, 4, 2, f) != 2) return fail("read header");
    std::vector<uint8_t> bgra(size_t(wh[0]) * wh[1] * 4);
    if (std::fread(bgra.data(), 1, bgra.size(), f) != bgra.size()) return fail("read texels");
    std::fclose(f);

    const nvtt::Format format = nvtt::Format(std::atoi(argv[4]));
    const nvtt::MipmapFilter filter = nvtt::MipmapFilter(std::atoi(argv[5]));
    const bool srgb = std::atoi(argv[6]) != 0;

    nvtt::Surface img;
    CHECK(img.setImage(nvtt::InputFormat_BGRA_8UB, wh[0], wh[1], 1, bgra.data()));
    img.setAlphaMode(nvtt::AlphaMode(std::atoi(argv[7])));
    img.setWrapMode(nvtt::WrapMode(std::atoi(argv[8])));
    if (srgb) img.toLinearFromSrgb();

    nvtt::Compressor compressor;
    CHECK(compressor.isCudaAccelerationEnabled());
    nvtt::CompressionOptions co;
    co.setFormat(format);
    co.setQuality(nvtt::Quality_Normal);
    Capture cap;
    Errors err;
    nvtt::OutputOptions oo;
    oo.setOutputHandler(&cap);
    oo.setErrorHandler(&err);

    FILE *o = std::fopen(argv[3], "wb");
    if (!o) return fail("open output");
    const int levels = img.countMipmaps();
    for (int m = 0; m < levels; m++) {
        cap.bytes.clear();
        CHECK(compressor.compress(img, 0, m, co, oo));
        int32_t hdr[3] = { img.width(), img.height(), int32_t(cap.bytes.size()) };
        std::fwrite(hdr, 4, 3, o);
        std::fwrite(cap.bytes.data(), 1, cap.bytes.size(), o);
        if (m + 1 < levels) CHECK(img.buildNextMipmap(filter));
    }
    CHECK(!img.buildNextMipmap(filter));  // 1x1: no further level
    CHECK(err.seen.empty());
    int32_t n = int32_t(cap.events.size() / 7);
    std::fwrite(&n, 4, 1, o);
    std::fwrite(cap.events.data(), 4, cap.events.size(), o);
    std::fclose(o);
    return 0;
}

int run_api()
{
    // Surface queries (Surface.cpp:155-207).
    nvtt::Surface s;
    CHECK(s.isNull());
    CHECK(!s.canMakeNextMipmap());
    CHECK(s.setImage(1024, 512, 1));
    CHECK(s.width() == 1024 && s.height() == 512 && s.depth() == 1);
    CHECK(s.countMipmaps() == 11);
    CHECK(s.countMipmaps(4) == 8);
    CHECK(s.wrapMode() == nvtt::WrapMode_Mirror && s.alphaMode() == nvtt::AlphaMode_None);
    const float *z = s.data();
    CHECK(z != nullptr);
    for (int i = 0; i < 4 * 1024 * 512; i += 4099) CHECK(z[i] == 0.0f);

    // Copy-on-write: mipmapping a copy leaves the original alone.
    nvtt::Surface t = s;
    CHECK(t.buildNextMipmap(nvtt::MipmapFilter_Box));
    CHECK(t.width() == 512 && t.height() == 256);
    CHECK(s.width() == 1024 && s.height() == 512);

    // Float inputs land in planar order.
    const int w = 5, h = 3;
    std::vector<float> rgba(4 * w * h), r(w * h);
    for (int i = 0; i < 4 * w * h; i++) rgba[i] = float(i) * 0.01f - 0.1f;
    for (int i = 0; i < w * h; i++) r[i] = float(i) * 0.5f;
    CHECK(s.setImage(nvtt::InputFormat_RGBA_32F, w, h, 1, rgba.data()));
    const float *p = s.data();
    for (int i = 0; i < w * h; i++)
        for (int c = 0; c < 4; c++) CHECK(p[c * w * h + i] == rgba[4 * i + c]);
    CHECK(s.setImage(nvtt::InputFormat_R_32F, w, h, 1, r.data()));
    p = s.data();
    for (int i = 0; i < w * h; i++) CHECK(p[i] == r[i] && p[w * h + i] == 0 && p[3 * w * h + i] == 0);
    std::vector<uint16_t> half(4 * w * h);
    const uint16_t halves[6] = { 0x3c00, 0x0000, 0x8000, 0x3800, 0x0001, 0x7bff };  // 1, 0, -0, 0.5, 2^-24, 65504
    for (int i = 0; i < 4 * w * h; i++) half[i] = halves[i % 6];
    CHECK(s.setImage(nvtt::InputFormat_RGBA_16F, w, h, 1, half.data()));
    p = s.data();
    const float expect[6] = { 1.0f, 0.0f, -0.0f, 0.5f, std::ldexp(1.0f, -24), 65504.0f };
    for (int i = 0; i < w * h; i++)
        for (int c = 0; c < 4; c++) CHECK(std::memcmp(&p[c * w * h + i], &expect[(4 * i + c) % 6], 4) == 0);

    // Handler sequence and errors (Context.cpp:486-516).
    nvtt::Compressor comp;
    nvtt::CompressionOptions co;
    Capture cap;
    Errors err;
    nvtt::OutputOptions oo;
    oo.setOutputHandler(&cap);
    oo.setErrorHandler(&err);
    std::vector<float> img(4 * 16 * 8, 0.25f);
    CHECK(comp.compress(16, 8, 1, 2, 3, img.data(), co, oo));
    CHECK(err.seen.empty());
    CHECK(cap.events.size() == 21);
    CHECK(cap.events[0] == 1 && cap.events[1] == 64 && cap.events[2] == 16 && cap.events[3] == 8 &&
          cap.events[4] == 1 && cap.events[5] == 2 && cap.events[6] == 3);
    CHECK(cap.events[7] == 2 && cap.events[8] == 64);
    CHECK(cap.events[14] == 3);
    CHECK(comp.estimateSize(16, 8, 1, 5, co) == 64 + 16 + 8 + 8 + 8);

    cap.events.clear();
    co.setFormat(nvtt::Format_BC7);
    CHECK(comp.compress(16, 8, 1, 0, 0, img.data(), co, oo));
    CHECK(err.seen.size() == 1 && err.seen[0] == nvtt::Error_UnsupportedFeature);
    CHECK(cap.events.size() == 14 && cap.events[0] == 1 && cap.events[7] == 3);  // begin, end, no data

    err.seen.clear();
    co.setFormat(nvtt::Format_BC1);
    co.setQuality(nvtt::Quality_Fastest);
    CHECK(comp.compress(16, 8, 1, 0, 0, img.data(), co, oo));
    CHECK(err.seen.size() == 1 && err.seen[0] == nvtt::Error_UnsupportedFeature);

    err.seen.clear();
    co.setQuality(nvtt::Quality_Normal);
    comp.enableCudaAcceleration(false);
    CHECK(!comp.isCudaAccelerationEnabled());
    CHECK(comp.compress(16, 8, 1, 0, 0, img.data(), co, oo));
    CHECK(err.seen.size() == 1 && err.seen[0] == nvtt::Error_CudaError);
    comp.enableCudaAcceleration(true);
    CHECK(comp.isCudaAccelerationEnabled());

    // No handlers at all is allowed.
    nvtt::OutputOptions none;
    CHECK(comp.compress(16, 8, 1, 0, 0, img.data(), co, none));
    std::printf("api ok\n");
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc >= 2 && std::strcmp(argv[1], "chain") == 0) return run_chain(argc, argv);
    if (argc >= 2 && std::strcmp(argv[1], "api") == 0) return run_api();
    return fail("usage: nvtt_api_test chain|api ...");
}
