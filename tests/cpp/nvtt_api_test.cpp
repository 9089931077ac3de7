// nvtt_api_test.cpp -- drives the nvtt.h front end the way an NVTT user does
// (compare src/nvtt/tests/imperativeapi.cpp): Surface::setImage, optional
// toLinearFromSrgb, a buildNextMipmap loop, Compressor::compress per level with
// an OutputHandler that captures the bytes.  Used by tests/test_gpu_nvtt_api.py.
//
//   nvtt_api_test chain <in.raw> <out.bin> <format> <filter> <srgb> <alphaMode> <wrapMode>
//       in.raw:  int32 w, int32 h, then w*h BGRA8 texels
//       out.bin: per level int32 w, h, size, then size bytes; then int32 event count
//                and the handler events (int32 code, six int32 arguments)
//   nvtt_api_test process <in.raw> <out.bin> <format> <filter> <cube> <alphaMode> <wrapMode> <normalMap>
//                         <inGamma> <outGamma> <container> <header> <mipmaps> <maxLevel> <normalize>
//       Compressor::process through InputOptions.  in.raw: int32 w, h, faces,
//       supplied levels L, then for each face its levels 0..L-1 as BGRA8 texels.
//       out.bin: int32 byte count, the bytes of every writeData in order, then
//                the handler events as above, then int32 error count and the errors
//   nvtt_api_test process-time <in.raw> <format> <filter> <reps>
//       the same call (2D, defaults otherwise) repeated: prints the best wall
//       time of one process() call in ms and the bytes it delivered
//   nvtt_api_test api      checks error reporting, handler order and Surface semantics
//   nvtt_api_test batch    Compressor::compress(BatchList) against per-image compress
//   nvtt_api_test device-error   run with NTX_SCRATCH_LIMIT_BYTES=1048576: an
//       image whose device staging exceeds the limit reports Error_CudaError
//       (begin, error, end, no data) and the next, smaller image compresses
#include <cfloat>
#include <chrono>
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

int run_process(int argc, char **argv)
{
    if (argc < 17) return fail("usage");
    FILE *f = std::fopen(argv[2], "rb");
    if (!f) return fail("open input");
    int32_t hd[4];
    if (std::fread(hd, 4, 4, f) != 4) return fail("read header");
    const int w = hd[0], h = hd[1], faces = hd[2], supplied = hd[3];
    nvtt::InputOptions io;
    io.setTextureLayout(std::atoi(argv[6]) ? nvtt::TextureType_Cube : nvtt::TextureType_2D, w, h);
    for (int face = 0; face < faces; face++) {
        int lw = w, lh = h;
        for (int m = 0; m < supplied; m++) {
            std::vector<uint8_t> bgra(size_t(lw) * lh * 4);
            if (std::fread(bgra.data(), 1, bgra.size(), f) != bgra.size()) return fail("read texels");
            CHECK(io.setMipmapData(bgra.data(), lw, lh, 1, face, m));
            lw = lw / 2 > 1 ? lw / 2 : 1;
            lh = lh / 2 > 1 ? lh / 2 : 1;
        }
    }
    std::fclose(f);
    io.setMipmapFilter(nvtt::MipmapFilter(std::atoi(argv[5])));
    io.setAlphaMode(nvtt::AlphaMode(std::atoi(argv[7])));
    io.setWrapMode(nvtt::WrapMode(std::atoi(argv[8])));
    io.setNormalMap(std::atoi(argv[9]) != 0);
    io.setGamma(float(std::atof(argv[10])), float(std::atof(argv[11])));
    io.setMipmapGeneration(std::atoi(argv[14]) != 0, std::atoi(argv[15]));
    io.setNormalizeMipmaps(std::atoi(argv[16]) != 0);

    nvtt::CompressionOptions co;
    co.setFormat(nvtt::Format(std::atoi(argv[4])));
    if (argc > 17) {
        // The learned BC1 form with the network parameters in file argv[17].
        FILE *pf = std::fopen(argv[17], "rb");
        if (!pf) return fail("open parameters");
        std::vector<float> params(8921);
        const size_t got = std::fread(params.data(), sizeof(float), params.size(), pf);
        std::fclose(pf);
        if (got != params.size()) return fail("read parameters");
        CHECK(!co.setLearnedBC1(params.data(), 17));
        CHECK(co.setLearnedBC1(params.data(), int(params.size())));
    }
    nvtt::OutputOptions oo;
    Capture cap;
    Errors err;
    oo.setOutputHandler(&cap);
    oo.setErrorHandler(&err);
    oo.setContainer(nvtt::Container(std::atoi(argv[12])));
    oo.setOutputHeader(std::atoi(argv[13]) != 0);
    oo.setUserVersion(7);
    nvtt::Compressor compressor;
    CHECK(compressor.isCudaAccelerationEnabled());
    const bool ok = compressor.process(io, co, oo);
    const int est = compressor.estimateSize(io, co);

    FILE *o = std::fopen(argv[3], "wb");
    if (!o) return fail("open output");
    int32_t nb = int32_t(cap.bytes.size());
    std::fwrite(&nb, 4, 1, o);
    std::fwrite(cap.bytes.data(), 1, cap.bytes.size(), o);
    int32_t n = int32_t(cap.events.size() / 7);
    std::fwrite(&n, 4, 1, o);
    std::fwrite(cap.events.data(), 4, cap.events.size(), o);
    int32_t ne = int32_t(err.seen.size());
    std::fwrite(&ne, 4, 1, o);
    for (nvtt::Error e : err.seen) {
        int32_t v = int32_t(e);
        std::fwrite(&v, 4, 1, o);
    }
    int32_t tail[2] = { ok ? 1 : 0, est };
    std::fwrite(tail, 4, 2, o);
    std::fclose(o);
    return 0;
}

struct Sink : nvtt::OutputHandler {
    size_t bytes = 0;
    void beginImage(int, int, int, int, int, int) override {}
    bool writeData(const void *, int size) override
    {
        bytes += size_t(size);
        return true;
    }
    void endImage() override {}
};

int run_process_time(int argc, char **argv)
{
    if (argc < 6) return fail("usage");
    FILE *f = std::fopen(argv[2], "rb");
    if (!f) return fail("open input");
    int32_t hd[4];
    if (std::fread(hd, 4, 4, f) != 4) return fail("read header");
    std::vector<uint8_t> bgra(size_t(hd[0]) * hd[1] * 4);
    if (std::fread(bgra.data(), 1, bgra.size(), f) != bgra.size()) return fail("read texels");
    std::fclose(f);
    nvtt::InputOptions io;
    io.setTextureLayout(nvtt::TextureType_2D, hd[0], hd[1]);
    CHECK(io.setMipmapData(bgra.data(), hd[0], hd[1]));
    io.setMipmapFilter(nvtt::MipmapFilter(std::atoi(argv[4])));
    nvtt::CompressionOptions co;
    co.setFormat(nvtt::Format(std::atoi(argv[3])));
    nvtt::OutputOptions oo;
    Sink sink;
    Errors err;
    oo.setOutputHandler(&sink);
    oo.setErrorHandler(&err);
    nvtt::Compressor compressor;
    CHECK(compressor.isCudaAccelerationEnabled());
    const int reps = std::atoi(argv[5]);
    double best = 1e30;
    for (int r = 0; r <= reps; r++) {  // the first call warms up
        sink.bytes = 0;
        const auto t0 = std::chrono::steady_clock::now();
        CHECK(compressor.process(io, co, oo));
        const auto t1 = std::chrono::steady_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        if (r > 0 && ms < best) best = ms;
    }
    CHECK(err.seen.empty());
    std::printf("process_ms %.4f bytes %zu\n", best, sink.bytes);
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
    co.setFormat(nvtt::Format_BC2);  // CompressorDXT3: not on this path
    CHECK(comp.compress(16, 8, 1, 0, 0, img.data(), co, oo));
    CHECK(err.seen.size() == 1 && err.seen[0] == nvtt::Error_UnsupportedFeature);

    err.seen.clear();
    co.setFormat(nvtt::Format_BC1);
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

    // Decoding compressed output and measuring it (Surface::setImage2D,
    // nvtt::rmsError / rmsAlphaError): a 16 x 8 ramp through BC3.
    {
        std::vector<uint8_t> px(16 * 8 * 4);
        for (int i = 0; i < 16 * 8; i++) {
            const int x = i % 16, y = i / 16;
            px[4 * i] = uint8_t(x * 16);        // B: a ramp along x (colinear in each block)
            px[4 * i + 1] = uint8_t(x * 8);     // G
            px[4 * i + 2] = uint8_t(128);       // R
            px[4 * i + 3] = uint8_t(y < 4 ? 255 : 128);
        }
        nvtt::Surface src;
        CHECK(src.setImage(nvtt::InputFormat_BGRA_8UB, 16, 8, 1, px.data()));
        Capture bc3;
        nvtt::OutputOptions oo3;
        oo3.setOutputHandler(&bc3);
        nvtt::CompressionOptions co3;
        co3.setFormat(nvtt::Format_BC3);
        CHECK(comp.compress(src, 0, 0, co3, oo3));
        CHECK(bc3.bytes.size() == 4 * 2 * 16);
        nvtt::Surface dec;
        CHECK(dec.setImage2D(nvtt::Format_BC3, nvtt::Decoder_D3D10, 16, 8, bc3.bytes.data()));
        CHECK(dec.width() == 16 && dec.height() == 8);
        const float e = nvtt::rmsError(src, dec), ea = nvtt::rmsAlphaError(src, dec);
        if (!(e > 0.0f && e < 0.05f)) std::fprintf(stderr, "rmsError %g rmsAlphaError %g\n", e, ea);
        CHECK(e > 0.0f && e < 0.05f);
        CHECK(ea >= 0.0f && ea < 0.01f);
        CHECK(nvtt::rmsError(src, src) == 0.0f);
        nvtt::Surface small;
        CHECK(small.setImage(8, 8, 1));
        CHECK(nvtt::rmsError(src, small) == FLT_MAX);  // different layouts
        CHECK(!dec.setImage2D(nvtt::Format_BC7, nvtt::Decoder_D3D10, 16, 8, bc3.bytes.data()));
    }
    std::printf("api ok\n");
    return 0;
}

// A deterministic BGRA8 image: gradients in B and G, hashed noise in R and A.
bool make_surface(nvtt::Surface &s, int w, int h, uint32_t seed)
{
    std::vector<uint8_t> bgra(size_t(w) * h * 4);
    uint32_t x = seed * 2654435761u + 1u;
    for (int i = 0; i < w * h; i++) {
        x = x * 1664525u + 1013904223u;
        const int px = i % w, py = i / w;
        bgra[4 * i + 0] = uint8_t((px * 255) / (w > 1 ? w - 1 : 1));
        bgra[4 * i + 1] = uint8_t((py * 255) / (h > 1 ? h - 1 : 1));
        bgra[4 * i + 2] = uint8_t(x >> 24);
        bgra[4 * i + 3] = uint8_t(255 - ((x >> 16) & 0x3f));
    }
    return s.setImage(nvtt::InputFormat_BGRA_8UB, w, h, 1, bgra.data());
}

// Compressor::compress(BatchList) against compress(Surface) per image and
// buildNextMipmap per level: same bytes and the same handler events, entry
// by entry, for the persistent-batch formats and for BC3.
int run_batch()
{
    const int sizes[4][2] = { { 128, 64 }, { 37, 90 }, { 256, 256 }, { 1, 16 } };
    const nvtt::MipmapFilter filters[4] = { nvtt::MipmapFilter_Box, nvtt::MipmapFilter_Box, nvtt::MipmapFilter_Kaiser,
                                            nvtt::MipmapFilter_Triangle };
    const int chain_levels[4] = { 0, 1, 0, 3 };  // entry 1 is a single image (Append)
    const nvtt::Format formats[4] = { nvtt::Format_BC1, nvtt::Format_BC4, nvtt::Format_BC5, nvtt::Format_BC3 };
    nvtt::Compressor compressor;
    CHECK(compressor.isCudaAccelerationEnabled());
    for (nvtt::Format format : formats) {
        nvtt::Surface surf[4];
        for (int i = 0; i < 4; i++) CHECK(make_surface(surf[i], sizes[i][0], sizes[i][1], uint32_t(17 * i + 3)));
        surf[3].setWrapMode(nvtt::WrapMode_Repeat);
        nvtt::CompressionOptions co;
        co.setFormat(format);
        co.setQuality(nvtt::Quality_Normal);

        Capture expect[4], got[4];
        for (int i = 0; i < 4; i++) {
            nvtt::Surface t = surf[i];
            nvtt::OutputOptions oo;
            oo.setOutputHandler(&expect[i]);
            int levels = i == 1 ? 1 : t.countMipmaps();
            if (chain_levels[i] > 0 && chain_levels[i] < levels) levels = chain_levels[i];
            for (int m = 0; m < levels; m++) {
                CHECK(compressor.compress(t, i, 5 + m, co, oo));
                if (m + 1 < levels) CHECK(t.buildNextMipmap(filters[i]));
            }
        }
        nvtt::OutputOptions oos[4];
        Errors err;
        nvtt::BatchList list;
        for (int i = 0; i < 4; i++) {
            oos[i].setOutputHandler(&got[i]);
            oos[i].setErrorHandler(&err);
            if (i == 1) list.Append(&surf[i], i, 5, &oos[i]);
            else list.AppendChain(&surf[i], i, 5, filters[i], chain_levels[i], &oos[i]);
        }
        CHECK(list.GetSize() == 4);
        CHECK(compressor.compress(list, co));
        CHECK(err.seen.empty());
        for (int i = 0; i < 4; i++) {
            CHECK(got[i].bytes == expect[i].bytes);
            CHECK(got[i].events == expect[i].events);
        }
    }
    // Unsupported format: every image reports, none writes.
    nvtt::Surface s;
    CHECK(make_surface(s, 16, 16, 9));
    nvtt::CompressionOptions co;
    co.setFormat(nvtt::Format_BC7);
    Capture cap;
    Errors err;
    nvtt::OutputOptions oo;
    oo.setOutputHandler(&cap);
    oo.setErrorHandler(&err);
    nvtt::BatchList list;
    list.AppendChain(&s, 0, 0, nvtt::MipmapFilter_Box, 0, &oo);
    CHECK(compressor.compress(list, co));
    CHECK(err.seen.size() == 5 && err.seen[0] == nvtt::Error_UnsupportedFeature && cap.bytes.empty());
    std::printf("batch ok\n");
    return 0;
}

}  // namespace

// CudaCompressorDXT.cpp:183-189: a device error is reported through the
// ErrorHandler and processing continues.  The backend's scratch limit makes
// the 512 x 512 image's upload staging (4 MiB) fail to allocate.
int run_device_error()
{
    nvtt::Compressor comp;
    nvtt::CompressionOptions co;
    co.setFormat(nvtt::Format_BC3);
    Capture cap;
    Errors err;
    nvtt::OutputOptions oo;
    oo.setOutputHandler(&cap);
    oo.setErrorHandler(&err);
    std::vector<float> big(size_t(4) * 512 * 512, 0.5f), small(size_t(4) * 64 * 64);
    for (size_t i = 0; i < small.size(); i++) small[i] = float((i * 37) % 251) / 250.0f;
    for (int round = 0; round < 2; round++) {
        cap.events.clear();
        cap.bytes.clear();
        err.seen.clear();
        CHECK(comp.compress(512, 512, 1, 0, 0, big.data(), co, oo));
        CHECK(err.seen.size() == 1 && err.seen[0] == nvtt::Error_CudaError);
        CHECK(cap.bytes.empty());
        CHECK(cap.events.size() == 14 && cap.events[0] == 1 && cap.events[1] == 512 * 512 && cap.events[7] == 3);
        cap.events.clear();
        err.seen.clear();
        CHECK(comp.compress(64, 64, 1, 0, 0, small.data(), co, oo));
        CHECK(err.seen.empty());
        CHECK(cap.bytes.size() == 16 * 16 * 16);
        CHECK(cap.events.size() == 21 && cap.events[7] == 2 && cap.events[14] == 3);
    }
    std::printf("device-error ok\n");
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && std::strcmp(argv[1], "device-error") == 0) return run_device_error();
    if (argc >= 2 && std::strcmp(argv[1], "chain") == 0) return run_chain(argc, argv);
    if (argc >= 2 && std::strcmp(argv[1], "process") == 0) return run_process(argc, argv);
    if (argc >= 2 && std::strcmp(argv[1], "process-time") == 0) return run_process_time(argc, argv);
    if (argc >= 2 && std::strcmp(argv[1], "api") == 0) return run_api();
    if (argc >= 2 && std::strcmp(argv[1], "batch") == 0) return run_batch();
    return fail("usage: nvtt_api_test chain|api|batch ...");
}
