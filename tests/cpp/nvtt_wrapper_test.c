/* nvtt_wrapper_test.c -- drives the C interface (nvtt_wrapper.h) the way a C
 * or FFI caller does, with the same arguments and output file as
 * nvtt_api_test's "process" mode, so the two can be compared byte for byte.
 * Used by tests/test_gpu_process.py.
 *
 *   nvtt_wrapper_test process <in.raw> <out.bin> <format> <filter> <cube> <alphaMode> <wrapMode> <normalMap>
 *                             <inGamma> <outGamma> <container> <header> <mipmaps> <maxLevel> <normalize>
 *                             [<quality> <colorDithering> <binaryAlpha> [<file>]]
 *       <file>: nvttSetOutputOptionsFileName after the handlers, so the file
 *       writer takes the stream (the callbacks then see nothing)
 *   nvtt_wrapper_test info    prints the version and every error string */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nvtt_wrapper.h"

static unsigned char *g_bytes;
static size_t g_nbytes, g_capbytes;
static int32_t g_events[4096 * 7];
static int g_nevents;
static int32_t g_errors[64];
static int g_nerrors;

static void event(int code, int a, int b, int c, int d, int e, int f)
{
    if (g_nevents >= 4096) return;
    int32_t *p = g_events + 7 * g_nevents++;
    p[0] = code; p[1] = a; p[2] = b; p[3] = c; p[4] = d; p[5] = e; p[6] = f;
}

static void on_begin(int size, int w, int h, int d, int face, int mip) { event(1, size, w, h, d, face, mip); }

static bool on_write(const void *data, int size)
{
    if (g_nbytes + (size_t)size > g_capbytes) {
        g_capbytes = (g_nbytes + (size_t)size) * 2;
        g_bytes = (unsigned char *)realloc(g_bytes, g_capbytes);
        if (!g_bytes) return false;
    }
    memcpy(g_bytes + g_nbytes, data, (size_t)size);
    g_nbytes += (size_t)size;
    event(2, size, 0, 0, 0, 0, 0);
    return true;
}

static void on_end(void) { event(3, 0, 0, 0, 0, 0, 0); }

static void on_error(NvttError e)
{
    if (g_nerrors < 64) g_errors[g_nerrors++] = (int32_t)e;
}

static int fail(const char *what)
{
    fprintf(stderr, "nvtt_wrapper_test: %s\n", what);
    return 1;
}

static int run_process(int argc, char **argv)
{
    if (argc < 17) return fail("usage");
    FILE *f = fopen(argv[2], "rb");
    if (!f) return fail("open input");
    int32_t hd[4];
    if (fread(hd, 4, 4, f) != 4) return fail("read header");
    const int w = hd[0], h = hd[1], faces = hd[2], supplied = hd[3];

    NvttInputOptions *io = nvttCreateInputOptions();
    nvttSetInputOptionsTextureLayout(io, atoi(argv[6]) ? NVTT_TextureType_Cube : NVTT_TextureType_2D, w, h, 1, 1);
    for (int face = 0; face < faces; face++) {
        int lw = w, lh = h;
        for (int m = 0; m < supplied; m++) {
            const size_t n = (size_t)lw * lh * 4;
            unsigned char *bgra = (unsigned char *)malloc(n);
            if (!bgra || fread(bgra, 1, n, f) != n) return fail("read texels");
            if (nvttSetInputOptionsMipmapData(io, bgra, lw, lh, 1, face, m) != NVTT_True) return fail("mipmap data");
            free(bgra);
            lw = lw / 2 > 1 ? lw / 2 : 1;
            lh = lh / 2 > 1 ? lh / 2 : 1;
        }
    }
    fclose(f);
    nvttSetInputOptionsMipmapFilter(io, (NvttMipmapFilter)atoi(argv[5]));
    nvttSetInputOptionsAlphaMode(io, (NvttAlphaMode)atoi(argv[7]));
    nvttSetInputOptionsWrapMode(io, (NvttWrapMode)atoi(argv[8]));
    nvttSetInputOptionsNormalMap(io, atoi(argv[9]) ? NVTT_True : NVTT_False);
    nvttSetInputOptionsGamma(io, (float)atof(argv[10]), (float)atof(argv[11]));
    nvttSetInputOptionsMipmapGeneration(io, atoi(argv[14]) ? NVTT_True : NVTT_False, atoi(argv[15]));
    nvttSetInputOptionsNormalizeMipmaps(io, atoi(argv[16]) ? NVTT_True : NVTT_False);

    NvttCompressionOptions *co = nvttCreateCompressionOptions();
    nvttSetCompressionOptionsFormat(co, (NvttFormat)atoi(argv[4]));
    if (argc >= 20) {
        nvttSetCompressionOptionsQuality(co, (NvttQuality)atoi(argv[17]));
        nvttSetCompressionOptionsQuantization(co, atoi(argv[18]) ? NVTT_True : NVTT_False, NVTT_False,
                                              atoi(argv[19]) ? NVTT_True : NVTT_False, 127);
    }

    NvttOutputOptions *oo = nvttCreateOutputOptions();
    nvttSetOutputOptionsOutputHandler(oo, on_begin, on_write, on_end);
    nvttSetOutputOptionsErrorHandler(oo, on_error);
    nvttSetOutputOptionsContainer(oo, (NvttContainer)atoi(argv[12]));
    nvttSetOutputOptionsOutputHeader(oo, atoi(argv[13]) ? NVTT_True : NVTT_False);
    if (argc >= 21 && strcmp(argv[20], "-") != 0) nvttSetOutputOptionsFileName(oo, argv[20]);

    NvttCompressor *c = nvttCreateCompressor();
    if (nvttIsCudaAccelerationEnabled(c) != NVTT_True) return fail("acceleration off");
    const NvttBoolean ok = nvttCompress(c, io, co, oo);
    const int est = nvttEstimateSize(c, io, co);

    FILE *o = fopen(argv[3], "wb");
    if (!o) return fail("open output");
    int32_t nb = (int32_t)g_nbytes;
    fwrite(&nb, 4, 1, o);
    if (g_nbytes) fwrite(g_bytes, 1, g_nbytes, o);
    int32_t ne = g_nevents;
    fwrite(&ne, 4, 1, o);
    fwrite(g_events, 4, (size_t)g_nevents * 7, o);
    int32_t nerr = g_nerrors;
    fwrite(&nerr, 4, 1, o);
    fwrite(g_errors, 4, (size_t)g_nerrors, o);
    int32_t tail[2] = { ok == NVTT_True ? 1 : 0, est };
    fwrite(tail, 4, 2, o);
    fclose(o);

    nvttDestroyCompressor(c);
    nvttDestroyOutputOptions(oo);
    nvttDestroyCompressionOptions(co);
    nvttDestroyInputOptions(io);
    free(g_bytes);
    return 0;
}

static int run_info(void)
{
    /* The reference header's spellings of the texture types. */
    const NvttTextureType t3 = TextureType_3D, ta = TextureType_Array;
    if (t3 != NVTT_TextureType_3D || ta != NVTT_TextureType_Array || (int)t3 != 2 || (int)ta != 3)
        return fail("texture type values");
    printf("version %u\n", nvttVersion());
    for (int e = NVTT_Error_Unknown; e <= NVTT_Error_UnsupportedOutputFormat; e++)
        printf("error %d %s\n", e, nvttErrorString((NvttError)e));
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && strcmp(argv[1], "process") == 0) return run_process(argc, argv);
    if (argc >= 2 && strcmp(argv[1], "info") == 0) return run_info();
    return fail("unknown mode");
}
