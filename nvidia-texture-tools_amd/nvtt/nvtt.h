// nvtt.h -- the part of NVTT 2.1.2's public API (src/nvtt/nvtt.h) that this
// backend serves: Compressor::compress for BC1/BC3/BC4/BC5, Surface with its
// mipmap builder, and the output/error handler plumbing.  Names, enum values,
// argument order and error behaviour follow the reference so code written
// against it compiles and runs unchanged for this path.
//
// Differences, all deliberate:
//  - Compression and mip filtering run on the MI355X through the C ABI in
//    include/ntx_hip.h.  There is no CPU compressor: without a device, or for
//    a format/quality the device path does not implement, compress() reports
//    through ErrorHandler (Error_CudaError / Error_UnsupportedFeature) and
//    writes no data.  The reference's "w*h < 512 -> CPU" threshold
//    (Context.cpp:494) therefore does not apply.
//  - Surface keeps its FP32 planes in device memory; data()/channel() copy
//    them back on demand.
//  - InputOptions / Compressor::process run the reference's input pipeline
//    (Context.cpp:217-344) on the device for 2D and cube textures: 8-bit and
//    float input, gamma to linear and back, generated or supplied mip levels,
//    normal-map renormalisation, DDS, DDS10 and KTX output.  Resizing
//    (setMaxExtents / setRoundMode changing the extents), height-to-normal
//    conversion and volume textures report Error_UnsupportedFeature.
//  - Surface transforms outside this path are not provided.
#ifndef NVTT_AMD_H
#define NVTT_AMD_H

#define NVTT_VERSION 20102

#if defined(__GNUC__)
#define NVTT_API __attribute__((visibility("default")))
#else
#define NVTT_API
#endif

namespace nvtt {

// nvtt.h:77-115 (values are ABI).
enum Format {
    Format_RGB,
    Format_RGBA = Format_RGB,
    Format_DXT1,
    Format_DXT1a,
    Format_DXT3,
    Format_DXT5,
    Format_DXT5n,
    Format_BC1 = Format_DXT1,
    Format_BC1a = Format_DXT1a,
    Format_BC2 = Format_DXT3,
    Format_BC3 = Format_DXT5,
    Format_BC3n = Format_DXT5n,
    Format_BC4,
    Format_BC5,
    Format_DXT1n,
    Format_CTX1,
    Format_BC6,
    Format_BC7,
    Format_BC3_RGBM,
    Format_ETC1,
    Format_ETC2_R,
    Format_ETC2_RG,
    Format_ETC2_RGB,
    Format_ETC2_RGBA,
    Format_ETC2_RGB_A1,
    Format_ETC2_RGBM,
    Format_PVR_2BPP_RGB,
    Format_PVR_4BPP_RGB,
    Format_PVR_2BPP_RGBA,
    Format_PVR_4BPP_RGBA,
    Format_Count
};

// nvtt.h:148-154.
enum Quality { Quality_Fastest, Quality_Normal, Quality_Production, Quality_Highest };

// nvtt.h:209-214, 217-223, 226-233, 236-241, 274-279.
enum WrapMode { WrapMode_Clamp, WrapMode_Repeat, WrapMode_Mirror };
enum TextureType { TextureType_2D, TextureType_Cube, TextureType_3D, TextureType_Array };
enum InputFormat { InputFormat_BGRA_8UB, InputFormat_RGBA_16F, InputFormat_RGBA_32F, InputFormat_R_32F };
enum MipmapFilter { MipmapFilter_Box, MipmapFilter_Triangle, MipmapFilter_Kaiser };
enum AlphaMode { AlphaMode_None, AlphaMode_Transparency, AlphaMode_Premultiplied };

// nvtt.h:149-155.
enum Decoder { Decoder_D3D10, Decoder_D3D9, Decoder_NV5x };

// nvtt.h:253-262.
enum RoundMode {
    RoundMode_None,
    RoundMode_ToNextPowerOfTwo,
    RoundMode_ToNearestPowerOfTwo,
    RoundMode_ToPreviousPowerOfTwo,
    RoundMode_ToNextMultipleOfFour,
    RoundMode_ToNearestMultipleOfFour,
    RoundMode_ToPreviousMultipleOfFour,
};

// nvtt.h:367-373.
enum Container { Container_DDS, Container_DDS10, Container_KTX };

// nvtt.h:345-355.
enum Error {
    Error_Unknown,
    Error_InvalidInput,
    Error_UnsupportedFeature,
    Error_CudaError,  // reported for any HIP device error
    Error_FileOpen,
    Error_FileWrite,
    Error_UnsupportedOutputFormat,
    Error_Count
};

struct Surface;

// nvtt.h:159-196 (the settings this path reads).
struct CompressionOptions {
    struct Private;
    NVTT_API CompressionOptions();
    NVTT_API ~CompressionOptions();
    NVTT_API void reset();
    NVTT_API void setFormat(Format format);
    NVTT_API void setQuality(Quality quality);
    NVTT_API void setColorWeights(float red, float green, float blue, float alpha = 1.0f);
    // Kept for the uncompressed and dithered outputs, which this path does not
    // produce: process() reports Error_UnsupportedFeature when dithering or
    // binary alpha is asked for a block format (Context.cpp:519-541).
    NVTT_API void setPixelFormat(unsigned int bitcount, unsigned int rmask, unsigned int gmask, unsigned int bmask,
                                 unsigned int amask);
    NVTT_API void setQuantization(bool colorDithering, bool alphaDithering, bool binaryAlpha, int alphaThreshold = 127);
    // Extension, no counterpart in the reference: BC1 at Quality_Normal (no
    // transparency weighting, unit colour weights) by the learned endpoint
    // proposals of ntx_hip_compress_bc1_mlp_images_device instead of the
    // cluster fit, with the network parameters `params` (count must be 8921,
    // NTX_MLP_PARAMS of ntx_hip.h; copied), in Compressor::process and Compressor::compress
    // (a batch list keeps the cluster fit).  Per image: the shipped settings
    // at mip level 0, the refined ones below it, the plain cluster fit under
    // 4096 blocks.  The output is then not the reference's bytes.  NULL (or
    // count 0) switches it off; returns false for a wrong count.
    NVTT_API bool setLearnedBC1(const float *params, int count);
    NVTT_API Format format() const;
    Private &m;

private:
    CompressionOptions(const CompressionOptions &);
    void operator=(const CompressionOptions &);
};

// nvtt.h:330-342.
struct OutputHandler {
    virtual ~OutputHandler() {}
    virtual void beginImage(int size, int width, int height, int depth, int face, int miplevel) = 0;
    virtual bool writeData(const void *data, int size) = 0;
    virtual void endImage() = 0;
};

// nvtt.h:358-364.
struct ErrorHandler {
    virtual ~ErrorHandler() {}
    virtual void error(Error e) = 0;
};

// nvtt.h:281-326 (all setters are accepted; process() reports the
// settings it does not implement, see above).
struct InputOptions {
    struct Private;
    NVTT_API InputOptions();
    NVTT_API ~InputOptions();
    NVTT_API void reset();
    NVTT_API void setTextureLayout(TextureType type, int w, int h, int d = 1, int arraySize = 1);
    NVTT_API void resetTextureLayout();
    // Copies the data.
    NVTT_API bool setMipmapData(const void *data, int w, int h, int d = 1, int face = 0, int mipmap = 0);
    NVTT_API void setFormat(InputFormat format);
    NVTT_API void setAlphaMode(AlphaMode alphaMode);
    NVTT_API void setGamma(float inputGamma, float outputGamma);
    NVTT_API void setWrapMode(WrapMode mode);
    NVTT_API void setMipmapFilter(MipmapFilter filter);
    NVTT_API void setMipmapGeneration(bool enabled, int maxLevel = -1);
    NVTT_API void setKaiserParameters(float width, float alpha, float stretch);
    NVTT_API void setNormalMap(bool b);
    NVTT_API void setConvertToNormalMap(bool convert);
    NVTT_API void setHeightEvaluation(float redScale, float greenScale, float blueScale, float alphaScale);
    NVTT_API void setNormalFilter(float sm, float medium, float big, float large);
    NVTT_API void setNormalizeMipmaps(bool b);
    NVTT_API void setMaxExtents(int d);
    NVTT_API void setRoundMode(RoundMode mode);
    Private &m;

private:
    InputOptions(const InputOptions &);
    void operator=(const InputOptions &);
};

// nvtt.h:378-399.
struct OutputOptions {
    struct Private;
    NVTT_API OutputOptions();
    NVTT_API ~OutputOptions();
    NVTT_API void reset();
    // Writes the output to a file (a handler of the library's own).
    NVTT_API void setFileName(const char *fileName);
    NVTT_API void setOutputHandler(OutputHandler *outputHandler);
    NVTT_API void setErrorHandler(ErrorHandler *errorHandler);
    NVTT_API void setOutputHeader(bool outputHeader);
    NVTT_API void setContainer(Container container);
    NVTT_API void setUserVersion(int version);
    NVTT_API void setSrgbFlag(bool b);
    Private &m;

private:
    OutputOptions(const OutputOptions &);
    void operator=(const OutputOptions &);
};

// nvtt.h:402-410.
typedef void Task(void *context, int id);
struct TaskDispatcher {
    virtual ~TaskDispatcher() {}
    virtual void dispatch(Task *task, void *context, int count) = 0;
};

// Extension with no counterpart in NVTT 2.1.2 (the shape of the BatchList
// of later NVTT versions): many surfaces compressed by one call.  Append adds
// one surface, output as (face, mipmap); AppendChain adds a surface and the
// mip chain Surface::buildNextMipmap(filter) would build from it, mipmapCount
// levels (0: all, countMipmaps()), output as mip levels mipmap, mipmap + 1,
// ...  The surfaces must stay alive and unchanged until compress returns.
struct BatchList {
    struct Private;
    NVTT_API BatchList();
    NVTT_API ~BatchList();
    NVTT_API void Clear();
    NVTT_API void Append(const Surface *img, int face, int mipmap, const OutputOptions *outputOptions);
    NVTT_API void AppendChain(const Surface *img, int face, int mipmap, MipmapFilter filter, int mipmapCount,
                              const OutputOptions *outputOptions);
    NVTT_API unsigned GetSize() const;
    Private *m;

private:
    BatchList(const BatchList &);
    void operator=(const BatchList &);
};

// nvtt.h:413-448 (Surface and raw APIs).
struct Compressor {
    struct Private;
    NVTT_API Compressor();
    NVTT_API ~Compressor();

    // Device selection: enabling takes the first HIP device (the reference's
    // CUDA switch; the method names are kept because they are ABI).
    NVTT_API void enableCudaAcceleration(bool enable);
    NVTT_API bool isCudaAccelerationEnabled() const;
    // Accepted for source compatibility; the device path does not use it.
    NVTT_API void setTaskDispatcher(TaskDispatcher *disp);

    // InputOptions API (nvtt.h:426-428): the whole texture, header first.
    NVTT_API bool process(const InputOptions &inputOptions, const CompressionOptions &compressionOptions,
                          const OutputOptions &outputOptions) const;
    NVTT_API int estimateSize(const InputOptions &inputOptions, const CompressionOptions &compressionOptions) const;

    NVTT_API bool outputHeader(const Surface &img, int mipmapCount, const CompressionOptions &compressionOptions,
                               const OutputOptions &outputOptions) const;
    NVTT_API bool compress(const Surface &img, int face, int mipmap, const CompressionOptions &compressionOptions,
                           const OutputOptions &outputOptions) const;
    NVTT_API int estimateSize(const Surface &img, int mipmapCount, const CompressionOptions &compressionOptions) const;
    NVTT_API bool compress(int w, int h, int d, int face, int mipmap, const float *rgba,
                           const CompressionOptions &compressionOptions, const OutputOptions &outputOptions) const;
    NVTT_API int estimateSize(int w, int h, int d, int mipmapCount, const CompressionOptions &compressionOptions) const;
    // Every image of the list (BatchList above), delivered to each entry's
    // OutputOptions in list order, level by level, with the same
    // beginImage / writeData / endImage / error sequence and the same bytes
    // as compress(Surface) per image.  BC1, BC4 and BC5 run as one batch
    // graph for the whole list (mip chains included); BC3 runs each entry's
    // chain through the chain scheduler.
    NVTT_API bool compress(const BatchList &list, const CompressionOptions &compressionOptions) const;
    Private &m;

private:
    Compressor(const Compressor &);
    void operator=(const Compressor &);
};
typedef Compressor Context;

// nvtt.h:476-600 (the members on this path).  Copies share storage until one
// of them is modified (copy-on-write, Surface.cpp:354-363).
struct Surface {
    NVTT_API Surface();
    NVTT_API Surface(const Surface &img);
    NVTT_API ~Surface();
    NVTT_API void operator=(const Surface &img);

    NVTT_API void setWrapMode(WrapMode mode);
    NVTT_API void setAlphaMode(AlphaMode alphaMode);
    NVTT_API void setNormalMap(bool isNormalMap);

    NVTT_API bool isNull() const;
    NVTT_API int width() const;
    NVTT_API int height() const;
    NVTT_API int depth() const;
    NVTT_API TextureType type() const;
    NVTT_API WrapMode wrapMode() const;
    NVTT_API AlphaMode alphaMode() const;
    NVTT_API bool isNormalMap() const;
    NVTT_API int countMipmaps() const;
    NVTT_API int countMipmaps(int min_size) const;
    NVTT_API const float *data() const;
    NVTT_API const float *channel(int i) const;

    NVTT_API bool setImage(int w, int h, int d);
    NVTT_API bool setImage(InputFormat format, int w, int h, int d, const void *data);
    // Surface.cpp:908-1118: decodes BC1/BC3/BC4/BC5 blocks (D3D10 or D3D9
    // decoder) on the device; other formats and Decoder_NV5x return false.
    NVTT_API bool setImage2D(Format format, Decoder decoder, int w, int h, const void *data);

    NVTT_API bool buildNextMipmap(MipmapFilter filter, int min_size = 1);
    NVTT_API bool buildNextMipmap(MipmapFilter filter, float filterWidth, const float *params = 0, int min_size = 1);
    NVTT_API bool canMakeNextMipmap(int min_size = 1);

    NVTT_API void toLinearFromSrgb();
    // Surface.cpp:1470-1487 (R, G, B; nothing for a gamma of 1).
    NVTT_API void toLinear(float gamma);
    NVTT_API void toGamma(float gamma);

    // The device copy of the planes (4 planes of width*height floats), for
    // callers that keep work on the device.  Null for a null surface.
    NVTT_API const float *deviceData() const;

    struct Private;
    Private *m;
};

// nvtt.h:451-455 (nvtt.cpp:31-59).
NVTT_API const char *errorString(Error e);
NVTT_API unsigned int version();

// nvtt.h:699-700 (Surface.cpp:3270-3280, ErrorMetric.cpp:13-69), on the
// device; FLT_MAX when the layouts differ.
NVTT_API float rmsError(const Surface &reference, const Surface &img);
NVTT_API float rmsAlphaError(const Surface &reference, const Surface &img);

}  // namespace nvtt

#endif  // NVTT_AMD_H
