/* nvtt_wrapper.h -- the C interface of NVTT 2.1.2 (src/nvtt/nvtt_wrapper.h:
 * 213-270) over this library's nvtt.h front end: InputOptions,
 * CompressionOptions, OutputOptions and Compressor handles, with
 * nvttCompress running Compressor::process.  Names, enumerant values and
 * argument order are the reference's (they are ABI). */
#ifndef NVTT_AMD_WRAPPER_H
#define NVTT_AMD_WRAPPER_H

#ifndef __cplusplus
#include <stdbool.h>
#endif

#if defined(__GNUC__)
#define NVTT_WRAPPER_API __attribute__((visibility("default")))
#else
#define NVTT_WRAPPER_API
#endif

#ifdef __cplusplus
namespace nvtt {
struct InputOptions;
struct CompressionOptions;
struct OutputOptions;
struct Compressor;
}  // namespace nvtt
typedef struct nvtt::InputOptions NvttInputOptions;
typedef struct nvtt::CompressionOptions NvttCompressionOptions;
typedef struct nvtt::OutputOptions NvttOutputOptions;
typedef struct nvtt::Compressor NvttCompressor;
#else
typedef struct NvttInputOptions NvttInputOptions;
typedef struct NvttCompressionOptions NvttCompressionOptions;
typedef struct NvttOutputOptions NvttOutputOptions;
typedef struct NvttCompressor NvttCompressor;
#endif

/* nvtt_wrapper.h:50-196 (values as nvtt.h). */
typedef enum {
    NVTT_Format_RGB,
    NVTT_Format_RGBA = NVTT_Format_RGB,
    NVTT_Format_DXT1,
    NVTT_Format_DXT1a,
    NVTT_Format_DXT3,
    NVTT_Format_DXT5,
    NVTT_Format_DXT5n,
    NVTT_Format_BC1 = NVTT_Format_DXT1,
    NVTT_Format_BC1a = NVTT_Format_DXT1a,
    NVTT_Format_BC2 = NVTT_Format_DXT3,
    NVTT_Format_BC3 = NVTT_Format_DXT5,
    NVTT_Format_BC3n = NVTT_Format_DXT5n,
    NVTT_Format_BC4,
    NVTT_Format_BC5,
    NVTT_Format_DXT1n,
    NVTT_Format_CTX1,
    NVTT_Format_BC6,
    NVTT_Format_BC7,
    NVTT_Format_BC3_RGBM,
    NVTT_Format_ETC1,
    NVTT_Format_ETC2_R,
    NVTT_Format_ETC2_RG,
    NVTT_Format_ETC2_RGB,
    NVTT_Format_ETC2_RGBA,
    NVTT_Format_ETC2_RGB_A1,
    NVTT_Format_ETC2_RGBM,
    NVTT_Format_PVR_2BPP_RGB,
    NVTT_Format_PVR_4BPP_RGB,
    NVTT_Format_PVR_2BPP_RGBA,
    NVTT_Format_PVR_4BPP_RGBA
} NvttFormat;

typedef enum { NVTT_Quality_Fastest, NVTT_Quality_Normal, NVTT_Quality_Production, NVTT_Quality_Highest } NvttQuality;
typedef enum { NVTT_WrapMode_Clamp, NVTT_WrapMode_Repeat, NVTT_WrapMode_Mirror } NvttWrapMode;
/* The reference spells the last two without the NVTT_ prefix
 * (nvtt_wrapper.h:131-135); both spellings name the same values. */
typedef enum {
    NVTT_TextureType_2D,
    NVTT_TextureType_Cube,
    TextureType_3D,
    TextureType_Array,
    NVTT_TextureType_3D = TextureType_3D,
    NVTT_TextureType_Array = TextureType_Array
} NvttTextureType;
typedef enum {
    NVTT_InputFormat_BGRA_8UB,
    NVTT_InputFormat_RGBA_16F,
    NVTT_InputFormat_RGBA_32F,
    NVTT_InputFormat_R_32F
} NvttInputFormat;
typedef enum { NVTT_MipmapFilter_Box, NVTT_MipmapFilter_Triangle, NVTT_MipmapFilter_Kaiser } NvttMipmapFilter;
typedef enum {
    NVTT_RoundMode_None,
    NVTT_RoundMode_ToNextPowerOfTwo,
    NVTT_RoundMode_ToNearestPowerOfTwo,
    NVTT_RoundMode_ToPreviousPowerOfTwo,
    NVTT_RoundMode_ToNextMultipleOfFour,
    NVTT_RoundMode_ToNearestMultipleOfFour,
    NVTT_RoundMode_ToPreviousMultipleOfFour
} NvttRoundMode;
typedef enum { NVTT_AlphaMode_None, NVTT_AlphaMode_Transparency, NVTT_AlphaMode_Premultiplied } NvttAlphaMode;
typedef enum {
    NVTT_Error_Unknown,
    NVTT_Error_InvalidInput,
    NVTT_Error_UnsupportedFeature,
    NVTT_Error_CudaError,
    NVTT_Error_FileOpen,
    NVTT_Error_FileWrite,
    NVTT_Error_UnsupportedOutputFormat
} NvttError;
typedef enum { NVTT_Container_DDS, NVTT_Container_DDS10, NVTT_Container_KTX } NvttContainer;
typedef enum { NVTT_False, NVTT_True } NvttBoolean;

#ifdef __cplusplus
extern "C" {
#endif

/* nvtt_wrapper.h:200-206 callbacks. */
typedef void (*nvttErrorHandler)(NvttError e);
typedef void (*nvttBeginImageHandler)(int size, int width, int height, int depth, int face, int miplevel);
typedef bool (*nvttOutputHandler)(const void *data, int size);
typedef void (*nvttEndImageHandler)(void);

/* InputOptions (nvtt_wrapper.h:213-232). */
NVTT_WRAPPER_API NvttInputOptions *nvttCreateInputOptions(void);
NVTT_WRAPPER_API void nvttDestroyInputOptions(NvttInputOptions *inputOptions);
NVTT_WRAPPER_API void nvttSetInputOptionsTextureLayout(NvttInputOptions *inputOptions, NvttTextureType type, int w, int h,
                                                       int d, int arraySize);
NVTT_WRAPPER_API void nvttResetInputOptionsTextureLayout(NvttInputOptions *inputOptions);
NVTT_WRAPPER_API NvttBoolean nvttSetInputOptionsMipmapData(NvttInputOptions *inputOptions, const void *data, int w, int h,
                                                           int d, int face, int mipmap);
NVTT_WRAPPER_API void nvttSetInputOptionsFormat(NvttInputOptions *inputOptions, NvttInputFormat format);
NVTT_WRAPPER_API void nvttSetInputOptionsAlphaMode(NvttInputOptions *inputOptions, NvttAlphaMode alphaMode);
NVTT_WRAPPER_API void nvttSetInputOptionsGamma(NvttInputOptions *inputOptions, float inputGamma, float outputGamma);
NVTT_WRAPPER_API void nvttSetInputOptionsWrapMode(NvttInputOptions *inputOptions, NvttWrapMode mode);
NVTT_WRAPPER_API void nvttSetInputOptionsMipmapFilter(NvttInputOptions *inputOptions, NvttMipmapFilter filter);
NVTT_WRAPPER_API void nvttSetInputOptionsMipmapGeneration(NvttInputOptions *inputOptions, NvttBoolean enabled, int maxLevel);
NVTT_WRAPPER_API void nvttSetInputOptionsKaiserParameters(NvttInputOptions *inputOptions, float width, float alpha,
                                                          float stretch);
NVTT_WRAPPER_API void nvttSetInputOptionsNormalMap(NvttInputOptions *inputOptions, NvttBoolean b);
NVTT_WRAPPER_API void nvttSetInputOptionsConvertToNormalMap(NvttInputOptions *inputOptions, NvttBoolean convert);
NVTT_WRAPPER_API void nvttSetInputOptionsHeightEvaluation(NvttInputOptions *inputOptions, float redScale, float greenScale,
                                                          float blueScale, float alphaScale);
NVTT_WRAPPER_API void nvttSetInputOptionsNormalFilter(NvttInputOptions *inputOptions, float sm, float medium, float big,
                                                      float large);
NVTT_WRAPPER_API void nvttSetInputOptionsNormalizeMipmaps(NvttInputOptions *inputOptions, NvttBoolean b);
NVTT_WRAPPER_API void nvttSetInputOptionsMaxExtents(NvttInputOptions *inputOptions, int dim);
NVTT_WRAPPER_API void nvttSetInputOptionsRoundMode(NvttInputOptions *inputOptions, NvttRoundMode mode);

/* CompressionOptions (nvtt_wrapper.h:236-243). */
NVTT_WRAPPER_API NvttCompressionOptions *nvttCreateCompressionOptions(void);
NVTT_WRAPPER_API void nvttDestroyCompressionOptions(NvttCompressionOptions *compressionOptions);
NVTT_WRAPPER_API void nvttSetCompressionOptionsFormat(NvttCompressionOptions *compressionOptions, NvttFormat format);
NVTT_WRAPPER_API void nvttSetCompressionOptionsQuality(NvttCompressionOptions *compressionOptions, NvttQuality quality);
NVTT_WRAPPER_API void nvttSetCompressionOptionsColorWeights(NvttCompressionOptions *compressionOptions, float red,
                                                            float green, float blue, float alpha);
NVTT_WRAPPER_API void nvttSetCompressionOptionsPixelFormat(NvttCompressionOptions *compressionOptions, unsigned int bitcount,
                                                           unsigned int rmask, unsigned int gmask, unsigned int bmask,
                                                           unsigned int amask);
NVTT_WRAPPER_API void nvttSetCompressionOptionsQuantization(NvttCompressionOptions *compressionOptions,
                                                            NvttBoolean colorDithering, NvttBoolean alphaDithering,
                                                            NvttBoolean binaryAlpha, int alphaThreshold);

/* OutputOptions (nvtt_wrapper.h:247-255). */
NVTT_WRAPPER_API NvttOutputOptions *nvttCreateOutputOptions(void);
NVTT_WRAPPER_API void nvttDestroyOutputOptions(NvttOutputOptions *outputOptions);
NVTT_WRAPPER_API void nvttSetOutputOptionsFileName(NvttOutputOptions *outputOptions, const char *fileName);
NVTT_WRAPPER_API void nvttSetOutputOptionsOutputHeader(NvttOutputOptions *outputOptions, NvttBoolean b);
NVTT_WRAPPER_API void nvttSetOutputOptionsContainer(NvttOutputOptions *outputOptions, NvttContainer containerFormat);
NVTT_WRAPPER_API void nvttSetOutputOptionsSrgbFlag(NvttOutputOptions *outputOptions, NvttBoolean b);
NVTT_WRAPPER_API void nvttSetOutputOptionsErrorHandler(NvttOutputOptions *outputOptions, nvttErrorHandler errorHandler);
NVTT_WRAPPER_API void nvttSetOutputOptionsOutputHandler(NvttOutputOptions *outputOptions,
                                                        nvttBeginImageHandler beginImageHandler,
                                                        nvttOutputHandler outputHandler,
                                                        nvttEndImageHandler endImageHandler);

/* Compressor (nvtt_wrapper.h:259-265). */
NVTT_WRAPPER_API NvttCompressor *nvttCreateCompressor(void);
NVTT_WRAPPER_API void nvttDestroyCompressor(NvttCompressor *compressor);
NVTT_WRAPPER_API void nvttEnableCudaAcceleration(NvttCompressor *compressor, NvttBoolean b);
NVTT_WRAPPER_API NvttBoolean nvttIsCudaAccelerationEnabled(const NvttCompressor *compressor);
NVTT_WRAPPER_API NvttBoolean nvttCompress(const NvttCompressor *compressor, const NvttInputOptions *inputOptions,
                                          const NvttCompressionOptions *compressionOptions,
                                          const NvttOutputOptions *outputOptions);
NVTT_WRAPPER_API int nvttEstimateSize(const NvttCompressor *compressor, const NvttInputOptions *inputOptions,
                                      const NvttCompressionOptions *compressionOptions);

/* Global functions (nvtt_wrapper.h:269-270). */
NVTT_WRAPPER_API const char *nvttErrorString(NvttError e);
NVTT_WRAPPER_API unsigned int nvttVersion(void);

#ifdef __cplusplus
}
#endif

#endif /* NVTT_AMD_WRAPPER_H */
