"""ctypes binding of the gfx950 backend's C ABI (include/ntx_hip.h).

This is how the Python tests and bench.py reach the HIP kernels: every call
goes through lib/libntx_hip.so.  There is no fallback: if the library is not
built, or no HIP device is present, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libntx_hip.so")
NVTT_LIB_PATH = os.path.join(LIB_DIR, "libnvtt_amd.so")

# nvtt::Format / Quality / AlphaMode / MipmapFilter / WrapMode values (nvtt.h).
FORMAT_BC1, FORMAT_BC3, FORMAT_BC4, FORMAT_BC5 = 1, 4, 6, 7
FORMAT_BC3N = 5  # Format_BC3n (DXT5n): X in the alpha block, Y in the colour block
FORMAT_BC1A = 2  # Format_BC1a (DXT1a): BC1 with the transparent index for alpha-0 texels
QUALITY_FASTEST, QUALITY_NORMAL, QUALITY_PRODUCTION, QUALITY_HIGHEST = 0, 1, 2, 3
DECODER_D3D10, DECODER_D3D9, DECODER_NV5X = 0, 1, 2
METRIC_RMS_COLOR, METRIC_RMS_COLOR_ALPHA_WEIGHTED, METRIC_RMS_ALPHA = 0, 1, 2
ALPHA_NONE, ALPHA_TRANSPARENCY = 0, 1
FILTER_BOX, FILTER_TRIANGLE, FILTER_KAISER = 0, 1, 2
WRAP_CLAMP, WRAP_REPEAT, WRAP_MIRROR = 0, 1, 2

# Surface.cpp:1121-1141 getDefaultFilterWidthAndParams.
DEFAULT_FILTER = {FILTER_BOX: (0.5, None), FILTER_TRIANGLE: (1.0, None), FILTER_KAISER: (3.0, (4.0, 1.0))}

# ntx_hip.h kernel ids for ntx_hip_profile_read.
KERNELS = ("bc1_kernel", "bc3_search_kernel", "channel_kernel", "box_kernel", "poly_x_kernel", "poly_y_kernel",
           "srgb_kernel", "unpack_kernel", "poly_fused_kernel", "bc3_front_kernel", "batch_kernel",
           "bc1_mlp_kernel", "bc1_list_kernel", "mlp_train_kernel", "bc3_search_split_kernel", "gamma_kernel",
           "normal_kernel", "bc1_classify_kernel", "bc1_search_kernel", "optimal_channel_kernel",
           "optimal_green_kernel", "bc3_fast_kernel", "bc1a_front_kernel+bc3_search_kernel", "mip_tail_kernel",
           "poly_slide_kernel")

# Every symbol include/ntx_hip.h declares.
EXPORTS = (
    "ntx_hip_is_hardware_present", "ntx_hip_device_count", "ntx_hip_status_string",
    "ntx_hip_create", "ntx_hip_destroy", "ntx_hip_last_error", "ntx_hip_stream",
    "ntx_hip_malloc", "ntx_hip_free", "ntx_hip_upload", "ntx_hip_download", "ntx_hip_synchronize",
    "ntx_hip_copy_device", "ntx_hip_memset_device",
    "ntx_hip_timer_begin", "ntx_hip_timer_end", "ntx_hip_kernel_name", "ntx_hip_profile_enable",
    "ntx_hip_profile_read", "ntx_hip_compressed_size",
    "ntx_hip_compress", "ntx_hip_compress_device", "ntx_hip_compress_bc1_stats_device", "ntx_hip_set_bc1_lanes_min",
    "ntx_hip_compress_device_noalloc", "ntx_hip_set_scratch_limit",
    "ntx_hip_compress_chain_device", "ntx_hip_compress_images_device",
    "ntx_hip_downsample", "ntx_hip_downsample_device", "ntx_hip_downsample_chain_device",
    "ntx_hip_srgb_to_linear_device", "ntx_hip_unpack_bgra8_device", "ntx_hip_unpack_bgra8_srgb_device",
    "ntx_hip_gamma_device", "ntx_hip_gamma_levels_device", "ntx_hip_unpack_bgra8_gamma_device",
    "ntx_hip_normalize_normals_device", "ntx_hip_normalize_vectors_device", "ntx_hip_decode_device",
    "ntx_hip_rms_error_device",
    "ntx_hip_batch_create", "ntx_hip_batch_destroy", "ntx_hip_batch_submit", "ntx_hip_batch_wait",
    "ntx_hip_batch_levels", "ntx_hip_batch_level", "ntx_hip_batch_get_tuning", "ntx_hip_batch_set_tuning",
    "ntx_hip_batch_autotune",
    "ntx_hip_mlp_create", "ntx_hip_mlp_destroy", "ntx_hip_mlp_get_params", "ntx_hip_mlp_set_params",
    "ntx_hip_mlp_features_device", "ntx_hip_mlp_labels_device", "ntx_hip_mlp_forward_device", "ntx_hip_mlp_train_step",
    "ntx_hip_compress_bc1_mlp_device", "ntx_hip_compress_bc1_mlp_images_device", "ntx_hip_bc1_sse_device",
)

MLP_PARAMS = 8921          # NTX_MLP_PARAMS
MLP_BOX_CANDIDATES = 1     # NTX_MLP_BOX_CANDIDATES


class MlpSettings(ctypes.Structure):
    """ntx_mlp_settings: one image's settings for ntx_hip_compress_bc1_mlp_images_device."""
    _fields_ = [("tau_pred", ctypes.c_float), ("tau_gap", ctypes.c_float), ("refine_iters", ctypes.c_int),
                ("greedy_rounds", ctypes.c_int), ("flags", ctypes.c_int), ("plain", ctypes.c_int)]

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_S = ctypes.c_size_t
_lib = None


class TextureDesc(ctypes.Structure):
    """ntx_texture_desc (include/ntx_hip.h)."""
    _fields_ = [("d_rgba", ctypes.c_void_p), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("format", ctypes.c_int), ("quality", ctypes.c_int), ("alpha_mode", ctypes.c_int),
                ("color_weights", ctypes.c_float * 4), ("mip_levels", ctypes.c_int), ("filter", ctypes.c_int),
                ("filter_width", ctypes.c_float), ("filter_params", ctypes.c_float * 2), ("wrap", ctypes.c_int)]


class BatchTuning(ctypes.Structure):
    """ntx_batch_tuning (include/ntx_hip.h)."""
    _fields_ = [("bc1_blocks_per_item", ctypes.c_int), ("channel_jobs_per_item", ctypes.c_int),
                ("mip_tiles_per_item", ctypes.c_int), ("mip_tile_width", ctypes.c_int),
                ("kernel_variant", ctypes.c_int), ("workgroups", ctypes.c_int), ("schedule", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


class NtxError(RuntimeError):
    def __init__(self, status: int, what: str, detail: str = ""):
        self.status = status
        super().__init__(f"{what} failed: status {status}" + (f" ({detail})" if detail else ""))


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libntx_hip.so and declare its signatures; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build() first")
    lib = ctypes.CDLL(path)
    sig = {
        "ntx_hip_is_hardware_present": (_I, []),
        "ntx_hip_device_count": (_I, [ctypes.POINTER(_I)]),
        "ntx_hip_status_string": (ctypes.c_char_p, [_I]),
        "ntx_hip_create": (_I, [_I, ctypes.POINTER(_P)]),
        "ntx_hip_destroy": (_I, [_P]),
        "ntx_hip_last_error": (ctypes.c_char_p, [_P]),
        "ntx_hip_stream": (_P, [_P]),
        "ntx_hip_malloc": (_I, [_P, _S, ctypes.POINTER(_P)]),
        "ntx_hip_free": (_I, [_P, _P]),
        "ntx_hip_upload": (_I, [_P, _P, _P, _S]),
        "ntx_hip_download": (_I, [_P, _P, _P, _S]),
        "ntx_hip_synchronize": (_I, [_P]),
        "ntx_hip_copy_device": (_I, [_P, _P, _P, _S]),
        "ntx_hip_memset_device": (_I, [_P, _P, _I, _S]),
        "ntx_hip_timer_begin": (_I, [_P]),
        "ntx_hip_timer_end": (_I, [_P, ctypes.POINTER(_F)]),
        "ntx_hip_kernel_name": (ctypes.c_char_p, [_I]),
        "ntx_hip_profile_enable": (_I, [_P, _I]),
        "ntx_hip_profile_read": (_I, [_P, _P, _P, _I]),
        "ntx_hip_compressed_size": (_S, [_I, _I, _I]),
        "ntx_hip_compress": (_I, [_P, _I, _I, _I, _P, _P, _I, _I, _P, _S]),
        "ntx_hip_compress_device": (_I, [_P, _I, _I, _I, _P, _P, _I, _I, _P]),
        "ntx_hip_compress_bc1_stats_device": (_I, [_P, _I, _P, _P, _I, _I, _P, _P]),
        "ntx_hip_set_bc1_lanes_min": (_I, [_P, _I]),
        "ntx_hip_compress_device_noalloc": (_I, [_P, _I, _I, _I, _P, _P, _I, _I, _P]),
        "ntx_hip_set_scratch_limit": (_I, [_P, _S]),
        "ntx_hip_compress_chain_device": (_I, [_P, _I, _I, _I, _P, _I, _F, _P, _I, _I, _I, _I, _P, _P]),
        "ntx_hip_compress_images_device": (_I, [_P, _I, _I, _I, _P, _I, _P, _P, _P, _P]),
        "ntx_hip_downsample": (_I, [_P, _I, _F, _P, _I, _I, _P, _I, _I, _P]),
        "ntx_hip_downsample_device": (_I, [_P, _I, _F, _P, _I, _I, _P, _I, _I, _P]),
        "ntx_hip_downsample_chain_device": (_I, [_P, _I, _F, _P, _I, _I, _I, _I, _I, _P]),
        "ntx_hip_srgb_to_linear_device": (_I, [_P, _P, _I, _I]),
        "ntx_hip_unpack_bgra8_device": (_I, [_P, _P, _I, _I, _P]),
        "ntx_hip_unpack_bgra8_srgb_device": (_I, [_P, _P, _I, _I, _P]),
        "ntx_hip_gamma_device": (_I, [_P, _I, _F, _P, _I, _I, _P]),
        "ntx_hip_gamma_levels_device": (_I, [_P, _I, _F, _I, _P, _P, _P, _P]),
        "ntx_hip_unpack_bgra8_gamma_device": (_I, [_P, _P, _I, _I, _F, _P]),
        "ntx_hip_normalize_normals_device": (_I, [_P, _P, _I, _I]),
        "ntx_hip_normalize_vectors_device": (_I, [_P, _P, _I, _I]),
        "ntx_hip_decode_device": (_I, [_P, _I, _I, _P, _I, _I, _P]),
        "ntx_hip_rms_error_device": (_I, [_P, _I, _P, _P, _I, _I, ctypes.POINTER(_F)]),
        "ntx_hip_batch_create": (_I, [_P, _P, _I, ctypes.POINTER(_P)]),
        "ntx_hip_batch_destroy": (_I, [_P]),
        "ntx_hip_batch_submit": (_I, [_P]),
        "ntx_hip_batch_wait": (_I, [_P]),
        "ntx_hip_batch_levels": (_I, [_P, _I]),
        "ntx_hip_batch_level": (_I, [_P, _I, _I, ctypes.POINTER(_P), ctypes.POINTER(_I), ctypes.POINTER(_I),
                                     ctypes.POINTER(_P), ctypes.POINTER(_S)]),
        "ntx_hip_batch_get_tuning": (_I, [_P, ctypes.POINTER(BatchTuning)]),
        "ntx_hip_batch_set_tuning": (_I, [_P, ctypes.POINTER(BatchTuning)]),
        "ntx_hip_batch_autotune": (_I, [_P, _I, ctypes.POINTER(BatchTuning), ctypes.POINTER(_F)]),
        "ntx_hip_mlp_create": (_I, [_P, _P, ctypes.c_uint32, ctypes.POINTER(_P)]),
        "ntx_hip_mlp_destroy": (_I, [_P]),
        "ntx_hip_mlp_get_params": (_I, [_P, _P]),
        "ntx_hip_mlp_set_params": (_I, [_P, _P]),
        "ntx_hip_mlp_features_device": (_I, [_P, _P, _I, _I, _P, _P]),
        "ntx_hip_mlp_labels_device": (_I, [_P, _P, _I, _I, _P, _P, _P]),
        "ntx_hip_mlp_forward_device": (_I, [_P, _P, _I, _P]),
        "ntx_hip_mlp_train_step": (_I, [_P, _P, _P, _I, _F, ctypes.POINTER(_F), _P]),
        "ntx_hip_compress_bc1_mlp_device": (_I, [_P, _P, _I, _I, _F, _F, _I, _I, _I, _P, ctypes.POINTER(_I)]),
        "ntx_hip_compress_bc1_mlp_images_device": (_I, [_P, _I, _P, _P, _P, _P, _P]),
        "ntx_hip_bc1_sse_device": (_I, [_P, _P, _I, _I, _P, ctypes.POINTER(ctypes.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def compressed_size(fmt: int, w: int, h: int) -> int:
    bs = 8 if fmt in (FORMAT_BC1, FORMAT_BC1A, FORMAT_BC4) else 16
    return ((w + 3) // 4) * ((h + 3) // 4) * bs


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _f32c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _need(buf: "DeviceBuffer", nbytes: int, what: str) -> None:
    """Raises ValueError unless buf is a live allocation of at least nbytes:
    every device entry point reads or writes exactly the extent its sizes
    imply, so a short buffer would be an out-of-bounds device access."""
    if not isinstance(buf, DeviceBuffer) or not buf.ptr:
        raise ValueError(f"{what}: not a live DeviceBuffer")
    if int(nbytes) > buf.nbytes:
        raise ValueError(f"{what}: the call needs {int(nbytes)} bytes, the buffer holds {buf.nbytes}")


def _dims(w: int, h: int, what: str) -> None:
    if int(w) < 0 or int(h) < 0:
        raise ValueError(f"{what}: negative extent {w}x{h}")


def _planar(w: int, h: int) -> int:
    return 16 * int(w) * int(h)


class DeviceBuffer:
    """A device allocation owned by a Context."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _P()
        ctx._check(ctx.lib.ntx_hip_malloc(ctx.handle, self.nbytes, ctypes.byref(p)), "ntx_hip_malloc")
        self.ptr = p.value

    def upload(self, host: np.ndarray) -> "DeviceBuffer":
        host = np.ascontiguousarray(host)
        _need(self, host.nbytes, "upload")
        self.ctx._check(self.ctx.lib.ntx_hip_upload(self.ctx.handle, self.ptr, _ptr(host), host.nbytes), "upload")
        return self

    def download(self, dtype, count: Optional[int] = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else int(count)
        if count < 0:
            raise ValueError("download: negative count")
        _need(self, count * dtype.itemsize, "download")
        out = np.empty(count, dtype)
        self.ctx._check(self.ctx.lib.ntx_hip_download(self.ctx.handle, _ptr(out), self.ptr, out.nbytes), "download")
        return out

    def free(self) -> None:
        if self.ptr:
            self.ctx.lib.ntx_hip_free(self.ctx.handle, self.ptr)
            self.ptr = None


class Context:
    """One HIP stream + device tables (ntx_hip_create)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _P()
        st = self.lib.ntx_hip_create(device, ctypes.byref(h))
        if st != 0:
            raise NtxError(st, "ntx_hip_create", self.lib.ntx_hip_status_string(st).decode())
        self.handle = h.value

    def close(self) -> None:
        if self.handle:
            self.lib.ntx_hip_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int, what: str) -> None:
        if st != 0:
            detail = self.lib.ntx_hip_last_error(self.handle).decode() if self.handle else ""
            raise NtxError(st, what, detail or self.lib.ntx_hip_status_string(st).decode())

    # -- host-pointer entry points ------------------------------------------
    def compress(self, fmt: int, rgba_planar: np.ndarray, w: int, h: int, quality: int = QUALITY_NORMAL,
                 alpha_mode: int = ALPHA_NONE, color_weights: Sequence[float] = (1, 1, 1, 1)) -> np.ndarray:
        _dims(w, h, "compress")
        src = _f32c(rgba_planar).reshape(-1)
        if src.size != 4 * w * h:
            raise ValueError(f"compress: {src.size} floats for a {w}x{h} planar RGBA image")
        out = np.zeros(compressed_size(fmt, w, h), np.uint8)
        cw = _f32c(color_weights)
        st = self.lib.ntx_hip_compress(self.handle, fmt, quality, alpha_mode, _ptr(cw), _ptr(src), w, h,
                                       _ptr(out), out.nbytes)
        self._check(st, "ntx_hip_compress")
        return out

    def downsample(self, filt: int, rgba_planar: np.ndarray, W: int, H: int, width: Optional[float] = None,
                   params: Optional[Sequence[float]] = None, wrap: int = WRAP_MIRROR,
                   alpha_mode: int = ALPHA_NONE) -> np.ndarray:
        dw, dp = DEFAULT_FILTER[filt]
        width = dw if width is None else width
        params = dp if params is None else params
        src = _f32c(rgba_planar).reshape(-1)
        if src.size != 4 * W * H:
            raise ValueError(f"downsample: {src.size} floats for a {W}x{H} planar RGBA image")
        w, h = max(1, W // 2), max(1, H // 2)
        out = np.empty(4 * w * h, np.float32)
        pp = _f32c(params) if params is not None else None
        st = self.lib.ntx_hip_downsample(self.handle, filt, width, _ptr(pp) if pp is not None else None, wrap,
                                         alpha_mode, _ptr(src), W, H, _ptr(out))
        self._check(st, "ntx_hip_downsample")
        return out

    # -- device-pointer entry points -----------------------------------------
    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def compress_device(self, fmt: int, d_rgba: DeviceBuffer, w: int, h: int, d_out: DeviceBuffer,
                        quality: int = QUALITY_NORMAL, alpha_mode: int = ALPHA_NONE,
                        color_weights: Sequence[float] = (1, 1, 1, 1)) -> None:
        _dims(w, h, "compress_device")
        _need(d_rgba, _planar(w, h), "compress_device input")
        _need(d_out, compressed_size(fmt, w, h), "compress_device output")
        cw = _f32c(color_weights)
        st = self.lib.ntx_hip_compress_device(self.handle, fmt, quality, alpha_mode, _ptr(cw), d_rgba.ptr, w, h,
                                              d_out.ptr)
        self._check(st, "ntx_hip_compress_device")

    def compress_chain_device(self, fmt: int, filt: int, d_levels: Sequence[DeviceBuffer], w: int, h: int,
                              d_out: Sequence[DeviceBuffer], quality: int = QUALITY_NORMAL,
                              alpha_mode: int = ALPHA_NONE, color_weights: Sequence[float] = (1, 1, 1, 1),
                              width: Optional[float] = None, params: Optional[Sequence[float]] = None,
                              wrap: int = WRAP_MIRROR) -> None:
        """Level 0 in d_levels[0]; levels 1.. are written to d_levels[1..] and
        every level's blocks to d_out[k] (ntx_hip_compress_chain_device)."""
        if len(d_levels) != len(d_out):
            raise ValueError("one output buffer per level")
        _dims(w, h, "compress_chain_device")
        for k, (a, b) in enumerate(mip_sizes(w, h)[:len(d_levels)]):
            _need(d_levels[k], _planar(a, b), f"compress_chain_device level {k}")
            _need(d_out[k], compressed_size(fmt, a, b), f"compress_chain_device output {k}")
        dw, dp = DEFAULT_FILTER[filt]
        width = dw if width is None else width
        params = dp if params is None else params
        pp = _f32c(params) if params is not None else None
        cw = _f32c(color_weights)
        lv = (_P * len(d_levels))(*[d.ptr for d in d_levels])
        out = (_P * len(d_out))(*[d.ptr for d in d_out])
        st = self.lib.ntx_hip_compress_chain_device(self.handle, fmt, quality, alpha_mode, _ptr(cw), filt, width,
                                                    _ptr(pp) if pp is not None else None, wrap, w, h, len(d_levels),
                                                    ctypes.cast(lv, _P), ctypes.cast(out, _P))
        self._check(st, "ntx_hip_compress_chain_device")

    def compress_images_device(self, fmt: int, d_rgba: Sequence[DeviceBuffer], sizes: Sequence[Tuple[int, int]],
                               d_out: Sequence[DeviceBuffer], quality: int = QUALITY_NORMAL,
                               alpha_mode: int = ALPHA_NONE, color_weights: Sequence[float] = (1, 1, 1, 1)) -> None:
        """Independent images compressed together (ntx_hip_compress_images_device)."""
        n = len(d_rgba)
        if len(sizes) != n or len(d_out) != n:
            raise ValueError("one size and one output buffer per image")
        for k, (a, b) in enumerate(sizes):
            _dims(a, b, "compress_images_device")
            _need(d_rgba[k], _planar(a, b), f"compress_images_device input {k}")
            _need(d_out[k], compressed_size(fmt, a, b), f"compress_images_device output {k}")
        cw = _f32c(color_weights)
        ins = (_P * n)(*[d.ptr for d in d_rgba])
        outs = (_P * n)(*[d.ptr for d in d_out])
        ws = np.array([a for a, _ in sizes], np.int32)
        hs = np.array([b for _, b in sizes], np.int32)
        st = self.lib.ntx_hip_compress_images_device(self.handle, fmt, quality, alpha_mode, _ptr(cw), n,
                                                     ctypes.cast(ins, _P), _ptr(ws), _ptr(hs), ctypes.cast(outs, _P))
        self._check(st, "ntx_hip_compress_images_device")

    def compress_device_noalloc(self, fmt: int, d_rgba: DeviceBuffer, w: int, h: int, d_out: DeviceBuffer,
                                quality: int = QUALITY_NORMAL, alpha_mode: int = ALPHA_NONE,
                                color_weights: Sequence[float] = (1, 1, 1, 1)) -> None:
        """compress_device without any device allocation (the library's own
        scratch array; images needing it are limited to 65,536 blocks)."""
        _dims(w, h, "compress_device_noalloc")
        _need(d_rgba, _planar(w, h), "compress_device_noalloc input")
        _need(d_out, compressed_size(fmt, w, h), "compress_device_noalloc output")
        cw = _f32c(color_weights)
        st = self.lib.ntx_hip_compress_device_noalloc(self.handle, fmt, quality, alpha_mode, _ptr(cw), d_rgba.ptr, w, h,
                                                      d_out.ptr)
        self._check(st, "ntx_hip_compress_device_noalloc")

    def set_bc1_lanes_min(self, min_blocks: int) -> None:
        """BC1 Level8 images of at least min_blocks blocks take the two-pass
        one-block-per-lane encoder (0: the default threshold)."""
        self._check(self.lib.ntx_hip_set_bc1_lanes_min(self.handle, int(min_blocks)), "set_bc1_lanes_min")

    def compress_bc1_stats_device(self, d_rgba: DeviceBuffer, w: int, h: int, d_out: DeviceBuffer,
                                  d_nm: DeviceBuffer, alpha_mode: int = ALPHA_NONE,
                                  color_weights: Sequence[float] = (1, 1, 1, 1)) -> None:
        _dims(w, h, "compress_bc1_stats_device")
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        _need(d_rgba, _planar(w, h), "compress_bc1_stats_device input")
        _need(d_out, 8 * nb, "compress_bc1_stats_device output")
        _need(d_nm, 4 * nb, "compress_bc1_stats_device counts")
        cw = _f32c(color_weights)
        st = self.lib.ntx_hip_compress_bc1_stats_device(self.handle, alpha_mode, _ptr(cw), d_rgba.ptr, w, h,
                                                        d_out.ptr, d_nm.ptr)
        self._check(st, "ntx_hip_compress_bc1_stats_device")

    def downsample_device(self, filt: int, d_src: DeviceBuffer, W: int, H: int, d_dst: DeviceBuffer,
                          width: Optional[float] = None, params: Optional[Sequence[float]] = None,
                          wrap: int = WRAP_MIRROR, alpha_mode: int = ALPHA_NONE) -> None:
        dw, dp = DEFAULT_FILTER[filt]
        width = dw if width is None else width
        params = dp if params is None else params
        _dims(W, H, "downsample_device")
        _need(d_src, _planar(W, H), "downsample_device input")
        _need(d_dst, _planar(max(1, W // 2), max(1, H // 2)), "downsample_device output")
        pp = _f32c(params) if params is not None else None
        st = self.lib.ntx_hip_downsample_device(self.handle, filt, width, _ptr(pp) if pp is not None else None,
                                                wrap, alpha_mode, d_src.ptr, W, H, d_dst.ptr)
        self._check(st, "ntx_hip_downsample_device")

    def downsample_chain_device(self, filt: int, d_levels: Sequence[DeviceBuffer], w: int, h: int,
                                width: Optional[float] = None, params: Optional[Sequence[float]] = None,
                                wrap: int = WRAP_MIRROR, alpha_mode: int = ALPHA_NONE) -> None:
        """Levels 1.. of the chain from d_levels[0] (ntx_hip_downsample_chain_device)."""
        dw, dp = DEFAULT_FILTER[filt]
        width = dw if width is None else width
        params = dp if params is None else params
        _dims(w, h, "downsample_chain_device")
        for k, (a, b) in enumerate(mip_sizes(w, h)[:len(d_levels)]):
            _need(d_levels[k], _planar(a, b), f"downsample_chain_device level {k}")
        pp = _f32c(params) if params is not None else None
        lv = (_P * len(d_levels))(*[d.ptr for d in d_levels])
        st = self.lib.ntx_hip_downsample_chain_device(self.handle, filt, width, _ptr(pp) if pp is not None else None,
                                                      wrap, alpha_mode, w, h, len(d_levels), ctypes.cast(lv, _P))
        self._check(st, "ntx_hip_downsample_chain_device")

    def srgb_to_linear_device(self, d_rgba: DeviceBuffer, w: int, h: int) -> None:
        _dims(w, h, "srgb_to_linear_device")
        _need(d_rgba, _planar(w, h), "srgb_to_linear_device")
        self._check(self.lib.ntx_hip_srgb_to_linear_device(self.handle, d_rgba.ptr, w, h), "srgb_to_linear")

    def unpack_bgra8_device(self, d_bgra: DeviceBuffer, w: int, h: int, d_rgba: DeviceBuffer) -> None:
        self._need_unpack(d_bgra, w, h, d_rgba, "unpack_bgra8_device")
        self._check(self.lib.ntx_hip_unpack_bgra8_device(self.handle, d_bgra.ptr, w, h, d_rgba.ptr), "unpack")

    def unpack_bgra8_srgb_device(self, d_bgra: DeviceBuffer, w: int, h: int, d_rgba: DeviceBuffer) -> None:
        """unpack_bgra8_device followed by srgb_to_linear_device, in one pass."""
        self._need_unpack(d_bgra, w, h, d_rgba, "unpack_bgra8_srgb_device")
        self._check(self.lib.ntx_hip_unpack_bgra8_srgb_device(self.handle, d_bgra.ptr, w, h, d_rgba.ptr), "unpack_srgb")

    def gamma_device(self, to_gamma: bool, gamma: float, d_src: DeviceBuffer, w: int, h: int,
                     d_dst: Optional[DeviceBuffer] = None) -> None:
        """FloatImage::toLinear / toGamma on R, G, B (ntx_hip_gamma_device); in place without d_dst."""
        dst = d_src if d_dst is None else d_dst
        _dims(w, h, "gamma_device")
        _need(d_src, _planar(w, h), "gamma_device input")
        _need(dst, _planar(w, h), "gamma_device output")
        self._check(self.lib.ntx_hip_gamma_device(self.handle, int(bool(to_gamma)), gamma, d_src.ptr, w, h, dst.ptr),
                    "ntx_hip_gamma_device")

    def gamma_levels_device(self, to_gamma: bool, gamma: float, d_src: Sequence[DeviceBuffer],
                            sizes: Sequence[Tuple[int, int]], d_dst: Sequence[DeviceBuffer]) -> None:
        """gamma_device over several images in one launch (ntx_hip_gamma_levels_device)."""
        n = len(d_src)
        if len(sizes) != n or len(d_dst) != n:
            raise ValueError("gamma_levels_device: one size and one output per level")
        for k, (a, b) in enumerate(sizes):
            _dims(a, b, "gamma_levels_device")
            _need(d_src[k], _planar(a, b), f"gamma_levels_device input {k}")
            _need(d_dst[k], _planar(a, b), f"gamma_levels_device output {k}")
        src = (_P * n)(*[d.ptr for d in d_src])
        dst = (_P * n)(*[d.ptr for d in d_dst])
        ws = (_I * n)(*[int(w) for w, _ in sizes])
        hs = (_I * n)(*[int(h) for _, h in sizes])
        st = self.lib.ntx_hip_gamma_levels_device(self.handle, int(bool(to_gamma)), gamma, n, ctypes.cast(src, _P),
                                                  ctypes.cast(ws, _P), ctypes.cast(hs, _P), ctypes.cast(dst, _P))
        self._check(st, "ntx_hip_gamma_levels_device")

    def unpack_bgra8_gamma_device(self, d_bgra: DeviceBuffer, w: int, h: int, gamma: float,
                                  d_rgba: DeviceBuffer) -> None:
        """unpack_bgra8_device followed by Surface::toLinear(gamma), in one pass."""
        self._need_unpack(d_bgra, w, h, d_rgba, "unpack_bgra8_gamma_device")
        self._check(self.lib.ntx_hip_unpack_bgra8_gamma_device(self.handle, d_bgra.ptr, w, h, gamma, d_rgba.ptr),
                    "ntx_hip_unpack_bgra8_gamma_device")

    def decode_device(self, fmt: int, d_blocks: DeviceBuffer, w: int, h: int, d_rgba: DeviceBuffer,
                      decoder: int = DECODER_D3D10) -> None:
        """Surface::setImage2D for BC1/BC3/BC4/BC5 (ntx_hip_decode_device)."""
        _dims(w, h, "decode_device")
        _need(d_blocks, compressed_size(fmt, w, h), "decode_device blocks")
        _need(d_rgba, _planar(w, h), "decode_device output")
        self._check(self.lib.ntx_hip_decode_device(self.handle, fmt, decoder, d_blocks.ptr, w, h, d_rgba.ptr),
                    "ntx_hip_decode_device")

    def rms_error_device(self, metric: int, d_ref: DeviceBuffer, d_img: DeviceBuffer, w: int, h: int) -> float:
        """nvtt::rmsError / rmsAlphaError (ntx_hip_rms_error_device)."""
        _dims(w, h, "rms_error_device")
        _need(d_ref, _planar(w, h), "rms_error_device reference")
        _need(d_img, _planar(w, h), "rms_error_device image")
        r = _F()
        self._check(self.lib.ntx_hip_rms_error_device(self.handle, metric, d_ref.ptr, d_img.ptr, w, h, ctypes.byref(r)),
                    "ntx_hip_rms_error_device")
        return float(r.value)

    def normalize_normals_device(self, d_rgba: DeviceBuffer, w: int, h: int) -> None:
        """expandNormals + normalizeNormalMap + packNormals in place (ntx_hip_normalize_normals_device)."""
        _dims(w, h, "normalize_normals_device")
        _need(d_rgba, _planar(w, h), "normalize_normals_device")
        self._check(self.lib.ntx_hip_normalize_normals_device(self.handle, d_rgba.ptr, w, h),
                    "ntx_hip_normalize_normals_device")

    def bc1_sse_device(self, d_rgba: DeviceBuffer, w: int, h: int, d_blocks: DeviceBuffer) -> float:
        """Sum over texels of the squared R, G, B differences (8-bit units) of a
        BC1 level against its FP32 source (ntx_hip_bc1_sse_device)."""
        _dims(w, h, "bc1_sse_device")
        _need(d_rgba, _planar(w, h), "bc1_sse_device image")
        _need(d_blocks, compressed_size(FORMAT_BC1, w, h), "bc1_sse_device blocks")
        sse = ctypes.c_double()
        self._check(self.lib.ntx_hip_bc1_sse_device(self.handle, d_rgba.ptr, w, h, d_blocks.ptr, ctypes.byref(sse)),
                    "bc1_sse")
        return float(sse.value)

    def bc1_psnr_device(self, d_rgba: DeviceBuffer, w: int, h: int, d_blocks: DeviceBuffer) -> float:
        """PSNR (dB) of a BC1 level against its FP32 source, computed on the device."""
        _dims(w, h, "bc1_psnr_device")
        _need(d_rgba, _planar(w, h), "bc1_psnr_device image")
        _need(d_blocks, compressed_size(FORMAT_BC1, w, h), "bc1_psnr_device blocks")
        sse = ctypes.c_double()
        self._check(self.lib.ntx_hip_bc1_sse_device(self.handle, d_rgba.ptr, w, h, d_blocks.ptr, ctypes.byref(sse)),
                    "bc1_sse")
        mse = sse.value / (w * h)
        return float("inf") if mse == 0 else float(20.0 * np.log10(255.0 / np.sqrt(mse)))

    @staticmethod
    def _need_unpack(d_bgra: DeviceBuffer, w: int, h: int, d_rgba: DeviceBuffer, what: str) -> None:
        _dims(w, h, what)
        _need(d_bgra, 4 * int(w) * int(h), what + " input")
        _need(d_rgba, _planar(w, h), what + " output")

    def synchronize(self) -> None:
        self._check(self.lib.ntx_hip_synchronize(self.handle), "synchronize")

    def timer_begin(self) -> None:
        self._check(self.lib.ntx_hip_timer_begin(self.handle), "timer_begin")

    def profile(self, enable: bool) -> None:
        self._check(self.lib.ntx_hip_profile_enable(self.handle, int(bool(enable))), "profile_enable")

    def profile_read(self, reset: bool = True) -> dict:
        """{kernel name: (total ms, launches)} since the last reset."""
        ms = np.zeros(len(KERNELS), np.float32)
        n = np.zeros(len(KERNELS), np.int32)
        self._check(self.lib.ntx_hip_profile_read(self.handle, _ptr(ms), _ptr(n), int(bool(reset))), "profile_read")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KERNELS) if n[i]}

    def timer_end(self) -> float:
        ms = _F()
        self._check(self.lib.ntx_hip_timer_end(self.handle, ctypes.byref(ms)), "timer_end")
        return float(ms.value)


class Batch:
    """Many textures, each with a mip chain and a block format, run as one
    HIP graph (ntx_hip_batch_*: mip stages and a compression queue, or one
    persistent launch for both, then the two-pass BC1 encoder).  `textures` is a sequence of dicts
    with keys d_rgba (DeviceBuffer of planar FP32 level 0), width, height,
    format and optionally quality, alpha_mode, color_weights, mip_levels,
    filter, filter_width, filter_params, wrap."""

    def __init__(self, ctx: Context, textures: Sequence[dict]):
        self.ctx = ctx
        self._keep = list(textures)  # the level-0 buffers must outlive the batch
        arr = (TextureDesc * len(textures))()
        for i, t in enumerate(textures):
            d = arr[i]
            d.d_rgba = t["d_rgba"].ptr
            d.width, d.height, d.format = int(t["width"]), int(t["height"]), int(t["format"])
            d.quality = int(t.get("quality", QUALITY_NORMAL))
            d.alpha_mode = int(t.get("alpha_mode", ALPHA_NONE))
            for k, v in enumerate(t.get("color_weights", (1, 1, 1, 1))):
                d.color_weights[k] = float(v)
            d.mip_levels = int(t.get("mip_levels", 0))
            d.filter = int(t.get("filter", FILTER_BOX))
            d.filter_width = float(t.get("filter_width", 0.0))
            for k, v in enumerate(t.get("filter_params", (0.0, 0.0))):
                d.filter_params[k] = float(v)
            d.wrap = int(t.get("wrap", WRAP_MIRROR))
        h = _P()
        ctx._check(ctx.lib.ntx_hip_batch_create(ctx.handle, ctypes.cast(arr, _P), len(textures), ctypes.byref(h)),
                   "ntx_hip_batch_create")
        self.handle = h.value
        self.n = len(textures)

    def close(self) -> None:
        if self.handle:
            self.ctx.lib.ntx_hip_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self) -> None:
        self.ctx._check(self.ctx.lib.ntx_hip_batch_submit(self.handle), "ntx_hip_batch_submit")

    def wait(self) -> None:
        self.ctx._check(self.ctx.lib.ntx_hip_batch_wait(self.handle), "ntx_hip_batch_wait")

    def run(self) -> None:
        self.submit()
        self.wait()

    def levels(self, tex: int) -> int:
        return int(self.ctx.lib.ntx_hip_batch_levels(self.handle, tex))

    def level_info(self, tex: int, level: int):
        """(device image pointer, w, h, device blocks pointer, block bytes)."""
        img, blocks = _P(), _P()
        w, h = _I(), _I()
        nb = _S()
        self.ctx._check(self.ctx.lib.ntx_hip_batch_level(self.handle, tex, level, ctypes.byref(img), ctypes.byref(w),
                                                         ctypes.byref(h), ctypes.byref(blocks), ctypes.byref(nb)),
                        "ntx_hip_batch_level")
        return img.value, w.value, h.value, blocks.value, nb.value

    def blocks(self, tex: int, level: int) -> np.ndarray:
        _, _, _, ptr, nb = self.level_info(tex, level)
        out = np.empty(nb, np.uint8)
        self.ctx._check(self.ctx.lib.ntx_hip_download(self.ctx.handle, _ptr(out), ptr, nb), "download")
        return out

    def image(self, tex: int, level: int) -> np.ndarray:
        ptr, w, h, _, _ = self.level_info(tex, level)
        out = np.empty(4 * w * h, np.float32)
        self.ctx._check(self.ctx.lib.ntx_hip_download(self.ctx.handle, _ptr(out), ptr, out.nbytes), "download")
        return out

    def tuning(self) -> dict:
        t = BatchTuning()
        self.ctx._check(self.ctx.lib.ntx_hip_batch_get_tuning(self.handle, ctypes.byref(t)), "batch_get_tuning")
        return t.as_dict()

    def set_tuning(self, **fields) -> None:
        t = BatchTuning()
        self.ctx._check(self.ctx.lib.ntx_hip_batch_get_tuning(self.handle, ctypes.byref(t)), "batch_get_tuning")
        for k, v in fields.items():
            setattr(t, k, int(v))
        self.ctx._check(self.ctx.lib.ntx_hip_batch_set_tuning(self.handle, ctypes.byref(t)), "batch_set_tuning")

    def autotune(self, reps: int = 2):
        """Returns (chosen tuning dict, ms per run)."""
        t = BatchTuning()
        ms = _F()
        self.ctx._check(self.ctx.lib.ntx_hip_batch_autotune(self.handle, reps, ctypes.byref(t), ctypes.byref(ms)),
                        "ntx_hip_batch_autotune")
        return t.as_dict(), float(ms.value)


class Mlp:
    """The BC1 endpoint-proposal network (ntx_hip_mlp_*): parameters and Adam
    state live on the device; training and inference run the gfx950 kernels."""

    def __init__(self, ctx: Context, params: Optional[np.ndarray] = None, seed: int = 1):
        self.ctx = ctx
        h = _P()
        pp = None if params is None else _f32c(params)
        if pp is not None and pp.size != MLP_PARAMS:
            raise ValueError(f"Mlp: {pp.size} parameters, the network has {MLP_PARAMS}")
        ctx._check(ctx.lib.ntx_hip_mlp_create(ctx.handle, _ptr(pp) if pp is not None else None, seed, ctypes.byref(h)),
                   "ntx_hip_mlp_create")
        self.handle = h.value

    def close(self) -> None:
        if self.handle:
            self.ctx.lib.ntx_hip_mlp_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def params(self) -> np.ndarray:
        out = np.empty(MLP_PARAMS, np.float32)
        self.ctx._check(self.ctx.lib.ntx_hip_mlp_get_params(self.handle, _ptr(out)), "mlp_get_params")
        return out

    def set_params(self, params: np.ndarray) -> None:
        pp = _f32c(params)
        if pp.size != MLP_PARAMS:
            raise ValueError(f"set_params: {pp.size} parameters, the network has {MLP_PARAMS}")
        self.ctx._check(self.ctx.lib.ntx_hip_mlp_set_params(self.handle, _ptr(pp)), "mlp_set_params")

    def forward(self, d_features: int, n: int) -> np.ndarray:
        """Network outputs (n, 32) for n blocks' features at device address d_features."""
        out = self.ctx.alloc(4 * 32 * max(n, 1))
        self.ctx._check(self.ctx.lib.ntx_hip_mlp_forward_device(self.handle, d_features, n, out.ptr), "mlp_forward")
        res = out.download(np.float32, 32 * n).reshape(n, 32)
        out.free()
        return res

    def train_step(self, d_features: int, d_labels: int, n: int, lr: float, want_grad: bool = False):
        """(mean loss, mean gradient or None); lr = 0 computes without updating."""
        loss = _F()
        grad = np.empty(MLP_PARAMS, np.float32) if want_grad else None
        self.ctx._check(self.ctx.lib.ntx_hip_mlp_train_step(self.handle, d_features, d_labels, n, float(lr),
                                                            ctypes.byref(loss), _ptr(grad) if want_grad else None),
                        "mlp_train_step")
        return float(loss.value), grad

    def compress_bc1_device(self, d_rgba: DeviceBuffer, w: int, h: int, d_out: DeviceBuffer, tau_pred: float = 4.0,
                            tau_gap: float = -1.0, refine_iters: int = 2, greedy_rounds: int = 2,
                            box_candidates: bool = False, count_fallback: bool = True) -> Optional[int]:
        _dims(w, h, "compress_bc1_device")
        _need(d_rgba, _planar(w, h), "compress_bc1_device input")
        _need(d_out, compressed_size(FORMAT_BC1, w, h), "compress_bc1_device output")
        fb = _I()
        st = self.ctx.lib.ntx_hip_compress_bc1_mlp_device(self.handle, d_rgba.ptr, w, h, float(tau_pred), float(tau_gap),
                                                          int(refine_iters), int(greedy_rounds),
                                                          MLP_BOX_CANDIDATES if box_candidates else 0,
                                                          d_out.ptr, ctypes.byref(fb) if count_fallback else None)
        self.ctx._check(st, "compress_bc1_mlp")
        return fb.value if count_fallback else None

    def compress_bc1_images_device(self, d_rgba: Sequence[DeviceBuffer], sizes: Sequence[Tuple[int, int]],
                                   d_out: Sequence[DeviceBuffer], settings: Optional[Sequence[Optional[dict]]] = None,
                                   levels: Optional[Sequence[int]] = None) -> None:
        """The learned form over many images (ntx_hip_compress_bc1_mlp_images_device).
        settings[k]: keyword settings of compress_bc1_device for image k, or
        None for the plain cluster fit; by default learned.level_settings of
        its size and its mip level (levels[k], 0 when not given)."""
        from . import learned
        n = len(d_rgba)
        if len(sizes) != n or len(d_out) != n:
            raise ValueError("one size and one output buffer per image")
        if settings is None:
            lv = levels if levels is not None else [0] * n
            if len(lv) != n:
                raise ValueError("one level per image")
            settings = [learned.level_settings(a, b, int(k)) for (a, b), k in zip(sizes, lv)]
        if len(settings) != n:
            raise ValueError("one settings entry per image")
        for k, (a, b) in enumerate(sizes):
            _dims(a, b, "compress_bc1_images_device")
            _need(d_rgba[k], _planar(a, b), "compress_bc1_images_device input")
            _need(d_out[k], compressed_size(FORMAT_BC1, a, b), "compress_bc1_images_device output")
        cs = (MlpSettings * max(n, 1))()
        for k, c in enumerate(settings):
            if c is None:
                cs[k].plain = 1
                continue
            cs[k].tau_pred = float(c.get("tau_pred", 4.0))
            cs[k].tau_gap = float(c.get("tau_gap", -1.0))
            cs[k].refine_iters = int(c.get("refine_iters", 2))
            cs[k].greedy_rounds = int(c.get("greedy_rounds", 2))
            cs[k].flags = MLP_BOX_CANDIDATES if c.get("box_candidates", False) else 0
            if cs[k].refine_iters < 0 or cs[k].greedy_rounds < 0:
                raise ValueError("compress_bc1_images_device: negative refine_iters or greedy_rounds")
        ins = (ctypes.c_void_p * max(n, 1))(*[d.ptr for d in d_rgba])
        outs = (ctypes.c_void_p * max(n, 1))(*[d.ptr for d in d_out])
        ws = (ctypes.c_int * max(n, 1))(*[a for a, _ in sizes])
        hs = (ctypes.c_int * max(n, 1))(*[b for _, b in sizes])
        st = self.ctx.lib.ntx_hip_compress_bc1_mlp_images_device(self.handle, n, ins, ws, hs, cs, outs)
        self.ctx._check(st, "compress_bc1_mlp_images")


def mlp_features_device(ctx: Context, d_rgba: DeviceBuffer, w: int, h: int, d_features: DeviceBuffer,
                        d_aux: DeviceBuffer) -> None:
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    _dims(w, h, "mlp_features_device")
    _need(d_rgba, _planar(w, h), "mlp_features_device input")
    _need(d_features, 2 * 48 * nb, "mlp_features_device features")
    _need(d_aux, 16 * nb, "mlp_features_device aux")
    ctx._check(ctx.lib.ntx_hip_mlp_features_device(ctx.handle, d_rgba.ptr, w, h, d_features.ptr, d_aux.ptr),
               "mlp_features")


def mlp_labels_device(ctx: Context, d_rgba: DeviceBuffer, w: int, h: int, d_bc1: DeviceBuffer, d_aux: DeviceBuffer,
                      d_labels: DeviceBuffer) -> None:
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    _dims(w, h, "mlp_labels_device")
    _need(d_rgba, _planar(w, h), "mlp_labels_device input")
    _need(d_bc1, 8 * nb, "mlp_labels_device blocks")
    _need(d_aux, 16 * nb, "mlp_labels_device aux")
    _need(d_labels, 4 * 8 * nb, "mlp_labels_device labels")
    ctx._check(ctx.lib.ntx_hip_mlp_labels_device(ctx.handle, d_rgba.ptr, w, h, d_bc1.ptr, d_aux.ptr, d_labels.ptr),
               "mlp_labels")


def mip_sizes(w: int, h: int):
    """Extents of the full chain down to 1x1 (Surface.cpp:155-207 countMipmaps)."""
    out = [(w, h)]
    while w != 1 or h != 1:
        w, h = max(1, w // 2), max(1, h // 2)
        out.append((w, h))
    return out
