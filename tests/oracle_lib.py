"""ctypes access to the oracle (oracle/liboracle.so) and to the reference's
own icbc/squish build (oracle/_ref/libntxref.so).  Test infrastructure."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libntxref.so")
_P = ctypes.c_void_p


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built; run `make -C oracle`")
        L = self.lib = ctypes.CDLL(path)
        L.ntxo_bc1_block.restype = ctypes.c_float
        L.ntxo_output_size.restype = ctypes.c_size_t
        L.ntxo_output_size.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ntxo_kernel_window.restype = ctypes.c_int
        L.ntxo_kernel_window.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, _P]
        L.ntxo_polyphase_kernel.argtypes = [ctypes.c_int, ctypes.c_float, _P, ctypes.c_int, ctypes.c_int, _P]
        L.ntxo_downsample.argtypes = [ctypes.c_int, ctypes.c_float, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P]
        L.ntxo_build_next_mipmap.argtypes = L.ntxo_downsample.argtypes
        L.ntxo_srgb_to_linear.argtypes = [ctypes.c_size_t, _P]
        L.ntxo_bc1_block_q.restype = ctypes.c_float
        L.ntxo_compress_image_q.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, _P,
                                            _P, ctypes.c_int]
        L.ntxo_gamma_table.argtypes = [ctypes.c_int, _P]
        L.ntxo_powf_5_11.argtypes = [ctypes.c_float]
        L.ntxo_powf_5_11.restype = ctypes.c_float
        L.ntxo_powf_11_5.argtypes = [ctypes.c_float]
        L.ntxo_powf_11_5.restype = ctypes.c_float
        L.ntxo_to_linear.argtypes = [ctypes.c_size_t, _P, ctypes.c_float]
        L.ntxo_to_gamma.argtypes = [ctypes.c_size_t, _P, ctypes.c_float]
        L.ntxo_normalize_normals.argtypes = [ctypes.c_size_t, _P]
        L.ntxo_normalize_vectors.argtypes = [ctypes.c_size_t, _P]
        L.ntxo_decode_image.argtypes = [ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, _P]
        L.ntxo_rms_error.argtypes = [ctypes.c_int, _P, _P, ctypes.c_size_t]
        L.ntxo_rms_error.restype = ctypes.c_float
        L.ntxo_init()

    # ---- tables
    def cluster_table4(self):
        t = np.zeros(976 * 3, np.uint8); tot = np.zeros(16, np.int32)
        self.lib.ntxo_cluster_table4(_p(t), _p(tot)); return t.reshape(-1, 3), tot

    def cluster_table3(self):
        t = np.zeros(160 * 2, np.uint8); tot = np.zeros(16, np.int32)
        self.lib.ntxo_cluster_table3(_p(t), _p(tot)); return t.reshape(-1, 2), tot

    def match_tables(self):
        a = np.zeros(512, np.uint8); b = np.zeros(512, np.uint8)
        self.lib.ntxo_icbc_match_tables(_p(a), _p(b)); return a.reshape(256, 2), b.reshape(256, 2)

    def omatch_tables(self):
        a = np.zeros(512, np.uint8); b = np.zeros(512, np.uint8)
        self.lib.ntxo_omatch_tables(_p(a), _p(b)); return a.reshape(256, 2), b.reshape(256, 2)

    def midpoints(self):
        a = np.zeros(32, np.float32); b = np.zeros(64, np.float32)
        self.lib.ntxo_midpoints(_p(a), _p(b)); return a, b

    # ---- blocks
    def bc1_block(self, colors, weights, cw=(1, 1, 1)):
        c = f32(colors).reshape(16, 4); w = f32(weights).reshape(16); cwv = f32(cw)
        out = np.zeros(8, np.uint8); n = ctypes.c_int(); m = ctypes.c_int()
        e = self.lib.ntxo_bc1_block(_p(c), _p(w), _p(cwv), _p(out), ctypes.byref(n), ctypes.byref(m))
        return out, float(e), n.value, m.value

    def bc1_block_q(self, colors, weights, quality, cw=(1, 1, 1)):
        c = f32(colors).reshape(16, 4); w = f32(weights).reshape(16); cwv = f32(cw)
        out = np.zeros(8, np.uint8)
        e = self.lib.ntxo_bc1_block_q(_p(c), _p(w), _p(cwv), int(quality), _p(out))
        return out, float(e)

    def compress_q(self, fmt, quality, planar, w, h, alpha_mode=0, cw=(1, 1, 1, 1), threads=8):
        src = f32(planar).reshape(-1); cwv = f32(cw)
        out = np.zeros(self.lib.ntxo_output_size(int(fmt), int(w), int(h)), np.uint8)
        self.lib.ntxo_compress_image_q(int(fmt), int(quality), int(alpha_mode), _p(cwv), int(w), int(h), _p(src), _p(out),
                                       int(threads))
        return out

    def bc1_decode(self, blk):
        blk = np.ascontiguousarray(blk, np.uint8); out = np.zeros(64, np.uint8)
        self.lib.ntxo_bc1_decode(_p(blk), _p(out)); return out.reshape(16, 4)

    def omatch_alpha_tables(self):
        a = np.zeros(512, np.uint8); b = np.zeros(512, np.uint8)
        self.lib.ntxo_omatch_alpha_tables(_p(a), _p(b)); return a.reshape(256, 2), b.reshape(256, 2)

    def bc1a_block(self, bgra, weight_by_alpha=0, cw=(1, 1, 1)):
        b = np.ascontiguousarray(bgra, np.uint8).reshape(64); cwv = f32(cw); out = np.zeros(8, np.uint8)
        self.lib.ntxo_bc1a_block(_p(b), int(weight_by_alpha), _p(cwv), _p(out)); return out

    def bc1a_squish(self, bgra, weight_by_alpha=0, cw=(1, 1, 1)):
        b = np.ascontiguousarray(bgra, np.uint8).reshape(64); cwv = f32(cw); out = np.zeros(8, np.uint8)
        self.lib.ntxo_bc1a_squish(_p(b), int(weight_by_alpha), _p(cwv), _p(out)); return out

    def bc3n_block(self, bgra, weight_by_alpha=0):
        b = np.ascontiguousarray(bgra, np.uint8).reshape(64); out = np.zeros(16, np.uint8)
        self.lib.ntxo_bc3n_block(_p(b), int(weight_by_alpha), _p(out)); return out

    def bc3_color_block(self, bgra, weight_by_alpha=0, cw=(1, 1, 1)):
        b = np.ascontiguousarray(bgra, np.uint8).reshape(64); cwv = f32(cw); out = np.zeros(8, np.uint8)
        self.lib.ntxo_bc3_color_block(_p(b), int(weight_by_alpha), _p(cwv), _p(out)); return out

    def bc4_channel(self, v):
        v = np.ascontiguousarray(v, np.uint8).reshape(16); out = np.zeros(8, np.uint8)
        self.lib.ntxo_bc4_channel(_p(v), _p(out)); return out

    def bc4_decode(self, blk):
        blk = np.ascontiguousarray(blk, np.uint8); out = np.zeros(16, np.uint8)
        self.lib.ntxo_bc4_decode(_p(blk), _p(out)); return out

    def bc4_channel_optimal(self, v):
        v = np.ascontiguousarray(v, np.uint8).reshape(16); out = np.zeros(8, np.uint8)
        self.lib.ntxo_bc4_channel_optimal(_p(v), _p(out)); return out

    def bc4_channels_optimal(self, v):
        v = np.ascontiguousarray(v, np.uint8); n = v.shape[0]; out = np.zeros((n, 8), np.uint8)
        self.lib.ntxo_bc4_channels_optimal(_p(v), ctypes.c_size_t(n), _p(out)); return out

    def bc1_green_optimal(self, bgra):
        b = np.ascontiguousarray(bgra, np.uint8).reshape(64); out = np.zeros(8, np.uint8)
        self.lib.ntxo_bc1_green_optimal(_p(b), _p(out)); return out

    def bc3_block_q(self, bgra, quality, weight_by_alpha=0, cw=(1, 1, 1)):
        b = np.ascontiguousarray(bgra, np.uint8).reshape(64); cwv = f32(cw); out = np.zeros(16, np.uint8)
        self.lib.ntxo_bc3_block_q(_p(b), int(quality), int(weight_by_alpha), _p(cwv), _p(out)); return out

    def bc3n_block_q(self, bgra, quality, weight_by_alpha=0):
        b = np.ascontiguousarray(bgra, np.uint8).reshape(64); out = np.zeros(16, np.uint8)
        self.lib.ntxo_bc3n_block_q(_p(b), int(quality), int(weight_by_alpha), _p(out)); return out

    # ---- images
    def compress(self, fmt, planar, w, h, alpha_mode=0, cw=(1, 1, 1, 1), threads=8):
        src = f32(planar).reshape(-1); assert src.size == 4 * w * h
        out = np.zeros(self.lib.ntxo_output_size(fmt, w, h), np.uint8); cwv = f32(cw)
        self.lib.ntxo_compress_image(int(fmt), int(alpha_mode), _p(cwv), int(w), int(h), _p(src), _p(out), int(threads))
        return out

    def compress_bc1_counts(self, planar, w, h, alpha_mode=0, cw=(1, 1, 1, 1), threads=8):
        src = f32(planar).reshape(-1)
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        out = np.zeros(nb * 8, np.uint8); nm = np.zeros(2 * nb, np.int16); cwv = f32(cw)
        self.lib.ntxo_compress_bc1_counts(int(alpha_mode), _p(cwv), int(w), int(h), _p(src), _p(out), _p(nm), int(threads))
        return out, nm.reshape(-1, 2)

    def fast_downsample(self, planar, W, H):
        src = f32(planar).reshape(-1); w, h = max(1, W // 2), max(1, H // 2)
        out = np.zeros(4 * w * h, np.float32)
        self.lib.ntxo_fast_downsample(int(W), int(H), _p(src), _p(out)); return out

    def polyphase(self, filt, width, params, src_len, dst_len):
        ws = self.lib.ntxo_kernel_window(int(filt), float(width), int(src_len), int(dst_len), None)
        out = np.zeros(ws * dst_len, np.float32)
        pp = f32(params) if params is not None else None
        self.lib.ntxo_polyphase_kernel(int(filt), float(width), _p(pp) if pp is not None else None, int(src_len), int(dst_len), _p(out))
        return out.reshape(dst_len, ws)

    def downsample(self, filt, width, params, wrap, planar, W, H):
        src = f32(planar).reshape(-1); w, h = max(1, W // 2), max(1, H // 2)
        out = np.zeros(4 * w * h, np.float32)
        pp = f32(params) if params is not None else None
        self.lib.ntxo_downsample(int(filt), float(width), _p(pp) if pp is not None else None, int(wrap), int(W), int(H), _p(src), _p(out))
        return out

    def build_next_mipmap(self, filt, width, params, wrap, planar, W, H):
        src = f32(planar).reshape(-1); w, h = max(1, W // 2), max(1, H // 2)
        out = np.zeros(4 * w * h, np.float32)
        pp = f32(params) if params is not None else None
        self.lib.ntxo_build_next_mipmap(int(filt), float(width), _p(pp) if pp is not None else None, int(wrap), int(W), int(H), _p(src), _p(out))
        return out

    def srgb_to_linear(self, x):
        a = f32(x).copy(); self.lib.ntxo_srgb_to_linear(a.size, _p(a)); return a

    def gamma_table(self, which):
        """Exponent table of powf(x, 5/11) (which=0) or powf(x, 11/5) (which=1)."""
        t = np.zeros(512, np.float32); self.lib.ntxo_gamma_table(int(which), _p(t)); return t

    def to_linear(self, x, gamma=2.2):
        a = f32(x).copy(); self.lib.ntxo_to_linear(a.size, _p(a), float(gamma)); return a

    def to_gamma(self, x, gamma=2.2):
        a = f32(x).copy(); self.lib.ntxo_to_gamma(a.size, _p(a), float(gamma)); return a

    def planar_to_linear(self, planar, gamma=2.2):
        """FloatImage::toLinear(0, 3, gamma) of planar RGBA: R, G, B transformed, A kept."""
        a = f32(planar).reshape(-1).copy(); n = a.size // 4
        a[:3 * n] = self.to_linear(a[:3 * n], gamma); return a

    def decode_image(self, fmt, blocks, w, h, d3d9=False):
        b = np.ascontiguousarray(blocks, np.uint8); out = np.zeros(4 * w * h, np.float32)
        self.lib.ntxo_decode_image(int(fmt), int(bool(d3d9)), _p(b), int(w), int(h), _p(out)); return out

    def rms_error(self, metric, ref, img):
        r = f32(ref).reshape(-1); i = f32(img).reshape(-1)
        return float(self.lib.ntxo_rms_error(int(metric), _p(r), _p(i), r.size // 4))

    def normalize_vectors(self, planar):
        a = f32(planar).reshape(-1).copy(); self.lib.ntxo_normalize_vectors(a.size // 4, _p(a)); return a

    def normalize_normals(self, planar):
        a = f32(planar).reshape(-1).copy(); self.lib.ntxo_normalize_normals(a.size // 4, _p(a)); return a

    def planar_to_gamma(self, planar, gamma=2.2):
        a = f32(planar).reshape(-1).copy(); n = a.size // 4
        a[:3 * n] = self.to_gamma(a[:3 * n], gamma); return a


class Reference:
    """The reference's own icbc.h (ICBC_SIMD=2) and nvsquish (scalar)."""

    @staticmethod
    def load(path=REF_SO):
        if not os.path.exists(path):
            return None
        return Reference(path)

    def __init__(self, path):
        L = self.lib = ctypes.CDLL(path)
        L.ntxref_bc1_block.restype = ctypes.c_float
        L.ntxref_bc1_block_q.restype = ctypes.c_float
        L.ntxref_init()

    def bc1_block_q(self, colors, weights, quality, cw=(1, 1, 1)):
        c = f32(colors).reshape(16, 4); w = f32(weights).reshape(16); cwv = f32(cw)
        out = np.zeros(8, np.uint8)
        e = self.lib.ntxref_bc1_block_q(_p(c), _p(w), _p(cwv), int(quality), _p(out))
        return out, float(e)

    def bc1_block(self, colors, weights, cw=(1, 1, 1)):
        c = f32(colors).reshape(16, 4); w = f32(weights).reshape(16); cwv = f32(cw)
        out = np.zeros(8, np.uint8)
        e = self.lib.ntxref_bc1_block(_p(c), _p(w), _p(cwv), _p(out))
        return out, float(e)

    def bc1_decode(self, blk):
        blk = np.ascontiguousarray(blk, np.uint8); out = np.zeros(64, np.uint8)
        self.lib.ntxref_bc1_decode(_p(blk), _p(out)); return out.reshape(16, 4)

    def bc1a_squish(self, bgra, weight_by_alpha=0, cw=(1, 1, 1)):
        b = np.ascontiguousarray(bgra, np.uint8).reshape(64); cwv = f32(cw); out = np.zeros(8, np.uint8)
        self.lib.ntxref_bc1a_squish(_p(b), int(weight_by_alpha), _p(cwv), _p(out)); return out

    def bc3_color_squish(self, bgra, weight_by_alpha=0, cw=(1, 1, 1)):
        b = np.ascontiguousarray(bgra, np.uint8).reshape(64); cwv = f32(cw); out = np.zeros(8, np.uint8)
        self.lib.ntxref_bc3_color_squish(_p(b), int(weight_by_alpha), _p(cwv), _p(out)); return out

    # ---- the reference's nvimage / nvtt CPU code (oracle/ref_harness_nv.cpp)
    KIND = {1: 0, 2: 1, 4: 2, 5: 3, 6: 4, 7: 5}  # nvtt::Format -> harness kind (BC1, BC1a, BC3, BC3n, BC4, BC5)

    def colorblock_init(self, planar, w, h, x, y):
        """ColorBlock::init(w, h, float*, x, y): 16 BGRA texels."""
        src = f32(planar).reshape(-1); out = np.zeros(64, np.uint8)
        self.lib.ntxref_colorblock_init(_p(src), int(w), int(h), int(x), int(y), _p(out)); return out

    def bc4_channel(self, v):
        """QuickCompress::compressDXT5A on one channel block."""
        v = np.ascontiguousarray(v, np.uint8).reshape(16); out = np.zeros(8, np.uint8)
        self.lib.ntxref_bc4_channel(_p(v), _p(out)); return out

    def bc4_channel_optimal(self, v):
        """OptimalCompress::compressDXT5A with its first error seed taken from (max, min)."""
        v = np.ascontiguousarray(v, np.uint8).reshape(16); out = np.zeros(8, np.uint8)
        self.lib.ntxref_bc4_channel_optimal(_p(v), _p(out)); return out

    def block(self, fmt, quality, bgra, alpha_mode=0, cw=(1, 1, 1, 1)):
        """One 8-bit block through the reference compressor Context.cpp picks for (fmt, quality)."""
        b = np.ascontiguousarray(bgra, np.uint8).reshape(64); cwv = f32(cw)
        out = np.zeros(8 if fmt in (1, 2, 6) else 16, np.uint8)
        self.lib.ntxref_block(self.KIND[fmt], int(quality), int(alpha_mode), _p(cwv), _p(b), _p(out)); return out

    def compress_image(self, fmt, quality, planar, w, h, alpha_mode=0, cw=(1, 1, 1, 1), threads=None):
        """A whole image through the reference's own block loop: its
        ParallelTaskDispatcher (nvthread's processor count), or, with
        `threads`, the same task loop on that many host threads."""
        src = f32(planar).reshape(-1); cwv = f32(cw)
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        out = np.zeros(nb * (8 if fmt in (1, 2, 6) else 16), np.uint8)
        if threads is None:
            fn = self.lib.ntxref_compress_image
            fn.restype = ctypes.c_size_t
            n = fn(self.KIND[fmt], int(quality), int(alpha_mode), _p(cwv), _p(src), int(w), int(h), _p(out),
                   ctypes.c_size_t(out.size))
        else:
            fn = self.lib.ntxref_compress_image_threads
            fn.restype = ctypes.c_size_t
            n = fn(self.KIND[fmt], int(quality), int(alpha_mode), _p(cwv), _p(src), int(w), int(h), _p(out),
                   ctypes.c_size_t(out.size), int(threads))
        assert n == out.size, (n, out.size)
        return out

    def thread_count(self):
        return int(self.lib.ntxref_thread_count())

    def fast_downsample(self, planar, W, H):
        src = f32(planar).reshape(-1); out = np.zeros(4 * max(1, W // 2) * max(1, H // 2), np.float32)
        self.lib.ntxref_fast_downsample(int(W), int(H), _p(src), _p(out)); return out

    def build_next_mipmap(self, filt, width, params, wrap, planar, W, H, alpha_mode=0):
        src = f32(planar).reshape(-1); out = np.zeros(4 * max(1, W // 2) * max(1, H // 2), np.float32)
        pp = f32(params) if params is not None else None
        self.lib.ntxref_build_next_mipmap(int(filt), ctypes.c_float(width), _p(pp) if pp is not None else None,
                                          int(wrap), int(alpha_mode), int(W), int(H), _p(src), _p(out))
        return out

    def polyphase(self, filt, width, params, src_len, dst_len):
        pp = f32(params) if params is not None else None
        fn = self.lib.ntxref_polyphase_kernel
        kw = ctypes.c_float()
        ws = fn(int(filt), ctypes.c_float(width), _p(pp) if pp is not None else None, int(src_len), int(dst_len), None,
                ctypes.byref(kw))
        out = np.zeros(ws * dst_len, np.float32)
        fn(int(filt), ctypes.c_float(width), _p(pp) if pp is not None else None, int(src_len), int(dst_len), _p(out),
           None)
        return out.reshape(dst_len, ws), kw.value

    def powf(self, which, x):
        """nvmath powf_5_11 (which=0) / powf_11_5 (which=1) over an array."""
        a = f32(x).copy(); self.lib.ntxref_powf_array(int(which), _p(a), int(a.size)); return a

    def planar_gamma(self, planar, gamma, to_gamma):
        """FloatImage::toLinear / toGamma(0, 3, gamma) of planar RGBA."""
        a = f32(planar).reshape(-1).copy()
        self.lib.ntxref_gamma(int(bool(to_gamma)), ctypes.c_size_t(a.size // 4), _p(a), ctypes.c_float(gamma))
        return a

    def normalize(self, planar):
        """FloatImage::normalize(0) of planar RGBA."""
        a = f32(planar).reshape(-1).copy(); self.lib.ntxref_normalize(ctypes.c_size_t(a.size // 4), _p(a)); return a

    def decode_block(self, fmt, blk, d3d9=False):
        """BlockDXT1/DXT5/ATI1/ATI2::decodeBlock: 16 BGRA texels."""
        b = np.ascontiguousarray(blk, np.uint8); out = np.zeros(64, np.uint8)
        self.lib.ntxref_decode_block(int(fmt), int(bool(d3d9)), _p(b), _p(out)); return out.reshape(16, 4)
