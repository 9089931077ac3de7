"""GPU parity: the gfx950 kernels, called through the C ABI, against the
golden vectors written by the reference's own encoders and against the
oracle on the same seeded inputs.  The bar is bit-exact output."""
import os

import numpy as np
import pytest

import blocks
import ntx_amd as N
from ntx_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")

ODD_SIZES = [(1, 1), (2, 3), (3, 5), (5, 4), (7, 1), (1, 6), (37, 23), (130, 66)]


def _blocks_as_image(cols, wts):
    """Lay blocks side by side in a 4 x 4n image; weights go to alpha so that
    AlphaMode_Transparency reproduces them (weight = saturate(a))."""
    n = len(cols)
    img = np.zeros((4, 4, 4 * n), np.float32)
    for i in range(n):
        for t in range(16):
            y, x = divmod(t, 4)
            img[:3, y, 4 * i + x] = cols[i, t, :3]
            img[3, y, 4 * i + x] = wts[i, t]
    return img, 4 * n, 4


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_bc1_golden_blocks(gpu):
    g = np.load(os.path.join(GOLD, "bc1_blocks.npz"))
    cws = g["color_weights"]
    keys = {tuple(c) for c in cws}
    bad = 0
    for key in keys:
        sel = np.nonzero(np.all(cws == np.array(key, np.float32), axis=1))[0]
        img, w, h = _blocks_as_image(g["colors"][sel], g["weights"][sel])
        out = gpu.compress(N.FORMAT_BC1, img, w, h, alpha_mode=N.ALPHA_TRANSPARENCY,
                           color_weights=(key[0], key[1], key[2], 1.0)).reshape(-1, 8)
        bad += int(np.sum(np.any(out != g["expected"][sel], axis=1)))
    assert bad == 0, f"{bad} of {len(cws)} golden BC1 blocks differ"


def test_bc1_box_chain_golden(gpu):
    g = np.load(os.path.join(GOLD, "chain_c1_64.npz"))
    planar = synth.to_planar(g["image"]).reshape(-1)
    w = h = 64
    while True:
        out = gpu.compress(N.FORMAT_BC1, planar, w, h).reshape(-1, 8)
        assert np.array_equal(out, g[f"bc1_{w}x{h}"]), (w, h)
        if w == 1 and h == 1:
            break
        planar = gpu.downsample(N.FILTER_BOX, planar, w, h)
        w, h = max(1, w // 2), max(1, h // 2)


@pytest.mark.parametrize("seed,size", [(1, (256, 256)), (2, (128, 64))] + [(3, s) for s in ODD_SIZES])
def test_bc1_chain_matches_oracle(gpu, oracle, seed, size):
    w, h = size
    planar = synth.to_planar(synth.noise_rect_rgba(w, h, seed=seed)).reshape(-1)
    while True:
        a = gpu.compress(N.FORMAT_BC1, planar, w, h)
        b = oracle.compress(N.FORMAT_BC1, planar, w, h)
        assert np.array_equal(a, b), (w, h, int(np.sum(np.any(a.reshape(-1, 8) != b.reshape(-1, 8), axis=1))))
        if w == 1 and h == 1:
            break
        nxt = gpu.downsample(N.FILTER_BOX, planar, w, h)
        assert np.array_equal(_bits(nxt), _bits(oracle.build_next_mipmap(N.FILTER_BOX, 0.5, None, N.WRAP_MIRROR, planar, w, h)))
        planar = nxt
        w, h = max(1, w // 2), max(1, h // 2)


@pytest.mark.parametrize("kind", range(len(blocks.KINDS)))
def test_bc1_adversarial_blocks(gpu, oracle, kind):
    rng = np.random.default_rng(100 + kind)
    cols = np.zeros((512, 16, 4), np.float32)
    wts = np.zeros((512, 16), np.float32)
    for i in range(512):
        cols[i], wts[i] = blocks.bc1_block(rng, blocks.KINDS[kind])
    img, w, h = _blocks_as_image(cols, wts)
    for cw in ((1, 1, 1, 1), (0.5, 1.0, 0.25, 1.0)):
        a = gpu.compress(N.FORMAT_BC1, img, w, h, alpha_mode=N.ALPHA_TRANSPARENCY, color_weights=cw)
        b = oracle.compress(N.FORMAT_BC1, img, w, h, alpha_mode=N.ALPHA_TRANSPARENCY, cw=cw)
        assert np.array_equal(a, b), (blocks.KINDS[kind], cw, int(np.sum(np.any(a.reshape(-1, 8) != b.reshape(-1, 8), axis=1))))


def test_bc3_color_golden_blocks(gpu):
    g = np.load(os.path.join(GOLD, "bc3_color_blocks.npz"))
    px, wba, exp = g["bgra"], g["weight_by_alpha"], g["expected"]
    for mode in (0, 1):
        sel = np.nonzero(wba == mode)[0]
        n = len(sel)
        img = np.zeros((4, 4, 4 * n), np.float32)
        for j, i in enumerate(sel):
            for t in range(16):
                y, x = divmod(t, 4)
                b, gg, r, a = px[i, t].astype(np.float32)
                # (v + 0.5) / 255 truncates back to v in ColorBlock::init
                img[:, y, 4 * j + x] = (np.array([r, gg, b, a], np.float32) + np.float32(0.5)) / np.float32(255)
        out = gpu.compress(N.FORMAT_BC3, img, 4 * n, 4,
                           alpha_mode=N.ALPHA_TRANSPARENCY if mode else N.ALPHA_NONE).reshape(-1, 16)
        assert np.array_equal(out[:, 8:], exp[sel]), int(np.sum(np.any(out[:, 8:] != exp[sel], axis=1)))


@pytest.mark.parametrize("fmt", [N.FORMAT_BC3, N.FORMAT_BC4, N.FORMAT_BC5])
@pytest.mark.parametrize("size", [(256, 256), (64, 128)] + ODD_SIZES)
def test_bc345_images_match_oracle(gpu, oracle, fmt, size):
    w, h = size
    if fmt == N.FORMAT_BC3:
        img = synth.noise_rect_rgba(w, h, seed=11, transparent=True)
    elif fmt == N.FORMAT_BC5:
        img = synth.normal_map_rgba(w, h, seed=12)
    else:
        img = synth.single_channel_rgba(w, h, seed=13)
    planar = synth.to_planar(img).reshape(-1)
    for alpha_mode in ((N.ALPHA_NONE, N.ALPHA_TRANSPARENCY) if fmt == N.FORMAT_BC3 else (N.ALPHA_NONE,)):
        a = gpu.compress(fmt, planar, w, h, alpha_mode=alpha_mode)
        b = oracle.compress(fmt, planar, w, h, alpha_mode=alpha_mode)
        assert np.array_equal(a, b), (fmt, size, alpha_mode)


def test_bc345_random_float_blocks(gpu, oracle):
    rng = np.random.default_rng(21)
    w, h = 64, 48
    planar = rng.random(4 * w * h, dtype=np.float32) * np.float32(1.2) - np.float32(0.1)
    for fmt in (N.FORMAT_BC3, N.FORMAT_BC4, N.FORMAT_BC5):
        assert np.array_equal(gpu.compress(fmt, planar, w, h), oracle.compress(fmt, planar, w, h)), fmt


@pytest.mark.parametrize("filt", [N.FILTER_BOX, N.FILTER_TRIANGLE, N.FILTER_KAISER])
@pytest.mark.parametrize("wrap", [N.WRAP_CLAMP, N.WRAP_REPEAT, N.WRAP_MIRROR])
@pytest.mark.parametrize("size", [(64, 64), (37, 23), (1, 7), (6, 1), (2, 2), (5, 4), (130, 66), (3, 3)])
def test_mip_filters_match_oracle(gpu, oracle, filt, wrap, size):
    W, H = size
    rng = np.random.default_rng(W * 131 + H)
    planar = rng.random(4 * W * H, dtype=np.float32)
    width, params = N.DEFAULT_FILTER[filt]
    a = gpu.downsample(filt, planar, W, H, wrap=wrap)
    b = oracle.build_next_mipmap(filt, width, params, wrap, planar, W, H)
    assert np.array_equal(_bits(a), _bits(b))
    if filt == N.FILTER_BOX:
        # the polyphase Box (width != 0.5 path is fastDownSample-free) too
        a2 = gpu.downsample(filt, planar, W, H, width=1.0, wrap=wrap)
        b2 = oracle.build_next_mipmap(filt, 1.0, None, wrap, planar, W, H)
        assert np.array_equal(_bits(a2), _bits(b2))


def test_kaiser_custom_parameters(gpu, oracle):
    W, H = 96, 40
    planar = np.random.default_rng(5).random(4 * W * H, dtype=np.float32)
    for width, params in ((2.0, (3.0, 1.0)), (3.0, (4.0, 0.75)), (4.0, (6.0, 1.0))):
        a = gpu.downsample(N.FILTER_KAISER, planar, W, H, width=width, params=params)
        b = oracle.build_next_mipmap(N.FILTER_KAISER, width, params, N.WRAP_MIRROR, planar, W, H)
        assert np.array_equal(_bits(a), _bits(b)), (width, params)


def test_srgb_to_linear(gpu, oracle):
    w, h = 256, 3
    v = np.concatenate([np.arange(256, dtype=np.float32) / np.float32(255)] * 12)
    d = gpu.alloc(v.nbytes).upload(v)
    gpu.srgb_to_linear_device(d, w, h)
    got = d.download(np.float32)
    exp = oracle.srgb_to_linear(v[: 3 * w * h])
    assert np.array_equal(_bits(got[: 3 * w * h]), _bits(exp))
    assert np.array_equal(got[3 * w * h:], v[3 * w * h:])  # alpha plane untouched
    # arbitrary floats: device powf, within 2 ulp of the host result
    r = np.random.default_rng(3).random(4 * 4096, dtype=np.float32) * np.float32(1.2) - np.float32(0.1)
    d2 = gpu.alloc(r.nbytes).upload(r)
    gpu.srgb_to_linear_device(d2, 4096, 1)
    got = d2.download(np.float32)[: 3 * 4096]
    exp = oracle.srgb_to_linear(r[: 3 * 4096])
    ulps = np.abs(_bits(got).astype(np.int64) - _bits(exp).astype(np.int64))
    assert ulps.max() <= 2
    d.free(); d2.free()


def test_unpack_bgra8(gpu):
    img = synth.noise_rect_rgba(40, 24, seed=9)
    bgra = synth.to_bgra(img)
    d_in = gpu.alloc(bgra.nbytes).upload(bgra)
    d_out = gpu.alloc(4 * 40 * 24 * 4)
    gpu.unpack_bgra8_device(d_in, 40, 24, d_out)
    got = d_out.download(np.float32)
    assert np.array_equal(got, synth.to_planar(img).reshape(-1))
    d_in.free(); d_out.free()


@pytest.mark.parametrize("w,h,offset", [(40, 24, 0), (37, 21, 0), (40, 24, 4), (37, 21, 4), (1, 1, 0), (3, 1, 12)])
def test_front_end_sizes_and_alignment(gpu, oracle, w, h, offset):
    """Unpack and sRGB conversion on texel counts that are not a multiple of
    four and on buffers that start off a 16-byte boundary (the kernels take
    four values per lane only when both allow it)."""
    img = synth.noise_rect_rgba(w, h, seed=w * h)
    bgra = synth.to_bgra(img)
    n = w * h
    d_in = gpu.alloc(bgra.nbytes + 16)
    d_out = gpu.alloc(16 * n + 16)
    src = np.zeros(bgra.nbytes + 16, np.uint8)
    src[offset:offset + bgra.nbytes] = bgra.reshape(-1)
    d_in.upload(src)
    gpu._check(gpu.lib.ntx_hip_unpack_bgra8_device(gpu.handle, d_in.ptr + offset, w, h, d_out.ptr + offset), "unpack")
    raw = d_out.download(np.uint8)[offset:offset + 16 * n].copy().view(np.float32)
    planar = synth.to_planar(img).reshape(-1)
    assert np.array_equal(raw, planar)
    gpu._check(gpu.lib.ntx_hip_srgb_to_linear_device(gpu.handle, d_out.ptr + offset, w, h), "srgb")
    raw = d_out.download(np.uint8)[offset:offset + 16 * n].copy().view(np.float32)
    assert np.array_equal(_bits(raw[: 3 * n]), _bits(oracle.srgb_to_linear(planar[: 3 * n])))
    assert np.array_equal(raw[3 * n:], planar[3 * n:])
    # both steps in one pass: the same bits
    d_out.upload(np.zeros(16 * n + 16, np.uint8))
    gpu._check(gpu.lib.ntx_hip_unpack_bgra8_srgb_device(gpu.handle, d_in.ptr + offset, w, h, d_out.ptr + offset),
               "unpack_srgb")
    fused = d_out.download(np.uint8)[offset:offset + 16 * n].copy().view(np.float32)
    assert np.array_equal(_bits(fused), _bits(raw))
    d_in.free(); d_out.free()


def test_unpack_srgb_every_byte_value(gpu, oracle):
    """All 256 byte values in every channel: the one-pass conversion equals
    setImage followed by toLinearFromSrgb."""
    w, h = 256, 4
    bgra = np.zeros((h, w, 4), np.uint8)
    v = np.arange(256, dtype=np.uint8)
    for c in range(4):
        bgra[c, :, :] = np.roll(v, 37 * c)[:, None]
    d_in = gpu.alloc(bgra.nbytes).upload(bgra)
    d_a, d_b = gpu.alloc(16 * w * h), gpu.alloc(16 * w * h)
    gpu.unpack_bgra8_device(d_in, w, h, d_a)
    gpu.srgb_to_linear_device(d_a, w, h)
    gpu.unpack_bgra8_srgb_device(d_in, w, h, d_b)
    a, b = d_a.download(np.float32), d_b.download(np.float32)
    assert np.array_equal(_bits(a), _bits(b))
    n = w * h
    planar = (bgra.reshape(n, 4)[:, [2, 1, 0, 3]].T.astype(np.float32) / np.float32(255)).reshape(-1)
    assert np.array_equal(_bits(b[: 3 * n]), _bits(oracle.srgb_to_linear(planar[: 3 * n])))
    d_in.free(); d_a.free(); d_b.free()


def test_unsupported_and_invalid_calls_report_status(gpu):
    planar = np.zeros(4 * 16, np.float32)
    with pytest.raises(N.NtxError) as e:
        gpu.compress(N.FORMAT_BC4, planar, 4, 4, quality=4)  # no such quality
    assert e.value.status == 2
    with pytest.raises(N.NtxError) as e:
        gpu.compress(3, planar, 4, 4)  # Format_DXT3: not on this path
    assert e.value.status == 2
    with pytest.raises(N.NtxError) as e:
        gpu.downsample(N.FILTER_BOX, planar[:4], 1, 1)
    assert e.value.status == 1


@pytest.mark.parametrize("size", [(257, 129), (1000, 333), (4096, 64), (64, 4096), (21, 1023), (2048, 2048), (1030, 516)])
@pytest.mark.parametrize("filt", [N.FILTER_TRIANGLE, N.FILTER_KAISER])
def test_polyphase_large_and_odd(gpu, oracle, size, filt):
    W, H = size
    planar = np.random.default_rng(W + 7 * H).random(4 * W * H, dtype=np.float32)
    width, params = N.DEFAULT_FILTER[filt]
    for wrap in (N.WRAP_MIRROR, N.WRAP_REPEAT):
        a = gpu.downsample(filt, planar, W, H, wrap=wrap)
        b = oracle.downsample(filt, width, params, wrap, planar, W, H)
        assert np.array_equal(_bits(a), _bits(b)), (size, wrap)


@pytest.mark.parametrize("filt,width,params,alpha_mode", [
    (N.FILTER_BOX, 0.5, None, N.ALPHA_TRANSPARENCY), (N.FILTER_BOX, 1.0, None, N.ALPHA_NONE),
    (N.FILTER_TRIANGLE, 1.0, None, N.ALPHA_NONE), (N.FILTER_TRIANGLE, 2.0, None, N.ALPHA_NONE),
    (N.FILTER_KAISER, 2.0, (4.0, 1.0), N.ALPHA_NONE), (N.FILTER_KAISER, 3.0, (4.0, 1.0), N.ALPHA_NONE),
    (N.FILTER_KAISER, 1.5, (4.0, 1.0), N.ALPHA_NONE), (N.FILTER_KAISER, 2.5, (4.0, 1.0), N.ALPHA_NONE),
    (N.FILTER_KAISER, 3.5, (4.0, 1.0), N.ALPHA_NONE), (N.FILTER_KAISER, 3.75, (4.0, 1.0), N.ALPHA_NONE)])
@pytest.mark.parametrize("size", [(2048, 2048), (2570, 1640), (1024, 4096), (131072, 2), (2, 131072), (256, 130),
                                  (2, 258), (514, 2), (1000, 500)])
def test_halving_windows_match_oracle(gpu, oracle, filt, width, params, alpha_mode, size):
    """Exact halving (even extents) with windows of 3, 5, 7, 9, 11, 13 and 15
    taps, the same taps for every output, in every wrap mode: on sources of
    2048^2 and up the sliding-window kernel (partial column tiles and a
    partial last strip at 2570 x 1640, rows and columns of two texels whose
    taps all wrap), on the smaller ones the tiled kernel (with the taps
    passed by value for windows of 5, 9 and 13; partial tiles at 1000 x 500)
    and for the even 16-tap window the tiled kernel with tables."""
    W, H = size
    planar = np.random.default_rng(W * 3 + H).random(4 * W * H, dtype=np.float32)
    for wrap in (N.WRAP_CLAMP, N.WRAP_REPEAT, N.WRAP_MIRROR):
        a = gpu.downsample(filt, planar, W, H, width=width, params=params, wrap=wrap, alpha_mode=alpha_mode)
        b = oracle.downsample(filt, width, params, wrap, planar, W, H)
        assert np.array_equal(_bits(a), _bits(b)), (size, wrap)


@pytest.mark.parametrize("size", [(4096, 2048), (4096, 4096)])
def test_halving_strip_heights_match_oracle(gpu, oracle, size):
    """The sliding-window kernel at the strip heights the larger sources
    take (32 output rows at 4096 x 2048, 64 at 4096^2), Kaiser(3) in every
    wrap mode."""
    W, H = size
    planar = np.random.default_rng(W + H).random(4 * W * H, dtype=np.float32)
    for wrap in (N.WRAP_CLAMP, N.WRAP_REPEAT, N.WRAP_MIRROR):
        a = gpu.downsample(N.FILTER_KAISER, planar, W, H, width=3.0, params=(4.0, 1.0), wrap=wrap)
        b = oracle.downsample(N.FILTER_KAISER, 3.0, (4.0, 1.0), wrap, planar, W, H)
        assert np.array_equal(_bits(a), _bits(b)), (size, wrap)


@pytest.mark.parametrize("size", [(64, 64), (37, 23), (5, 4)])
def test_box_with_transparency_uses_polyphase_box(gpu, oracle, size):
    """Surface.cpp:1356-1362: with AlphaMode_Transparency a Box mip goes through
    downSample(BoxFilter) (polyphase), not fastDownSample."""
    W, H = size
    planar = np.random.default_rng(2).random(4 * W * H, dtype=np.float32)
    a = gpu.downsample(N.FILTER_BOX, planar, W, H, alpha_mode=N.ALPHA_TRANSPARENCY)
    b = oracle.downsample(N.FILTER_BOX, 0.5, None, N.WRAP_MIRROR, planar, W, H)
    assert np.array_equal(_bits(a), _bits(b))


@pytest.mark.parametrize("size", [8, 64, 256, 512, 1024, 2048])
def test_bc3_lane_split_widths(gpu, oracle, size):
    """The BC3 colour search splits each block over 64, 32, ... 2 lanes as the
    image shrinks (one block per lane at 4096x4096).  Colours from a small
    palette give many partitions of equal error, so the winner on ties (the
    first in the reference's loop order) is exercised at every width."""
    rng = np.random.default_rng(size)
    pal = rng.integers(0, 256, (6, 3))
    img = np.empty((size, size, 4), np.uint8)
    img[:, :, :3] = pal[rng.integers(0, 6, (size, size))]
    img[:, :, 3] = rng.integers(0, 256, (size, size))
    planar = synth.to_planar(img).reshape(-1)
    for am in (N.ALPHA_NONE, N.ALPHA_TRANSPARENCY):
        got = gpu.compress(N.FORMAT_BC3, planar, size, size, alpha_mode=am)
        exp = oracle.compress(N.FORMAT_BC3, planar, size, size, alpha_mode=am)
        diff = np.nonzero(np.any(got.reshape(-1, 16) != exp.reshape(-1, 16), axis=1))[0]
        assert diff.size == 0, f"{diff.size} blocks differ (alpha mode {am}), first {diff[:5]}"


@pytest.mark.parametrize("size", [(256, 256), (64, 128), (2048, 1024)] + ODD_SIZES)
def test_bc3n_images_match_oracle(gpu, oracle, size):
    """Format_BC3n (CompressorDXT5n): normal maps and noise (whose red, blue
    and alpha must not leak into the colour block), both alpha modes, colour
    weights that the format overrides with (0, 1, 0); 2048x1024 takes the
    one-block-per-lane search, the others the lane-split one."""
    w, h = size
    for img in (synth.normal_map_rgba(w, h, seed=14), synth.noise_rect_rgba(w, h, seed=15, transparent=True)):
        planar = synth.to_planar(img).reshape(-1)
        for alpha_mode in (N.ALPHA_NONE, N.ALPHA_TRANSPARENCY):
            exp = oracle.compress(N.FORMAT_BC3N, planar, w, h, alpha_mode=alpha_mode)
            for cw in ((1, 1, 1, 1), (0.3, 0.59, 0.11, 1)):
                got = gpu.compress(N.FORMAT_BC3N, planar, w, h, alpha_mode=alpha_mode, color_weights=cw)
                diff = np.nonzero(np.any(got.reshape(-1, 16) != exp.reshape(-1, 16), axis=1))[0]
                assert diff.size == 0, f"{size} am {alpha_mode} cw {cw}: {diff.size} blocks differ, first {diff[:5]}"


def test_bc3n_single_green_blocks(gpu, oracle):
    """Blocks whose texels share one green but differ in red, blue and alpha
    take the single-colour path (compressDXT1G) on the device too."""
    rng = np.random.default_rng(16)
    w, h = 64, 64
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    for by in range(h // 4):
        for bx in range(w // 4):
            img[4 * by:4 * by + 4, 4 * bx:4 * bx + 4, 1] = (by * 16 + bx) % 256
    planar = synth.to_planar(img).reshape(-1)
    assert np.array_equal(gpu.compress(N.FORMAT_BC3N, planar, w, h), oracle.compress(N.FORMAT_BC3N, planar, w, h))


def test_bc3n_chain_matches_per_level(gpu, oracle):
    """A Kaiser chain of a normal map as BC3n through the chain scheduler."""
    w, h = 128, 96
    img = synth.normal_map_rgba(w, h, seed=17)
    planar = synth.to_planar(img).reshape(-1)
    sizes = N.mip_sizes(w, h)
    d_lv = [gpu.alloc(16 * a * b) for a, b in sizes]
    d_lv[0].upload(planar)
    d_out = [gpu.alloc(N.compressed_size(N.FORMAT_BC3N, a, b)) for a, b in sizes]
    gpu.compress_chain_device(N.FORMAT_BC3N, N.FILTER_KAISER, d_lv, w, h, d_out)
    for k, (a, b) in enumerate(sizes):
        lv = d_lv[k].download(np.float32)
        assert np.array_equal(d_out[k].download(np.uint8), oracle.compress(N.FORMAT_BC3N, lv, a, b)), (a, b)
    for d in d_lv + d_out:
        d.free()
