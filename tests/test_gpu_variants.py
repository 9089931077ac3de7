"""GPU parity of the Production / Highest / Fastest compressors of BC3, BC3n,
BC4 and BC5 (Context.cpp:1075-1116) against the oracle, which is pinned to
the reference's own QuickCompress / OptimalCompress code in
test_oracle_vs_reference.py.  Bit-exact output is the bar.  The reference's
OptimalCompress::compressDXT5A seed (OptimalCompressDXT.cpp:546) is fixed at
the error of (max, min) on both sides."""
import numpy as np
import pytest

import blocks
import ntx_amd as N
from ntx_amd import synth

pytestmark = pytest.mark.gpu

VARIANTS = [(N.FORMAT_BC3, N.QUALITY_FASTEST), (N.FORMAT_BC3, N.QUALITY_HIGHEST),
            (N.FORMAT_BC3N, N.QUALITY_FASTEST), (N.FORMAT_BC3N, N.QUALITY_HIGHEST),
            (N.FORMAT_BC4, N.QUALITY_PRODUCTION), (N.FORMAT_BC4, N.QUALITY_HIGHEST),
            (N.FORMAT_BC5, N.QUALITY_PRODUCTION), (N.FORMAT_BC5, N.QUALITY_HIGHEST)]
SIZES = [(64, 64), (37, 23), (5, 3), (1, 1), (7, 1), (256, 128)]


def _inputs(fmt, w, h):
    if fmt == N.FORMAT_BC3:
        return [synth.noise_rect_rgba(w, h, seed=31, transparent=True)]
    if fmt == N.FORMAT_BC3N:
        return [synth.normal_map_rgba(w, h, seed=32), synth.noise_rect_rgba(w, h, seed=33, transparent=True)]
    if fmt == N.FORMAT_BC4:
        return [synth.single_channel_rgba(w, h, seed=34), synth.noise_rect_rgba(w, h, seed=35)]
    return [synth.normal_map_rgba(w, h, seed=36), synth.noise_rect_rgba(w, h, seed=37)]


@pytest.mark.parametrize("fmt,quality", VARIANTS)
@pytest.mark.parametrize("size", SIZES)
def test_variant_images_match_oracle(gpu, oracle, fmt, quality, size):
    w, h = size
    for img in _inputs(fmt, w, h):
        planar = synth.to_planar(img).reshape(-1)
        for am in (N.ALPHA_NONE, N.ALPHA_TRANSPARENCY):
            exp = oracle.compress_q(fmt, quality, planar, w, h, alpha_mode=am)
            got = gpu.compress(fmt, planar, w, h, quality=quality, alpha_mode=am)
            bs = 8 if fmt == N.FORMAT_BC4 else 16
            diff = np.nonzero(np.any(got.reshape(-1, bs) != exp.reshape(-1, bs), axis=1))[0]
            assert diff.size == 0, f"{diff.size} blocks differ (alpha mode {am}), first {diff[:5]}"


def _channel_image(v):
    """Channel blocks (n x 16 bytes) side by side in a 4 x 4n image, in red
    and green (green reversed) with alpha = red."""
    n = len(v)
    img = np.zeros((4, 4 * n, 4), np.uint8)
    for t in range(16):
        y, x = divmod(t, 4)
        img[y, x::4, 0] = v[:, t]
        img[y, x::4, 1] = v[::-1, t]
        img[y, x::4, 3] = v[:, t]
    return img


@pytest.mark.parametrize("fmt,quality", VARIANTS)
def test_variant_adversarial_channel_blocks(gpu, oracle, fmt, quality):
    """Full-range noise, clusters, two and three levels, 0/255 with a middle
    value, ramps, near-constant and constant channels: both OptimalCompress
    search loops, the seed fallback and the early-return branches."""
    v = blocks.channel_set(77 + fmt + quality, 2048)
    img = _channel_image(v)
    h, w = img.shape[:2]
    planar = synth.to_planar(img).reshape(-1)
    exp = oracle.compress_q(fmt, quality, planar, w, h)
    got = gpu.compress(fmt, planar, w, h, quality=quality)
    bs = 8 if fmt == N.FORMAT_BC4 else 16
    diff = np.nonzero(np.any(got.reshape(-1, bs) != exp.reshape(-1, bs), axis=1))[0]
    assert diff.size == 0, f"{diff.size} blocks differ, first {diff[:5]}"


def test_bc3n_highest_green_edge_cases(gpu, oracle):
    """compressDXT1G: greens of 255 (quantised to 64, stored as 0 in the
    6-bit field), near-flat greens that keep the seed, single greens with
    other channels varying, and random greens."""
    rng = np.random.default_rng(91)
    n = 1024
    img = rng.integers(0, 256, (4, 4 * n, 4), dtype=np.uint8)
    g = img[:, :, 1].reshape(4, n, 4)
    g[:, 0::4, :] = 255 - rng.integers(0, 2, (4, n // 4, 4))
    g[:, 1::4, :] = 128 + rng.integers(0, 3, (4, n // 4, 4))
    g[:, 2::4, :] = rng.integers(0, 256, (1, n // 4, 1))
    img[:, :, 1] = g.reshape(4, 4 * n)
    planar = synth.to_planar(img).reshape(-1)
    exp = oracle.compress_q(N.FORMAT_BC3N, N.QUALITY_HIGHEST, planar, 4 * n, 4)
    got = gpu.compress(N.FORMAT_BC3N, planar, 4 * n, 4, quality=N.QUALITY_HIGHEST)
    diff = np.nonzero(np.any(got.reshape(-1, 16) != exp.reshape(-1, 16), axis=1))[0]
    assert diff.size == 0, f"{diff.size} blocks differ, first {diff[:5]}"


@pytest.mark.parametrize("fmt,quality", [(N.FORMAT_BC4, N.QUALITY_PRODUCTION), (N.FORMAT_BC3, N.QUALITY_HIGHEST),
                                         (N.FORMAT_BC3N, N.QUALITY_FASTEST)])
def test_variant_large_image(gpu, oracle, fmt, quality):
    """1024 x 1024 (65,536 blocks): the grid-stride loops cover every block."""
    w = h = 1024
    img = synth.noise_rect_rgba(w, h, seed=41, transparent=True)
    planar = synth.to_planar(img).reshape(-1)
    exp = oracle.compress_q(fmt, quality, planar, w, h)
    got = gpu.compress(fmt, planar, w, h, quality=quality)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("fmt,quality", [(N.FORMAT_BC5, N.QUALITY_HIGHEST), (N.FORMAT_BC3, N.QUALITY_FASTEST),
                                         (N.FORMAT_BC3N, N.QUALITY_HIGHEST), (N.FORMAT_BC3, N.QUALITY_HIGHEST)])
def test_variant_chain_and_batch_match_per_level(gpu, oracle, fmt, quality):
    """A Kaiser chain through the chain scheduler and the levels again through
    compress_images_device, each level against the oracle."""
    w, h = 96, 80
    img = synth.normal_map_rgba(w, h, seed=43) if fmt != N.FORMAT_BC3 else synth.noise_rect_rgba(w, h, seed=44,
                                                                                                     transparent=True)
    planar = synth.to_planar(img).reshape(-1)
    sizes = N.mip_sizes(w, h)
    d_lv = [gpu.alloc(16 * a * b) for a, b in sizes]
    d_lv[0].upload(planar)
    d_out = [gpu.alloc(N.compressed_size(fmt, a, b)) for a, b in sizes]
    gpu.compress_chain_device(fmt, N.FILTER_KAISER, d_lv, w, h, d_out, quality=quality)
    levels = [d.download(np.float32) for d in d_lv]
    for k, (a, b) in enumerate(sizes):
        assert np.array_equal(d_out[k].download(np.uint8), oracle.compress_q(fmt, quality, levels[k], a, b)), (a, b)
    d_out2 = [gpu.alloc(N.compressed_size(fmt, a, b)) for a, b in sizes]
    gpu.compress_images_device(fmt, d_lv, sizes, d_out2, quality=quality)
    for k in range(len(sizes)):
        assert np.array_equal(d_out2[k].download(np.uint8), d_out[k].download(np.uint8)), sizes[k]
    for d in d_lv + d_out + d_out2:
        d.free()


# ------------------------------------------------------------------ BC1a --

def _bc1a_blocks(seed, n):
    """CompressorDXT1a's cases: several colours with some alpha-0 texels,
    one RGB with transparent texels, one RGB with alphas 0 / 128 / 255, one
    RGB all transparent, several RGB all transparent, and opaque noise."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 16, 4), np.uint8)
    for i in range(n):
        b = rng.integers(0, 256, (16, 4)).astype(np.uint8)
        k = i % 6
        if k == 0:
            b[:, 3] = np.where(rng.random(16) < 0.3, 0, 255)
        elif k == 1:
            b[:] = b[0]
            b[:, 3] = np.where(rng.random(16) < 0.4, 0, rng.integers(1, 256))
        elif k == 2:
            b[:, :3] = b[0, :3]
            b[:, 3] = rng.choice([0, 128, 255], 16)
        elif k == 3:
            b[:] = b[0]
            b[:, 3] = 0
        elif k == 4:
            b[:, 3] = 0
        else:
            b[:, 3] = 255
            b[:, :3] = np.clip(b[0, :3].astype(int) + rng.integers(-20, 21, (16, 3)), 0, 255)
        out[i] = b
    return out


def _blocks_rgba_image(px):
    """n blocks of 16 RGBA texels (R, G, B, A order) side by side in a 4 x 4n image."""
    n = len(px)
    img = np.zeros((4, 4 * n, 4), np.uint8)
    for t in range(16):
        y, x = divmod(t, 4)
        img[y, x::4] = px[:, t]
    return img


@pytest.mark.parametrize("quality", [N.QUALITY_NORMAL, N.QUALITY_PRODUCTION, N.QUALITY_HIGHEST])
@pytest.mark.parametrize("size", SIZES + [(1024, 512)])
def test_bc1a_images_match_oracle(gpu, oracle, quality, size):
    w, h = size
    img = synth.noise_rect_rgba(w, h, seed=51, transparent=True)
    planar = synth.to_planar(img).reshape(-1)
    for am in (N.ALPHA_NONE, N.ALPHA_TRANSPARENCY):
        for cw in ((1, 1, 1, 1), (0.3, 0.59, 0.11, 1)):
            exp = oracle.compress_q(N.FORMAT_BC1A, quality, planar, w, h, alpha_mode=am, cw=cw)
            got = gpu.compress(N.FORMAT_BC1A, planar, w, h, quality=quality, alpha_mode=am, color_weights=cw)
            diff = np.nonzero(np.any(got.reshape(-1, 8) != exp.reshape(-1, 8), axis=1))[0]
            assert diff.size == 0, f"am {am} cw {cw}: {diff.size} blocks differ, first {diff[:5]}"


def test_bc1a_adversarial_blocks(gpu, oracle):
    px = _bc1a_blocks(52, 3000)
    img = _blocks_rgba_image(px)
    h, w = img.shape[:2]
    planar = synth.to_planar(img).reshape(-1)
    for am in (N.ALPHA_NONE, N.ALPHA_TRANSPARENCY):
        exp = oracle.compress_q(N.FORMAT_BC1A, N.QUALITY_NORMAL, planar, w, h, alpha_mode=am)
        got = gpu.compress(N.FORMAT_BC1A, planar, w, h, alpha_mode=am)
        diff = np.nonzero(np.any(got.reshape(-1, 8) != exp.reshape(-1, 8), axis=1))[0]
        assert diff.size == 0, f"am {am}: {diff.size} blocks differ, first {diff[:5]} (kinds {diff[:5] % 6})"


def test_bc1a_chain_and_fastest(gpu, oracle):
    """A Kaiser chain of a transparent texture as BC1a through the chain
    scheduler; Fastest (FastCompressorDXT1a) is refused."""
    w, h = 96, 80
    planar = synth.to_planar(synth.noise_rect_rgba(w, h, seed=53, transparent=True)).reshape(-1)
    sizes = N.mip_sizes(w, h)
    d_lv = [gpu.alloc(16 * a * b) for a, b in sizes]
    d_lv[0].upload(planar)
    d_out = [gpu.alloc(N.compressed_size(N.FORMAT_BC1A, a, b)) for a, b in sizes]
    gpu.compress_chain_device(N.FORMAT_BC1A, N.FILTER_KAISER, d_lv, w, h, d_out, alpha_mode=N.ALPHA_TRANSPARENCY)
    for k, (a, b) in enumerate(sizes):
        lv = d_lv[k].download(np.float32)
        exp = oracle.compress_q(N.FORMAT_BC1A, N.QUALITY_NORMAL, lv, a, b, alpha_mode=N.ALPHA_TRANSPARENCY)
        assert np.array_equal(d_out[k].download(np.uint8), exp), (a, b)
    for d in d_lv + d_out:
        d.free()
    with pytest.raises(N.NtxError) as e:
        gpu.compress(N.FORMAT_BC1A, planar, w, h, quality=N.QUALITY_FASTEST)
    assert e.value.status == 2
