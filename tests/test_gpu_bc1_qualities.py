"""BC1 at Quality_Production (icbc's Quality_Level9: the Level8 cluster fit
followed by refine_endpoints, icbc.h:3329-3406) and Quality_Fastest
(Quality_Level1: box fit + least squares, icbc.h:2984-3153), on the device,
against the golden vectors written by the reference's own icbc and against
the oracle for images, adversarial blocks, chains, image sets and batches."""
import os

import numpy as np
import pytest

import blocks
import ntx_amd as N
from ntx_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
PROD, FAST = N.QUALITY_PRODUCTION, N.QUALITY_FASTEST


def _blocks_as_image(cols, wts):
    """Blocks side by side in one row of an image; weight = alpha
    (AlphaMode_Transparency gives each texel its weight)."""
    n = len(cols)
    w, h = 4 * n, 4
    img = np.zeros((4, h, w), np.float32)
    for i in range(n):
        c = np.asarray(cols[i], np.float32).reshape(4, 4, 4)
        img[:3, :, 4 * i:4 * i + 4] = np.moveaxis(c[..., :3], -1, 0)
        img[3, :, 4 * i:4 * i + 4] = np.asarray(wts[i], np.float32).reshape(4, 4)
    return img.reshape(-1), w, h


@pytest.mark.parametrize("name,quality", [("bc1_level9_blocks.npz", PROD), ("bc1_level1_blocks.npz", FAST)])
def test_quality_golden_blocks(gpu, name, quality):
    g = np.load(os.path.join(GOLD, name))
    cws = g["color_weights"]
    bad = 0
    for key in {tuple(c) for c in cws}:
        sel = np.nonzero(np.all(cws == np.array(key, np.float32), axis=1))[0]
        img, w, h = _blocks_as_image(g["colors"][sel], g["weights"][sel])
        out = gpu.compress(N.FORMAT_BC1, img, w, h, quality=quality, alpha_mode=N.ALPHA_TRANSPARENCY,
                           color_weights=(key[0], key[1], key[2], 1.0)).reshape(-1, 8)
        bad += int(np.sum(np.any(out != g["expected"][sel], axis=1)))
    assert bad == 0, f"{bad} of {len(cws)} golden blocks of {name} differ"


@pytest.mark.parametrize("quality", [PROD, FAST])
@pytest.mark.parametrize("kind", range(len(blocks.KINDS)))
def test_quality_adversarial_blocks(gpu, oracle, kind, quality):
    rng = np.random.default_rng(300 + kind)
    cols = np.zeros((256, 16, 4), np.float32)
    wts = np.zeros((256, 16), np.float32)
    for i in range(256):
        cols[i], wts[i] = blocks.bc1_block(rng, blocks.KINDS[kind])
    img, w, h = _blocks_as_image(cols, wts)
    for cw in ((1, 1, 1, 1), (0.5, 1.0, 0.25, 1.0)):
        a = gpu.compress(N.FORMAT_BC1, img, w, h, quality=quality, alpha_mode=N.ALPHA_TRANSPARENCY, color_weights=cw)
        b = oracle.compress_q(N.FORMAT_BC1, quality, img, w, h, alpha_mode=N.ALPHA_TRANSPARENCY, cw=cw)
        assert np.array_equal(a, b), (blocks.KINDS[kind], cw)


@pytest.mark.parametrize("size", [(256, 192), (37, 23), (5, 3), (1, 1)])
def test_quality_images_match_oracle(gpu, oracle, size):
    w, h = size
    planar = synth.to_planar(synth.noise_rect_rgba(w, h, seed=w + h, transparent=True)).reshape(-1)
    for q in (PROD, FAST):
        for am in (N.ALPHA_NONE, N.ALPHA_TRANSPARENCY):
            a = gpu.compress(N.FORMAT_BC1, planar, w, h, quality=q, alpha_mode=am)
            b = oracle.compress_q(N.FORMAT_BC1, q, planar, w, h, alpha_mode=am)
            assert np.array_equal(a, b), (size, q, am)
    # Normal and Highest stay Level8
    for q in (N.QUALITY_NORMAL, N.QUALITY_HIGHEST):
        assert np.array_equal(gpu.compress(N.FORMAT_BC1, planar, w, h, quality=q),
                              oracle.compress(N.FORMAT_BC1, planar, w, h))


@pytest.mark.parametrize("quality", [PROD, FAST])
def test_quality_chain_images_and_batch(gpu, oracle, quality):
    w, h = 96, 80
    img = synth.noise_rect_rgba(w, h, seed=12)
    planar = synth.to_planar(img).reshape(-1)
    sizes = N.mip_sizes(w, h)
    d_lv = [gpu.alloc(16 * a * b) for a, b in sizes]
    d_lv[0].upload(planar)
    d_out = [gpu.alloc(N.compressed_size(N.FORMAT_BC1, a, b)) for a, b in sizes]
    gpu.compress_chain_device(N.FORMAT_BC1, N.FILTER_BOX, d_lv, w, h, d_out, quality=quality)
    levels = [d.download(np.float32) for d in d_lv]
    chain = [d.download(np.uint8) for d in d_out]
    for k, (a, b) in enumerate(sizes):
        assert np.array_equal(chain[k], oracle.compress_q(N.FORMAT_BC1, quality, levels[k], a, b)), (a, b)
    # the image scheduler
    d_out2 = [gpu.alloc(N.compressed_size(N.FORMAT_BC1, a, b)) for a, b in sizes]
    gpu.compress_images_device(N.FORMAT_BC1, d_lv, sizes, d_out2, quality=quality)
    for k in range(len(sizes)):
        assert np.array_equal(d_out2[k].download(np.uint8), chain[k])
    # a batch of one texture with its chain
    tex = dict(d_rgba=d_lv[0], width=w, height=h, format=N.FORMAT_BC1, quality=quality, filter=N.FILTER_BOX)
    with N.Batch(gpu, [tex]) as b:
        b.run()
        for k in range(len(sizes)):
            assert np.array_equal(b.blocks(0, k), chain[k]), k
    for d in d_lv + d_out + d_out2:
        d.free()


def test_batch_mixing_qualities(gpu, oracle):
    """One batch holding BC1 textures at Fastest, Normal and Production and a
    BC4 texture (the batch kernel instance with every BC1 level, one shared
    cluster-fit path for levels 8 and 9): every level equals the oracle at
    its texture's quality."""
    texs, bufs, plan = [], [], []
    for i, (q, fmt, (w, h)) in enumerate([(FAST, N.FORMAT_BC1, (64, 48)), (N.QUALITY_NORMAL, N.FORMAT_BC1, (40, 72)),
                                          (PROD, N.FORMAT_BC1, (96, 64)), (N.QUALITY_NORMAL, N.FORMAT_BC4, (32, 32))]):
        img = synth.single_channel_rgba(w, h, seed=70 + i) if fmt == N.FORMAT_BC4 else \
            synth.noise_rect_rgba(w, h, seed=70 + i)
        planar = synth.to_planar(img).reshape(-1)
        d = gpu.alloc(planar.nbytes)
        d.upload(planar)
        bufs.append(d)
        texs.append(dict(d_rgba=d, width=w, height=h, format=fmt, quality=q, filter=N.FILTER_TRIANGLE))
        plan.append((q, fmt, w, h, planar))
    with N.Batch(gpu, texs) as b:
        b.run()
        for t, (q, fmt, w, h, planar) in enumerate(plan):
            lv = planar
            sizes = N.mip_sizes(w, h)
            for k, (a, c) in enumerate(sizes):
                if k > 0:
                    pw, ph = sizes[k - 1]
                    width, params = N.DEFAULT_FILTER[N.FILTER_TRIANGLE]
                    lv = oracle.build_next_mipmap(N.FILTER_TRIANGLE, width, params, N.WRAP_MIRROR, lv, pw, ph)
                exp = oracle.compress_q(fmt, q, lv, a, c) if fmt == N.FORMAT_BC1 else oracle.compress(fmt, lv, a, c)
                assert np.array_equal(b.blocks(t, k), exp), (t, k)
    for d in bufs:
        d.free()
