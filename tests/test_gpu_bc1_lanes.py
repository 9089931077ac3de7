"""GPU parity of the two-pass BC1 Level8 encoder (bc1_classify_kernel +
bc1_search_kernel, one block per lane over blocks bucketed by colour count).

A context with the threshold at one block sends every image through it; the
checks are the same as for the one-pass encoder: the reference icbc's golden
blocks, the oracle on adversarial blocks and seeded images (byte for byte),
the per-block colour counts (n, m) against the oracle's, and the two
encoders against each other on larger images."""
import os

import numpy as np
import pytest

import blocks
import ntx_amd as N
from ntx_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def lanes():
    ctx = N.Context(0)
    ctx.set_bc1_lanes_min(1)
    yield ctx
    ctx.close()


def _blocks_as_image(cols, wts):
    n = len(cols)
    img = np.zeros((4, 4, 4 * n), np.float32)
    for i in range(n):
        for t in range(16):
            y, x = divmod(t, 4)
            img[:3, y, 4 * i + x] = cols[i, t, :3]
            img[3, y, 4 * i + x] = wts[i, t]
    return img, 4 * n, 4


def _diff(a, b):
    return int(np.sum(np.any(a.reshape(-1, 8) != b.reshape(-1, 8), axis=1)))


def test_golden_blocks(lanes):
    """2048 blocks written by the reference's own icbc at Level8."""
    g = np.load(os.path.join(GOLD, "bc1_blocks.npz"))
    cws = g["color_weights"]
    bad = 0
    for key in {tuple(c) for c in cws}:
        sel = np.nonzero(np.all(cws == np.array(key, np.float32), axis=1))[0]
        img, w, h = _blocks_as_image(g["colors"][sel], g["weights"][sel])
        out = lanes.compress(N.FORMAT_BC1, img, w, h, alpha_mode=N.ALPHA_TRANSPARENCY,
                             color_weights=(key[0], key[1], key[2], 1.0))
        bad += _diff(out, g["expected"][sel])
    assert bad == 0, f"{bad} of {len(cws)} golden BC1 blocks differ"


@pytest.mark.parametrize("kind", range(len(blocks.KINDS)))
def test_adversarial_blocks(lanes, oracle, kind):
    rng = np.random.default_rng(300 + kind)
    n = 700  # several wavefronts per colour count, ragged tails
    cols = np.zeros((n, 16, 4), np.float32)
    wts = np.zeros((n, 16), np.float32)
    for i in range(n):
        cols[i], wts[i] = blocks.bc1_block(rng, blocks.KINDS[kind])
    img, w, h = _blocks_as_image(cols, wts)
    for mode in (N.ALPHA_TRANSPARENCY, N.ALPHA_NONE):
        for cw in ((1, 1, 1, 1), (0.5, 1.0, 0.25, 1.0)):
            a = lanes.compress(N.FORMAT_BC1, img, w, h, alpha_mode=mode, color_weights=cw)
            b = oracle.compress(N.FORMAT_BC1, img, w, h, alpha_mode=mode, cw=cw)
            assert np.array_equal(a, b), (blocks.KINDS[kind], mode, cw, _diff(a, b))


def test_mixed_kinds_in_one_image(lanes, oracle):
    """Every kind interleaved: waves of each colour count mix unit-weight,
    alpha-weighted and partial blocks."""
    rng = np.random.default_rng(77)
    n = 64 * 40
    cols = np.zeros((n, 16, 4), np.float32)
    wts = np.zeros((n, 16), np.float32)
    for i in range(n):
        cols[i], wts[i] = blocks.bc1_block(rng, blocks.KINDS[int(rng.integers(0, len(blocks.KINDS)))])
    img, w, h = _blocks_as_image(cols, wts)
    for mode in (N.ALPHA_TRANSPARENCY, N.ALPHA_NONE):
        a = lanes.compress(N.FORMAT_BC1, img, w, h, alpha_mode=mode)
        b = oracle.compress(N.FORMAT_BC1, img, w, h, alpha_mode=mode)
        assert np.array_equal(a, b), (mode, _diff(a, b))


@pytest.mark.parametrize("size", [(256, 256), (128, 64), (37, 23), (130, 66), (1, 1), (3, 5), (7, 1), (260, 9)])
@pytest.mark.parametrize("transparent", [False, True])
def test_images_match_oracle(lanes, oracle, size, transparent):
    w, h = size
    planar = synth.to_planar(synth.noise_rect_rgba(w, h, seed=41, transparent=transparent)).reshape(-1)
    for mode in (N.ALPHA_NONE, N.ALPHA_TRANSPARENCY):
        a = lanes.compress(N.FORMAT_BC1, planar, w, h, alpha_mode=mode)
        b = oracle.compress(N.FORMAT_BC1, planar, w, h, alpha_mode=mode)
        assert np.array_equal(a, b), (size, transparent, mode, _diff(a, b))


def test_dark_image(lanes, oracle):
    """Mostly black texels: skip_blacks drops clusters, the three-colour fit
    runs on a rebuilt table with its own count."""
    w, h = 192, 128
    img = synth.noise_rect_rgba(w, h, seed=5)
    img[:, :, :3] = (img[:, :, :3].astype(np.uint32) * 40 // 255).astype(np.uint8)
    planar = synth.to_planar(img).reshape(-1)
    a = lanes.compress(N.FORMAT_BC1, planar, w, h)
    b = oracle.compress(N.FORMAT_BC1, planar, w, h)
    assert np.array_equal(a, b), _diff(a, b)


def test_colour_counts_match_oracle(lanes, oracle):
    w, h = 160, 96
    planar = synth.to_planar(synth.noise_rect_rgba(w, h, seed=9)).reshape(-1)
    img = planar.astype(np.float32)
    d_in = lanes.alloc(img.nbytes).upload(img)
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    d_out = lanes.alloc(nb * 8)
    d_nm = lanes.alloc(nb * 4)
    lanes.compress_bc1_stats_device(d_in, w, h, d_out, d_nm)
    got = d_out.download(np.uint8)
    nm = d_nm.download(np.int16).reshape(-1, 2)
    exp, exp_nm = oracle.compress_bc1_counts(planar, w, h)
    assert np.array_equal(got, exp)
    assert np.array_equal(nm, exp_nm)
    for b in (d_in, d_out, d_nm):
        b.free()


def test_two_encoders_agree_on_large_images(lanes, gpu):
    """1024^2 (65,536 blocks) and 2048x512, default context (one-pass below
    2^17 blocks) against the two-pass context."""
    for (w, h, seed) in ((1024, 1024, 3), (2048, 512, 4)):
        planar = synth.to_planar(synth.noise_rect_rgba(w, h, seed=seed, transparent=seed == 4)).reshape(-1)
        for mode in (N.ALPHA_NONE, N.ALPHA_TRANSPARENCY):
            a = lanes.compress(N.FORMAT_BC1, planar, w, h, alpha_mode=mode)
            b = gpu.compress(N.FORMAT_BC1, planar, w, h, alpha_mode=mode)
            assert np.array_equal(a, b), (w, h, mode, _diff(a, b))


def test_repeated_calls_reuse_cleared_tallies(lanes, oracle):
    """The search's last workgroup clears the per-count tallies: a second and
    a smaller third image on the same scratch give the same bytes."""
    res = []
    for (w, h) in ((96, 96), (96, 96), (40, 24)):
        planar = synth.to_planar(synth.noise_rect_rgba(w, h, seed=w + h)).reshape(-1)
        a = lanes.compress(N.FORMAT_BC1, planar, w, h)
        assert np.array_equal(a, oracle.compress(N.FORMAT_BC1, planar, w, h)), (w, h)
        res.append(a)
    assert np.array_equal(res[0], res[1])


def test_chain_and_image_set(lanes, oracle):
    """The chain entry point and the image scheduler (per-image scratch
    regions) through the two-pass encoder."""
    w = h = 128
    planar = synth.to_planar(synth.noise_rect_rgba(w, h, seed=8)).reshape(-1)
    levels = [planar]
    sizes = [(w, h)]
    while sizes[-1] != (1, 1):
        W, H = sizes[-1]
        levels.append(oracle.build_next_mipmap(N.FILTER_TRIANGLE, 1.0, None, N.WRAP_MIRROR, levels[-1], W, H))
        sizes.append((max(1, W // 2), max(1, H // 2)))
    d_lv = [lanes.alloc(4 * 4 * a * b) for (a, b) in sizes]
    d_lv[0].upload(planar)
    d_out = [lanes.alloc(N.compressed_size(N.FORMAT_BC1, a, b)) for (a, b) in sizes]
    lanes.compress_chain_device(N.FORMAT_BC1, N.FILTER_TRIANGLE, d_lv, w, h, d_out)
    for k, (a, b) in enumerate(sizes):
        exp = oracle.compress(N.FORMAT_BC1, levels[k], a, b)
        assert np.array_equal(d_out[k].download(np.uint8), exp), (a, b)
    for buf in d_lv + d_out:
        buf.free()
