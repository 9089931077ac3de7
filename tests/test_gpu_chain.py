"""GPU tests of ntx_hip_compress_chain_device: a whole mip chain filtered and
compressed with each level's compression on its own stream must give the
same level images (FP32 bits) and the same blocks as the per-level calls
(themselves checked against the oracle and the reference's vectors)."""
import numpy as np
import pytest

import ntx_amd as N
from ntx_amd import synth

pytestmark = pytest.mark.gpu


def _image(fmt, w, h, seed):
    if fmt == N.FORMAT_BC5:
        img = synth.normal_map_rgba(w, h, seed=seed)
    elif fmt == N.FORMAT_BC4:
        img = synth.single_channel_rgba(w, h, seed=seed)
    else:
        img = synth.noise_rect_rgba(w, h, seed=seed, transparent=True)
    return synth.to_planar(img).reshape(-1)


def _per_level(gpu, fmt, filt, planar, w, h, levels, alpha_mode):
    sizes, lv, blocks = [], [planar], []
    for k in range(levels):
        sizes.append((w, h))
        d = gpu.alloc(lv[k].nbytes).upload(lv[k])
        out = gpu.alloc(N.compressed_size(fmt, w, h))
        gpu.compress_device(fmt, d, w, h, out, alpha_mode=alpha_mode)
        blocks.append(out.download(np.uint8))
        if k + 1 < levels:
            lv.append(gpu.downsample(filt, lv[k], w, h, alpha_mode=alpha_mode))
            w, h = max(1, w // 2), max(1, h // 2)
        d.free(); out.free()
    return sizes, lv, blocks


def _chain(gpu, fmt, filt, planar, sizes, alpha_mode):
    d_lv = [gpu.alloc(16 * w * h) for w, h in sizes]
    d_lv[0].upload(planar)
    d_out = [gpu.alloc(N.compressed_size(fmt, w, h)) for w, h in sizes]
    w, h = sizes[0]
    gpu.compress_chain_device(fmt, filt, d_lv, w, h, d_out, alpha_mode=alpha_mode)
    lv = [d.download(np.float32) for d in d_lv]
    blocks = [d.download(np.uint8) for d in d_out]
    for d in d_lv + d_out:
        d.free()
    return lv, blocks


@pytest.mark.parametrize("fmt", [N.FORMAT_BC1, N.FORMAT_BC3, N.FORMAT_BC4, N.FORMAT_BC5])
@pytest.mark.parametrize("filt", [N.FILTER_BOX, N.FILTER_KAISER])
@pytest.mark.parametrize("size,levels", [((256, 128), None), ((37, 23), None), ((64, 64), 3), ((1, 1), None)])
def test_chain_matches_per_level_calls(gpu, fmt, filt, size, levels):
    w, h = size
    full = len(N.mip_sizes(w, h))
    levels = full if levels is None else levels
    planar = _image(fmt, w, h, seed=w + h + fmt)
    for alpha_mode in ((N.ALPHA_NONE, N.ALPHA_TRANSPARENCY) if fmt == N.FORMAT_BC3 else (N.ALPHA_NONE,)):
        sizes, lv_ref, bl_ref = _per_level(gpu, fmt, filt, planar, w, h, levels, alpha_mode)
        lv, bl = _chain(gpu, fmt, filt, planar, sizes, alpha_mode)
        for k in range(levels):
            assert np.array_equal(lv[k].view(np.uint32), lv_ref[k].view(np.uint32)), (k, alpha_mode)
            assert np.array_equal(bl[k], bl_ref[k]), (k, alpha_mode)


def test_chain_scratch_reuse_across_layouts(gpu):
    """Chains of different shapes one after another on one context (the BC3
    tallies of each level's scratch region must start at zero every time),
    interleaved with per-level calls that use the context's other scratch."""
    cases = [(512, 512, None), (96, 40, 4), (512, 512, None), (512, 512, 6), (33, 65, None)]
    for i, (w, h, levels) in enumerate(cases):
        full = len(N.mip_sizes(w, h))
        levels = full if levels is None else levels
        planar = _image(N.FORMAT_BC3, w, h, seed=100 + i)
        sizes, _, bl_ref = _per_level(gpu, N.FORMAT_BC3, N.FILTER_KAISER, planar, w, h, levels, N.ALPHA_NONE)
        for rep in range(2):
            _, bl = _chain(gpu, N.FORMAT_BC3, N.FILTER_KAISER, planar, sizes, N.ALPHA_NONE)
            for k in range(levels):
                assert np.array_equal(bl[k], bl_ref[k]), (w, h, levels, rep, k)


def test_chain_rejects_bad_arguments(gpu):
    d = gpu.alloc(16 * 8 * 8)
    out = gpu.alloc(N.compressed_size(N.FORMAT_BC1, 8, 8))
    with pytest.raises(N.NtxError) as e:  # 8x8 has 4 levels
        gpu.compress_chain_device(N.FORMAT_BC1, N.FILTER_BOX, [d] * 5, 8, 8, [out] * 5)
    assert e.value.status == 1
    with pytest.raises(N.NtxError) as e:
        gpu.compress_chain_device(N.FORMAT_BC4, N.FILTER_BOX, [d], 8, 8, [out], quality=4)  # no such quality
    assert e.value.status == 2
    d.free(); out.free()


@pytest.mark.parametrize("filt", [N.FILTER_BOX, N.FILTER_TRIANGLE, N.FILTER_KAISER])
@pytest.mark.parametrize("wrap", [N.WRAP_CLAMP, N.WRAP_REPEAT, N.WRAP_MIRROR])
@pytest.mark.parametrize("size", [(512, 256), (130, 66), (37, 23), (1, 300), (300, 1), (96, 1), (2, 2), (1, 1),
                                  (90, 90), (129, 63), (1, 8192), (8193, 1), (256, 33)])
def test_downsample_chain_matches_per_level(gpu, filt, wrap, size):
    """ntx_hip_downsample_chain_device against per-level downsampling, FP32
    bit patterns, including 1-D images and the Transparency box (polyphase)
    case.  Levels from a source of at most 8192 texels run in one launch
    (the tail kernel); the sizes put that source at level 0, at an odd
    extent, at exactly 8192 texels and just above."""
    w, h = size
    planar = np.random.default_rng(w * 7 + h).random(4 * w * h, dtype=np.float32)
    sizes = N.mip_sizes(w, h)
    for alpha_mode in ((N.ALPHA_NONE, N.ALPHA_TRANSPARENCY) if filt == N.FILTER_BOX else (N.ALPHA_NONE,)):
        ref = [planar]
        for k in range(1, len(sizes)):
            ref.append(gpu.downsample(filt, ref[-1], *sizes[k - 1], wrap=wrap, alpha_mode=alpha_mode))
        d_lv = [gpu.alloc(16 * a * b) for a, b in sizes]
        d_lv[0].upload(planar)
        gpu.downsample_chain_device(filt, d_lv, w, h, wrap=wrap, alpha_mode=alpha_mode)
        for k, d in enumerate(d_lv):
            got = d.download(np.float32)
            assert np.array_equal(got.view(np.uint32), ref[k].view(np.uint32)), (k, sizes[k], alpha_mode)
            d.free()


@pytest.mark.parametrize("fmt", [N.FORMAT_BC1, N.FORMAT_BC3, N.FORMAT_BC4, N.FORMAT_BC5])
def test_images_match_per_image_calls(gpu, fmt):
    """ntx_hip_compress_images_device: unordered images of mixed sizes
    (largest scheduled first, the rest on the side streams) give the
    per-image bytes; run twice to reuse the per-image scratch regions."""
    sizes = [(37, 23), (256, 192), (1, 1), (64, 64), (130, 66), (8, 4), (512, 128)]
    imgs = [_image(fmt, w, h, seed=3 * i + fmt) for i, (w, h) in enumerate(sizes)]
    d_in = [gpu.alloc(x.nbytes).upload(x) for x in imgs]
    ref = []
    for d, (w, h) in zip(d_in, sizes):
        out = gpu.alloc(N.compressed_size(fmt, w, h))
        gpu.compress_device(fmt, d, w, h, out)
        ref.append(out.download(np.uint8))
        out.free()
    d_out = [gpu.alloc(N.compressed_size(fmt, w, h)) for w, h in sizes]
    for rep in range(2):
        gpu.compress_images_device(fmt, d_in, sizes, d_out)
        for k, d in enumerate(d_out):
            assert np.array_equal(d.download(np.uint8), ref[k]), (rep, sizes[k])
    for d in d_in + d_out:
        d.free()


def test_scheduler_scratch_reused_across_formats():
    """One context compresses the same set of level sizes as BC3 (regions
    laid out for BC3's scratch), then as BC1 with the two-pass encoder
    (regions of other sizes at other offsets), then as BC3 again: each run
    must find its own tallies zero, and every level must equal the per-level
    calls."""
    ctx = N.Context(0)
    ctx.set_bc1_lanes_min(1)
    try:
        w = h = 128
        sizes = N.mip_sizes(w, h)
        planar = _image(N.FORMAT_BC3, w, h, seed=77)
        d_lv = [ctx.alloc(16 * a * b) for a, b in sizes]
        d_lv[0].upload(planar)
        ctx.downsample_chain_device(N.FILTER_BOX, d_lv, w, h)
        lv = [d.download(np.float32) for d in d_lv]
        for fmt in (N.FORMAT_BC3, N.FORMAT_BC1, N.FORMAT_BC3, N.FORMAT_BC1):
            d_out = [ctx.alloc(N.compressed_size(fmt, a, b)) for a, b in sizes]
            ctx.compress_images_device(fmt, d_lv, sizes, d_out)
            for k, (a, b) in enumerate(sizes):
                exp = ctx.compress(fmt, lv[k], a, b)
                assert np.array_equal(d_out[k].download(np.uint8), exp), (fmt, a, b)
            for d in d_out:
                d.free()
        for d in d_lv:
            d.free()
    finally:
        ctx.close()
