"""ntx_hip_compress_device_noalloc (no device allocation; the library's own
scratch array) against ntx_hip_compress_device, byte for byte, on two calls
in a row, the second with a smaller image, for every encoder that uses the
scratch (BC3, BC3n, two-pass BC1) and one that does not (BC4); the size
limit and BC1's fallback to its one-pass encoder above it."""
import numpy as np
import pytest

import ntx_amd as N
from ntx_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = N.Context(0)
    yield c
    c.close()


def _both(ctx, fmt, w, h, seed, **kw):
    img = synth.noise_rect_rgba(w, h, seed=seed, transparent=fmt == N.FORMAT_BC3)
    planar = synth.to_planar(img).reshape(-1)
    d_in = ctx.alloc(planar.nbytes).upload(planar)
    size = N.compressed_size(fmt, w, h)
    d_a, d_b = ctx.alloc(size), ctx.alloc(size)
    ctx.compress_device(fmt, d_in, w, h, d_a, **kw)
    ctx.compress_device_noalloc(fmt, d_in, w, h, d_b, **kw)
    a, b = d_a.download(np.uint8), d_b.download(np.uint8)
    for buf in (d_in, d_a, d_b):
        buf.free()
    return a, b


@pytest.mark.parametrize("fmt,alpha", [(N.FORMAT_BC3, N.ALPHA_TRANSPARENCY), (N.FORMAT_BC3, N.ALPHA_NONE),
                                       (N.FORMAT_BC3N, N.ALPHA_NONE), (N.FORMAT_BC4, N.ALPHA_NONE)])
def test_matches_compress_device(ctx, fmt, alpha):
    for (w, h, seed) in ((256, 256, 61), (120, 66, 62)):
        a, b = _both(ctx, fmt, w, h, seed, alpha_mode=alpha)
        assert np.array_equal(a, b), (fmt, w, h)


def test_two_pass_bc1(ctx):
    ctx.set_bc1_lanes_min(1)
    try:
        for (w, h, seed) in ((512, 256, 63), (72, 40, 64)):
            a, b = _both(ctx, N.FORMAT_BC1, w, h, seed)
            assert np.array_equal(a, b), (w, h)
    finally:
        ctx.set_bc1_lanes_min(0)


def test_size_limit(ctx):
    w, h = 1040, 1024  # 66,560 blocks
    d_in = ctx.alloc(16 * w * h)
    d_out = ctx.alloc(N.compressed_size(N.FORMAT_BC3, w, h))
    with pytest.raises(N.NtxError) as e:
        ctx.compress_device_noalloc(N.FORMAT_BC3, d_in, w, h, d_out)
    assert e.value.status == 1  # NTX_ERR_INVALID_ARGUMENT
    d_in.free()
    d_out.free()
    # BC1 above the limit takes its one-pass encoder, which needs no scratch.
    ctx.set_bc1_lanes_min(1)
    try:
        a, b = _both(ctx, N.FORMAT_BC1, w, h, 65)
        assert np.array_equal(a, b)
    finally:
        ctx.set_bc1_lanes_min(0)
