"""The Python binding refuses device buffers too small for the extent a call
implies, before any device call: a short buffer would otherwise become an
out-of-bounds device access.  Runs without a GPU: the buffers and the
context are stand-ins whose library raises if it is ever reached."""
import numpy as np
import pytest

import ntx_amd as N


class _NoLib:
    def __getattr__(self, name):
        raise AssertionError(f"{name} reached the library")


def _ctx():
    c = N.Context.__new__(N.Context)
    c.lib = _NoLib()
    c.handle = 1
    return c


def _buf(ctx, nbytes, ptr=0x1000):
    b = N.DeviceBuffer.__new__(N.DeviceBuffer)
    b.ctx, b.nbytes, b.ptr = ctx, nbytes, ptr
    return b


def test_need_rejects_short_and_freed_buffers():
    c = _ctx()
    N._need(_buf(c, 64), 64, "ok")
    with pytest.raises(ValueError):
        N._need(_buf(c, 63), 64, "short")
    with pytest.raises(ValueError):
        N._need(_buf(c, 64, ptr=None), 1, "freed")
    with pytest.raises(ValueError):
        N._need(object(), 1, "not a buffer")


def test_device_calls_check_extents():
    c = _ctx()
    w, h = 37, 23
    img = _buf(c, 16 * w * h)
    blocks = _buf(c, N.compressed_size(N.FORMAT_BC1, w, h))
    with pytest.raises(ValueError):
        c.compress_device(N.FORMAT_BC1, _buf(c, 16 * w * h - 4), w, h, blocks)
    with pytest.raises(ValueError):
        c.compress_device(N.FORMAT_BC3, img, w, h, blocks)  # BC3 needs 16 B per block
    with pytest.raises(ValueError):
        c.downsample_device(N.FILTER_BOX, img, w, h, _buf(c, 16 * (w // 2) * (h // 2) - 1))
    with pytest.raises(ValueError):
        c.decode_device(N.FORMAT_BC1, _buf(c, 8), w, h, img)
    with pytest.raises(ValueError):
        c.compress_bc1_stats_device(img, w, h, blocks, _buf(c, 4))
    with pytest.raises(ValueError):
        c.unpack_bgra8_device(_buf(c, 4 * w * h - 1), w, h, img)
    with pytest.raises(ValueError):
        c.compress_device(N.FORMAT_BC1, img, -1, h, blocks)
    lv = [_buf(c, 16 * a * b) for a, b in N.mip_sizes(w, h)]
    outs = [_buf(c, N.compressed_size(N.FORMAT_BC1, a, b)) for a, b in N.mip_sizes(w, h)]
    outs[3] = _buf(c, 0)
    with pytest.raises(ValueError):
        c.compress_chain_device(N.FORMAT_BC1, N.FILTER_BOX, lv, w, h, outs)


def test_host_calls_check_sizes():
    c = _ctx()
    with pytest.raises(ValueError):
        c.compress(N.FORMAT_BC1, np.zeros(4 * 8 * 8 - 1, np.float32), 8, 8)
    with pytest.raises(ValueError):
        c.downsample(N.FILTER_BOX, np.zeros(10, np.float32), 8, 8)
    b = _buf(c, 16)
    with pytest.raises(ValueError):
        b.download(np.float32, 5)
    with pytest.raises(ValueError):
        b.upload(np.zeros(5, np.float32))
