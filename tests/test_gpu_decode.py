"""GPU parity of ntx_hip_decode_device (Surface::setImage2D for BC1/BC3/BC4/BC5,
D3D10 and D3D9 decoders) and ntx_hip_rms_error_device (nvtt::rmsError /
rmsAlphaError) against the oracle."""
import numpy as np
import pytest

import ntx_amd as N
from ntx_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", [N.FORMAT_BC1, N.FORMAT_BC3, N.FORMAT_BC4, N.FORMAT_BC5])
@pytest.mark.parametrize("decoder", [N.DECODER_D3D10, N.DECODER_D3D9])
@pytest.mark.parametrize("size", [(64, 32), (37, 23), (1, 1), (130, 5)])
def test_decode_matches_oracle(gpu, oracle, fmt, decoder, size):
    w, h = size
    nbytes = N.compressed_size(fmt, w, h)
    raw = np.random.default_rng(w * h + fmt).integers(0, 256, nbytes, dtype=np.uint8)
    d_blk = gpu.alloc(nbytes).upload(raw)
    d_out = gpu.alloc(16 * w * h)
    gpu.decode_device(fmt, d_blk, w, h, d_out, decoder=decoder)
    got = d_out.download(np.float32)
    exp = oracle.decode_image(fmt, raw, w, h, d3d9=(decoder == N.DECODER_D3D9))
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    d_blk.free(); d_out.free()


def test_decode_of_compressed_levels(gpu, oracle):
    """Decoding the device's own BC1/BC3 output: what Surface::setImage2D of the
    compressed bytes gives, and its error against the source."""
    w, h = 96, 64
    planar = synth.to_planar(synth.noise_rect_rgba(w, h, seed=3, transparent=True)).reshape(-1)
    d_src = gpu.alloc(planar.nbytes).upload(planar)
    for fmt in (N.FORMAT_BC1, N.FORMAT_BC3):
        d_blk = gpu.alloc(N.compressed_size(fmt, w, h))
        gpu.compress_device(fmt, d_src, w, h, d_blk)
        d_dec = gpu.alloc(16 * w * h)
        gpu.decode_device(fmt, d_blk, w, h, d_dec)
        dec = d_dec.download(np.float32)
        assert np.array_equal(dec, oracle.decode_image(fmt, d_blk.download(np.uint8), w, h))
        for metric in (N.METRIC_RMS_COLOR, N.METRIC_RMS_COLOR_ALPHA_WEIGHTED, N.METRIC_RMS_ALPHA):
            got = gpu.rms_error_device(metric, d_src, d_dec, w, h)
            exp = oracle.rms_error(metric, planar, dec)
            assert got == pytest.approx(exp, rel=2e-7, abs=0), (fmt, metric)
        d_blk.free(); d_dec.free()
    d_src.free()


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_rms_error_large(gpu, oracle, metric):
    w, h = 1024, 777
    rng = np.random.default_rng(metric)
    a = rng.random(4 * w * h, dtype=np.float32)
    b = (a + rng.normal(0, 0.01, a.size)).astype(np.float32)
    da, db = gpu.alloc(a.nbytes).upload(a), gpu.alloc(b.nbytes).upload(b)
    got = gpu.rms_error_device(metric, da, db, w, h)
    exp = oracle.rms_error(metric, a, b)
    assert got == pytest.approx(exp, rel=2e-7, abs=0)
    assert gpu.rms_error_device(metric, da, da, w, h) == 0.0
    da.free(); db.free()


def test_bc3n_decodes_as_bc3(gpu, oracle):
    """Surface::setImage2D(Format_BC3n) decodes the BlockDXT5 bytes as BC3,
    with no swizzle (Surface.cpp:1058-1072): the normal's X comes back in
    alpha, its Y in green, red 255 and blue 0 for the device's own BC3n."""
    w, h = 64, 48
    planar = synth.to_planar(synth.normal_map_rgba(w, h, seed=4)).reshape(-1)
    d_src = gpu.alloc(planar.nbytes).upload(planar)
    d_blk = gpu.alloc(N.compressed_size(N.FORMAT_BC3N, w, h))
    gpu.compress_device(N.FORMAT_BC3N, d_src, w, h, d_blk)
    raw = d_blk.download(np.uint8)
    d_dec = gpu.alloc(16 * w * h)
    gpu.decode_device(N.FORMAT_BC3N, d_blk, w, h, d_dec)
    dec = d_dec.download(np.float32)
    assert np.array_equal(dec, oracle.decode_image(N.FORMAT_BC3, raw, w, h))
    n = w * h
    assert np.all(dec[:n] == 1.0) and np.all(dec[2 * n:3 * n] == 0.0)
    for got, src in ((dec[3 * n:], planar[:n]), (dec[n:2 * n], planar[n:2 * n])):
        err = np.abs(got - src)
        assert np.mean(err) < 0.02 and np.max(err) < 0.15, (np.mean(err), np.max(err))
    d_src.free(); d_blk.free(); d_dec.free()
