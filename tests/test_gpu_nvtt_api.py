"""The nvtt.h front end (lib/libnvtt_amd.so) driven from C++ exactly as an
NVTT user drives the reference: Surface::setImage -> [toLinearFromSrgb] ->
Compressor::compress per level with an OutputHandler -> buildNextMipmap.
The captured bytes must equal the oracle's for the same pipeline."""
import os
import struct
import subprocess

import numpy as np
import pytest

import ntx_amd as N
from ntx_amd import synth

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(N.LIB_DIR, "nvtt_api_test")


def run_chain(tmp_path, img, fmt, filt, srgb, alpha_mode, wrap):
    h, w = img.shape[:2]
    src = tmp_path / "in.raw"
    dst = tmp_path / "out.bin"
    with open(src, "wb") as f:
        f.write(struct.pack("<ii", w, h))
        f.write(synth.to_bgra(img).tobytes())
    r = subprocess.run([DRIVER, "chain", str(src), str(dst), str(fmt), str(filt), str(int(srgb)), str(alpha_mode), str(wrap)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = dst.read_bytes()
    off, levels = 0, []
    for _ in N.mip_sizes(w, h):
        lw, lh, size = struct.unpack_from("<iii", data, off)
        off += 12
        levels.append((lw, lh, np.frombuffer(data, np.uint8, size, off)))
        off += size
    (n,) = struct.unpack_from("<i", data, off)
    events = np.frombuffer(data, np.int32, 7 * n, off + 4).reshape(n, 7)
    return levels, events


def oracle_chain(oracle, img, fmt, filt, srgb, alpha_mode, wrap):
    h, w = img.shape[:2]
    x = synth.to_planar(img).reshape(-1)
    if srgb:
        x = np.concatenate([oracle.srgb_to_linear(x[: 3 * w * h]), x[3 * w * h:]])
    width, params = N.DEFAULT_FILTER[filt]
    out = []
    for k, (lw, lh) in enumerate(N.mip_sizes(w, h)):
        out.append(oracle.compress(fmt, x, lw, lh, alpha_mode=alpha_mode))
        if (lw, lh) != (1, 1):
            if alpha_mode == N.ALPHA_TRANSPARENCY:
                x = oracle.downsample(filt, width, params, wrap, x, lw, lh)  # Surface.cpp:1356-1375
            else:
                x = oracle.build_next_mipmap(filt, width, params, wrap, x, lw, lh)
    return out


CASES = [
    # (format, filter, srgb, alpha mode, wrap, size)
    (N.FORMAT_BC1, N.FILTER_BOX, False, N.ALPHA_NONE, N.WRAP_MIRROR, (128, 96)),
    (N.FORMAT_BC1, N.FILTER_KAISER, True, N.ALPHA_NONE, N.WRAP_MIRROR, (64, 64)),
    (N.FORMAT_BC1, N.FILTER_BOX, False, N.ALPHA_TRANSPARENCY, N.WRAP_CLAMP, (70, 33)),
    (N.FORMAT_BC3, N.FILTER_KAISER, True, N.ALPHA_NONE, N.WRAP_MIRROR, (128, 128)),
    (N.FORMAT_BC3, N.FILTER_TRIANGLE, False, N.ALPHA_TRANSPARENCY, N.WRAP_REPEAT, (37, 61)),
    (N.FORMAT_BC4, N.FILTER_TRIANGLE, False, N.ALPHA_NONE, N.WRAP_REPEAT, (96, 64)),
    (N.FORMAT_BC5, N.FILTER_KAISER, False, N.ALPHA_NONE, N.WRAP_CLAMP, (64, 130)),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"fmt{c[0]}-f{c[1]}-s{int(c[2])}-a{c[3]}-w{c[4]}-{c[5][0]}x{c[5][1]}")
def test_front_end_chain_matches_oracle(tmp_path, oracle, case):
    fmt, filt, srgb, alpha_mode, wrap, (w, h) = case
    if fmt == N.FORMAT_BC5:
        img = synth.normal_map_rgba(w, h, seed=5)
    elif fmt == N.FORMAT_BC4:
        img = synth.single_channel_rgba(w, h, seed=6)
    else:
        img = synth.noise_rect_rgba(w, h, seed=4, transparent=True)
    levels, events = run_chain(tmp_path, img, fmt, filt, srgb, alpha_mode, wrap)
    expect = oracle_chain(oracle, img, fmt, filt, srgb, alpha_mode, wrap)
    sizes = N.mip_sizes(w, h)
    assert [(lw, lh) for lw, lh, _ in levels] == sizes
    for (lw, lh, got), exp in zip(levels, expect):
        assert np.array_equal(got, exp), (lw, lh)
    # beginImage(size, w, h, 1, face 0, mip m), one writeData(size), endImage, per level
    assert len(events) == 3 * len(sizes)
    for m, (lw, lh) in enumerate(sizes):
        size = N.compressed_size(fmt, lw, lh)
        assert list(events[3 * m]) == [1, size, lw, lh, 1, 0, m]
        assert list(events[3 * m + 1][:2]) == [2, size]
        assert events[3 * m + 2][0] == 3


def test_front_end_api_semantics():
    r = subprocess.run([DRIVER, "api"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "api ok" in r.stdout, r.stderr


def test_front_end_batch_list():
    """Compressor::compress(BatchList): one call for several surfaces and
    their mip chains gives every image's bytes and handler events exactly as
    compress(Surface) per image does (BC1, BC4, BC5 through the persistent
    batch launch; BC3 through the chain scheduler)."""
    r = subprocess.run([DRIVER, "batch"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "batch ok" in r.stdout, r.stdout + r.stderr


def test_device_error_reported_and_next_image_compresses(oracle):
    """A device failure (here an allocation over the backend's scratch limit)
    reaches the ErrorHandler as Error_CudaError with no data written, and the
    same Compressor then compresses the next image, twice over
    (CudaCompressorDXT.cpp:183-189)."""
    env = dict(os.environ, NTX_SCRATCH_LIMIT_BYTES=str(1 << 20))
    r = subprocess.run([DRIVER, "device-error"], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0 and "device-error ok" in r.stdout, r.stdout + r.stderr
