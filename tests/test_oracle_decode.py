"""The oracle's block decoders (Surface::setImage2D's nvimage decoders) and
error metrics (ErrorMetric.cpp) pinned against the reference's compiled icbc
decoder (BC1, D3D10 palette), the DXT5 alpha-palette definition at values
where the D3D9 bias changes the result, encoder round trips, and a float64
evaluation of the metrics."""
import numpy as np
import pytest

import blocks

BC1, BC3, BC4, BC5 = 1, 4, 6, 7


def test_bc1_decode_matches_reference_icbc(oracle, ref):
    rng = np.random.default_rng(4)
    raw = rng.integers(0, 256, (2000, 8), dtype=np.uint8)
    raw[::3, 0:2], raw[::3, 2:4] = raw[::3, 2:4].copy(), raw[::3, 0:2].copy()  # both endpoint orders
    for b in raw:
        got = oracle.decode_image(BC1, b, 4, 4).reshape(4, 16)
        exp = ref.bc1_decode(b).reshape(16, 4).T.astype(np.float32) / np.float32(255)
        assert np.array_equal(got[:3], exp[:3])
        # alpha as nv::BlockDXT1::evaluatePalette sets it (BlockDXT.cpp:43-98):
        # 255 but for index 3 of a three-colour block (icbc's own decoder
        # leaves the endpoints' alpha at 0, so only RGB is compared above)
        c0, c1 = int(b[0]) | int(b[1]) << 8, int(b[2]) | int(b[3]) << 8
        idx = (int.from_bytes(b[4:8].tobytes(), "little") >> (2 * np.arange(16))) & 3
        assert np.array_equal(got[3], np.where((c0 <= c1) & (idx == 3), 0.0, 1.0).astype(np.float32))


def _alpha_block(a0, a1, idx):
    u = a0 | (a1 << 8)
    for t, i in enumerate(idx):
        u |= int(i) << (16 + 3 * t)
    return np.frombuffer(int(u).to_bytes(8, "little"), np.uint8)


@pytest.mark.parametrize("d3d9", [False, True])
def test_alpha_palettes(oracle, d3d9):
    """DXT5 alpha palettes: 8 values when a0 > a1 ((7-k) a0 + k a1) / 7, else
    6 values ((5-k) a0 + k a1) / 5 plus 0 and 255; the D3D9 decoder adds 3
    (resp. 2) before dividing."""
    idx = list(range(8)) * 2
    exp8 = [255, 0, 218, 182, 145, 109, 72, 36] if not d3d9 else [255, 0, 219, 182, 146, 109, 73, 36]
    v8 = oracle.decode_image(BC4, _alpha_block(255, 0, idx), 4, 4, d3d9=d3d9)[:8] * 255
    assert np.allclose(v8, exp8), (v8, exp8)
    v6 = oracle.decode_image(BC4, _alpha_block(0, 254, idx), 4, 4, d3d9=d3d9)[:8] * 255
    exp6 = [0, 254, 50, 101, 152, 203, 0, 255] if not d3d9 else [0, 254, 51, 102, 152, 203, 0, 255]
    assert np.allclose(v6, exp6), (v6, exp6)


def test_bc4_bc5_round_trip(oracle):
    """The QuickCompress encoder's blocks decode to within its palette error;
    single-valued channels come back exactly; BC5 puts x in R, y in G, B = 0."""
    px = blocks.bgra_set(seed=12, count=300)
    for b in px:
        v = b[:, 2].copy()
        enc = oracle.bc4_channel(v)
        dec = np.round(oracle.decode_image(BC4, enc, 4, 4).reshape(4, 16) * 255).astype(int)
        assert np.array_equal(dec[0], dec[1]) and np.array_equal(dec[1], dec[2]) and np.all(dec[3] == 255)
        if len(set(v.tolist())) == 1:
            assert np.array_equal(dec[0], v)
        assert np.max(np.abs(dec[0] - v.astype(int))) <= 36
    v = np.full(16, 200, np.uint8)
    w = np.full(16, 30, np.uint8)
    enc = np.concatenate([oracle.bc4_channel(v), oracle.bc4_channel(w)])
    dec = np.round(oracle.decode_image(BC5, enc, 4, 4).reshape(4, 16) * 255).astype(int)
    assert np.array_equal(dec[0], v) and np.array_equal(dec[1], w) and np.all(dec[2] == 0) and np.all(dec[3] == 255)


def test_bc3_decode_composes_alpha_and_colour(oracle):
    rng = np.random.default_rng(8)
    for _ in range(200):
        b = rng.integers(0, 256, 16, dtype=np.uint8)
        got = oracle.decode_image(BC3, b, 4, 4).reshape(4, 16)
        col = oracle.decode_image(BC1, b[8:], 4, 4).reshape(4, 16)
        alp = oracle.decode_image(BC4, b[:8], 4, 4).reshape(4, 16)
        assert np.array_equal(got[:3], col[:3]) and np.array_equal(got[3], alp[0])


def test_partial_blocks_write_inside_texels_only(oracle):
    rng = np.random.default_rng(9)
    b = rng.integers(0, 256, 8 * 4, dtype=np.uint8)  # 2 x 2 blocks
    full = oracle.decode_image(BC1, b, 8, 8).reshape(4, 8, 8)
    # the same blocks for a 5 x 6 image: block raster is 2 x 2 as well
    part = oracle.decode_image(BC1, b, 5, 6).reshape(4, 6, 5)
    assert np.array_equal(part, full[:, :6, :5])


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_rms_metrics(oracle, metric):
    rng = np.random.default_rng(10 + metric)
    n = 3001
    a = rng.random(4 * n, dtype=np.float32)
    b = (a + rng.normal(0, 0.02, 4 * n)).astype(np.float32)
    A, B = a.reshape(4, n), b.reshape(4, n)
    if metric == 2:
        terms = ((B[3] - A[3]) ** 2).astype(np.float64)
    else:
        d = (A[:3] - B[:3])
        wgt = (A[3] * A[3]) if metric == 1 else np.float32(1)
        terms = ((d * d) * wgt).astype(np.float64)
    exp = np.float32(np.sqrt(terms.sum() / n))
    assert oracle.rms_error(metric, a, b) == pytest.approx(float(exp), rel=1e-6)
    assert oracle.rms_error(metric, a, a) == 0.0
