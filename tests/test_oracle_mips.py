"""Pins the oracle's mip filters (FloatImage / Filter) by golden CRCs written by
the reference's own compiled FloatImage (the live bit-for-bit comparisons are
in test_oracle_vs_reference.py), closed forms, library routines (numpy i0 /
sinc) and exact-arithmetic inputs."""
import os

import numpy as np
import pytest

import ntx_amd as N


def _planes(rng, W, H, ints=False):
    if ints:
        return rng.integers(0, 64, 4 * W * H).astype(np.float32)
    return rng.random(4 * W * H, dtype=np.float32)


def test_box_even_is_exact_2x2_mean(oracle):
    rng = np.random.default_rng(0)
    W, H = 16, 10
    x = _planes(rng, W, H, ints=True)  # small integers: every sum is exact
    got = oracle.fast_downsample(x, W, H).reshape(4, H // 2, W // 2)
    p = x.reshape(4, H, W).astype(np.float64)
    exp = 0.25 * (p[:, 0::2, 0::2] + p[:, 0::2, 1::2] + p[:, 1::2, 0::2] + p[:, 1::2, 1::2])
    assert np.array_equal(got, exp.astype(np.float32))


@pytest.mark.parametrize("W,H", [(7, 5), (7, 4), (6, 5), (1, 9), (9, 1), (1, 8), (3, 3)])
def test_box_odd_sizes_polyphase_weights(oracle, W, H):
    """np2 mipmapping: odd extents use 3-tap weights (n-x, n, 1+x)/(2n+1)."""
    rng = np.random.default_rng(W * 10 + H)
    x = _planes(rng, W, H, ints=True)
    w, h = max(1, W // 2), max(1, H // 2)
    got = oracle.fast_downsample(x, W, H).reshape(4, h, w)

    def taps(n_src, n_dst):
        m = np.zeros((n_dst, n_src))
        if n_src == 1:
            m[0, 0] = 1
        elif n_src % 2 == 0:
            for i in range(n_dst):
                m[i, 2 * i] = m[i, 2 * i + 1] = 0.5
        else:
            for i in range(n_dst):
                m[i, 2 * i:2 * i + 3] = np.array([n_dst - i, n_dst, 1 + i]) / (2 * n_dst + 1)
        return m
    p = x.reshape(4, H, W).astype(np.float64)
    if W == 1 or H == 1:
        n_src = W * H
        n_dst = w * h
        flat = p.reshape(4, -1)
        exp = (taps(n_src, n_dst) @ flat.T).T.reshape(4, h, w)
    else:
        exp = np.einsum("yi,cij,xj->cyx", taps(H, h), p, taps(W, w))
    assert np.allclose(got, exp, rtol=2e-7, atol=1e-6)


def test_box_polyphase_weights_closed_form(oracle):
    """Box, width 0.5, halving: taps (1/2, 1/2, 0) on each output."""
    k = oracle.polyphase(N.FILTER_BOX, 0.5, None, 64, 32)
    assert k.shape == (32, 3)
    assert np.array_equal(k, np.tile(np.array([0.5, 0.5, 0.0], np.float32), (32, 1)))


def test_triangle_weights_closed_form(oracle):
    """Triangle, width 1, halving: the box-integrated tent gives
    (1/8, 3/8, 3/8, 1/8, 0) exactly (midpoint rule is exact on linear pieces)."""
    k = oracle.polyphase(N.FILTER_TRIANGLE, 1.0, None, 64, 32)
    assert k.shape == (32, 5)
    assert np.array_equal(k, np.tile(np.array([0.125, 0.375, 0.375, 0.125, 0.0], np.float32), (32, 1)))


@pytest.mark.parametrize("src,dst,width,alpha,stretch", [(64, 32, 3.0, 4.0, 1.0), (37, 18, 3.0, 4.0, 1.0), (9, 4, 2.0, 3.0, 0.8), (2, 1, 3.0, 4.0, 1.0)])
def test_kaiser_weights_match_library_routines(oracle, src, dst, width, alpha, stretch):
    """Filter.cpp:257-263: Kaiser-windowed sinc, sinc(pi x s) * I0(a sqrt(1-t^2)) / I0(a),
    box-integrated with 32 midpoint samples, rows normalised; computed here in
    float64 with numpy's sinc and i0."""
    k = oracle.polyphase(N.FILTER_KAISER, width, (alpha, stretch), src, dst)
    scale = dst / src
    iscale = 1 / scale
    kw = width * iscale
    ws = int(np.ceil(kw * 2)) + 1
    assert k.shape == (dst, ws)
    for i in range(dst):
        center = (0.5 + i) * iscale
        left = int(np.floor(center - kw))
        row = []
        for j in range(ws):
            x = left + j - center
            s = (x + (np.arange(32) + 0.5) / 32) * scale
            t = s / width
            val = np.where(1 - t * t >= 0, np.sinc(s * stretch) * np.i0(alpha * np.sqrt(np.clip(1 - t * t, 0, None))) / np.i0(alpha), 0)
            row.append(val.mean())
        row = np.array(row) / np.sum(row)
        assert np.allclose(k[i], row, atol=2e-6), i


@pytest.mark.parametrize("wrap,mode", [(N.WRAP_CLAMP, "edge"), (N.WRAP_REPEAT, "wrap"), (N.WRAP_MIRROR, "reflect")])
@pytest.mark.parametrize("filt", [N.FILTER_TRIANGLE, N.FILTER_KAISER])
def test_separable_resize_matches_padded_convolution(oracle, wrap, mode, filt):
    """FloatImage::resize: X pass then Y pass with the kernel rows; borders by
    the wrap mode (FloatImage.h:298-330), written here with numpy padding."""
    W, H = 24, 18
    rng = np.random.default_rng(4)
    x = _planes(rng, W, H)
    width, params = N.DEFAULT_FILTER[filt]
    got = oracle.downsample(filt, width, params, wrap, x, W, H).reshape(4, H // 2, W // 2)
    kx = oracle.polyphase(filt, width, params, W, W // 2).astype(np.float64)
    ky = oracle.polyphase(filt, width, params, H, H // 2).astype(np.float64)
    p = x.reshape(4, H, W).astype(np.float64)
    pad = 16
    pp = np.pad(p, ((0, 0), (pad, pad), (pad, pad)), mode=mode)

    def resample(arr, k, n_src, axis):
        n_dst, ws = k.shape
        iscale = n_src / n_dst
        kw = (width * iscale)
        out = []
        for i in range(n_dst):
            left = int(np.floor((0.5 + i) * iscale - kw))
            idx = np.arange(left, left + ws) + pad
            seg = np.take(arr, idx, axis=axis)
            out.append(np.tensordot(seg, k[i], axes=([axis], [0])))
        return np.stack(out, axis=axis)
    tmp = resample(pp, kx, W, axis=2)            # (4, H+2pad, w)
    exp = resample(tmp, ky, H, axis=1)           # (4, h, w)
    assert np.allclose(got, exp, atol=2e-6)


def test_constant_image_stays_constant(oracle):
    for filt in (N.FILTER_BOX, N.FILTER_TRIANGLE, N.FILTER_KAISER):
        width, params = N.DEFAULT_FILTER[filt]
        for wrap in (N.WRAP_CLAMP, N.WRAP_REPEAT, N.WRAP_MIRROR):
            x = np.full(4 * 20 * 12, 0.3, np.float32)
            y = oracle.build_next_mipmap(filt, width, params, wrap, x, 20, 12)
            assert np.allclose(y, 0.3, rtol=1e-6)


def test_srgb_to_linear_closed_form(oracle):
    v = np.linspace(-0.2, 1.2, 20001).astype(np.float32)
    got = oracle.srgb_to_linear(v).astype(np.float64)
    f = np.clip(v.astype(np.float64), 0, None)
    exp = np.where(f < 0.04045, f / 12.92, ((f + 0.055) / 1.055) ** 2.4)
    exp = np.where(v > 1, 1.0, exp)
    assert np.allclose(got, exp, rtol=3e-7, atol=1e-9)


def test_mip_chain_golden_crcs(oracle):
    """Every level of a 37x23 chain to 1x1, CRC-32 of the FP32 bytes as the
    reference's FloatImage wrote them (tests/golden/mip_chains.npz): Box (fast
    and polyphase), Triangle, Kaiser x three wrap modes x both alpha modes
    (AlphaMode_Transparency always takes the polyphase resize)."""
    import zlib
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "mip_chains.npz"))
    filters = {"f0_w0.5": (0, 0.5, None), "f0_w1.0": (0, 1.0, None), "f1_w1.0": (1, 1.0, None),
               "f2_w3.0": (2, 3.0, (4.0, 1.0))}
    seen = 0
    for key in g.files:
        if not key.startswith("f"):
            continue
        fw, wrap, am = key.rsplit("_", 2)
        filt, width, params = filters[fw]
        wrap, am = int(wrap[4:]), int(am[1:])
        p, W, H = g["image"], int(g["width"]), int(g["height"])
        crcs = [zlib.crc32(p.tobytes())]
        while (W, H) != (1, 1):
            if am:
                p = oracle.downsample(filt, width, params, wrap, p, W, H)
            else:
                p = oracle.build_next_mipmap(filt, width, params, wrap, p, W, H)
            W, H = max(1, W // 2), max(1, H // 2)
            crcs.append(zlib.crc32(p.tobytes()))
        assert np.array_equal(np.array(crcs, np.uint32), g[key]), key
        seen += 1
    assert seen == 24
