"""Pins the BC3/BC4/BC5 parts of the oracle: squish colour against the
reference's own nvsquish build and its golden vectors; QuickCompress and the
8-bit block compressors against golden vectors written by the reference's
own compiled code (tests/golden/bc4_channels.npz, bc8bit_blocks.npz; the live
comparisons are in test_oracle_vs_reference.py), plus closed forms and
invariants."""
import os

import numpy as np

import blocks

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_bc3_color_golden_vectors(oracle):
    g = np.load(os.path.join(GOLD, "bc3_color_blocks.npz"))
    for i in range(len(g["expected"])):
        out = oracle.bc3_color_block(g["bgra"][i], g["weight_by_alpha"][i])
        assert np.array_equal(out, g["expected"][i]), i


def test_bc4_channel_golden_vectors(oracle):
    """QuickCompress::compressDXT5A as the reference's compiled code wrote it."""
    g = np.load(os.path.join(GOLD, "bc4_channels.npz"))
    for i in range(len(g["channels"])):
        assert np.array_equal(oracle.bc4_channel(g["channels"][i]), g["quick"][i]), i


def test_8bit_block_golden_vectors(oracle):
    """CompressorDXT5 / CompressorDXT5n / CompressorDXT1a at Normal and
    FastCompressorBC4 / BC5 (compressBlock of the reference's compiled code)
    against the oracle's block functions."""
    g = np.load(os.path.join(GOLD, "bc8bit_blocks.npz"))
    px, am = g["bgra"], g["alpha_mode"]
    ones = (1, 1, 1)
    for i in range(len(px)):
        a = int(am[i])
        out16 = np.zeros(16, np.uint8)
        oracle.lib.ntxo_bc3_block(px[i].ctypes.data_as(__import__("ctypes").c_void_p), a,
                                  np.ones(3, np.float32).ctypes.data_as(__import__("ctypes").c_void_p),
                                  out16.ctypes.data_as(__import__("ctypes").c_void_p))
        assert np.array_equal(out16, g["bc3_normal"][i]), ("bc3", i)
        assert np.array_equal(oracle.bc3n_block(px[i], a), g["bc3n_normal"][i]), ("bc3n", i)
        assert np.array_equal(oracle.bc1a_block(px[i], a, ones), g["bc1a_normal"][i]), ("bc1a", i)
        assert np.array_equal(oracle.bc4_channel(px[i, :, 2]), g["bc4_normal"][i]), ("bc4", i)
        assert np.array_equal(np.concatenate([oracle.bc4_channel(px[i, :, 2]), oracle.bc4_channel(px[i, :, 1])]),
                              g["bc5_normal"][i]), ("bc5", i)


def test_bc3_color_against_compiled_squish(oracle, ref):
    px = blocks.bgra_set(seed=77, count=3000)
    for i in range(len(px)):
        if np.all(px[i, :, :3] == px[i, 0, :3]):
            continue  # single colour goes to OptimalCompress, not squish
        for wba in (0, 1):
            for cw in ((1, 1, 1), (0.3, 0.59, 0.11)):
                assert np.array_equal(oracle.bc3_color_block(px[i], wba, cw), ref.bc3_color_squish(px[i], wba, cw)), i


def test_omatch_tables_brute_force(oracle):
    """SingleColorLookup.cpp:34-89: each entry minimises
    100*|(2*max+min)/3 - v| + 3*|max-min| over expanded endpoints, first in loop order."""
    m5, m6 = oracle.omatch_tables()
    for bits, tab in ((5, m5), (6, m6)):
        n = 1 << bits
        ex = np.array([(i << 3) | (i >> 2) for i in range(n)] if bits == 5 else [(i << 2) | (i >> 4) for i in range(n)])
        mn, mx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        lerp = (2 * ex[mx] + ex[mn]) // 3
        for v in range(256):
            err = np.abs(lerp - v) * 100 + np.abs(mx - mn) * 3
            k = int(np.argmin(err.reshape(-1)))
            assert (tab[v, 0], tab[v, 1]) == (mx.reshape(-1)[k], mn.reshape(-1)[k])


def test_single_colour_bc3_decodes_close(oracle):
    """OptimalCompress::compressDXT1: index 2 of a single-colour block
    interpolates to within the table's error of every channel."""
    for v in range(0, 256, 3):
        px = np.zeros((16, 4), np.uint8)
        px[:, 0], px[:, 1], px[:, 2], px[:, 3] = v, 255 - v, (v * 7) % 256, 255
        out = oracle.bc3_color_block(px)
        dec = oracle.bc1_decode(out)
        assert np.all(np.abs(dec[:, :3].astype(int) - px[0, [2, 1, 0]].astype(int)) <= 4)


def test_bc4_constant_block_closed_form(oracle):
    """All texels v: inset endpoints are (v, v), every index 0, error 0; the
    least-squares step cannot improve on it (QuickCompressDXT.cpp:772-821)."""
    for v in range(256):
        out = oracle.bc4_channel(np.full(16, v, np.uint8))
        assert list(out) == [v, v, 0, 0, 0, 0, 0, 0]


def test_bc4_two_level_closed_form(oracle):
    """Values {lo, hi} with hi-lo < 34: the inset leaves (hi, lo), the 8-entry
    palette holds both exactly, and the block is lossless."""
    rng = np.random.default_rng(1)
    for _ in range(500):
        lo = int(rng.integers(0, 222)); hi = lo + int(rng.integers(1, 34))
        v = np.where(rng.random(16) < 0.5, lo, hi).astype(np.uint8)
        v[0], v[1] = lo, hi
        out = oracle.bc4_channel(v)
        assert (out[0], out[1]) == (hi, lo)
        assert np.array_equal(oracle.bc4_decode(out), v)


def test_bc4_error_never_worse_than_inset_start(oracle):
    """The refinement keeps the best block seen, so the result's squared error
    is at most that of the initial inset min/max block."""
    rng = np.random.default_rng(2)
    for _ in range(2000):
        v = rng.integers(0, 256, 16).astype(np.uint8)
        if _ % 3 == 0:
            v = (v // 40 * 40).astype(np.uint8)
        mx, mn = int(v.max()), int(v.min())
        a0, a1 = mx - (mx - mn) // 34, mn + (mx - mn) // 34
        if a0 > a1:
            pal = [a0, a1] + [((7 - k) * a0 + k * a1) // 7 for k in range(1, 7)]
        else:
            pal = [a0, a1] + [((5 - k) * a0 + k * a1) // 5 for k in range(1, 5)] + [0, 255]
        start = int(sum(min((p - int(x)) ** 2 for p in pal) for x in v))
        dec = oracle.bc4_decode(oracle.bc4_channel(v)).astype(int)
        assert int(((dec - v.astype(int)) ** 2).sum()) <= start


def test_bc4_indices_are_nearest(oracle):
    rng = np.random.default_rng(3)
    for _ in range(2000):
        v = rng.integers(0, 256, 16).astype(np.uint8)
        out = oracle.bc4_channel(v)
        a0, a1 = int(out[0]), int(out[1])
        if a0 > a1:
            pal = [a0, a1] + [((7 - k) * a0 + k * a1) // 7 for k in range(1, 7)]
        else:
            pal = [a0, a1] + [((5 - k) * a0 + k * a1) // 5 for k in range(1, 5)] + [0, 255]
        dec = oracle.bc4_decode(out).astype(int)
        for t in range(16):
            assert (dec[t] - int(v[t])) ** 2 == min((p - int(v[t])) ** 2 for p in pal)


def test_gather_truncates_and_repeats_edges(oracle):
    """ColorBlock::init (ColorBlock.cpp:80-110): uint8(255*clamp(x)) truncation,
    and pixel repetition (index modulo the valid extent) at image edges."""
    w, h = 3, 2
    planar = np.zeros((4, h, w), np.float32)
    planar[0] = [[0.0, 0.5, 1.0], [-1.0, 2.0, 0.999]]
    planar[1] = 0.25
    planar[2] = np.float32(1) / np.float32(255)
    planar[3] = 1.0
    bgra = np.zeros(64, np.uint8)
    oracle.lib.ntxo_gather_u8(w, h, planar.ctypes.data_as(__import__("ctypes").c_void_p), 0, 0,
                              bgra.ctypes.data_as(__import__("ctypes").c_void_p))
    px = bgra.reshape(4, 4, 4)  # [row][col][bgra]
    r = px[:, :, 2]
    assert list(r[0]) == [0, 127, 255, 0]       # x = 3 repeats x = 0 (3 % 3)
    assert list(r[1]) == [0, 255, 254, 0]
    assert np.array_equal(r[2], r[0]) and np.array_equal(r[3], r[1])  # y % 2
    assert np.all(px[:, :, 1] == 63)
    assert np.all(px[:, :, 0] == int(np.float32(255) * (np.float32(1) / np.float32(255))))


def _bc3n_tile(px):
    """swizzle(4, 1, 5, 3) of BGRA texels: red 0xFF, green kept, blue 0, alpha kept."""
    t = np.array(px, np.uint8).reshape(16, 4).copy()
    t[:, 0] = 0
    t[:, 2] = 0xFF
    return t


def test_bc3n_single_green_is_the_single_colour_path(oracle):
    """compressDXT1G(g) (OptimalCompressDXT.cpp:294-311: red 31, blue 0, green
    from OMatch6) is compressDXT1 of (255, g, 0), since the brute-force OMatch5
    maps 255 to (31, 31) and 0 to (0, 0): the device's BC3n front pass relies
    on this.  Red and blue of the input do not matter."""
    m5, _ = oracle.omatch_tables()
    assert tuple(m5[255]) == (31, 31) and tuple(m5[0]) == (0, 0)
    rng = np.random.default_rng(5)
    for g in range(256):
        px = rng.integers(0, 256, (16, 4), dtype=np.uint8)
        px[:, 1] = g
        got = oracle.bc3n_block(px)
        assert np.array_equal(got[8:], oracle.bc3_color_block(_bc3n_tile(px))), g


def test_bc3n_against_compiled_squish(oracle, ref):
    """CompressorDXT5n at Normal (CompressorDX9.cpp:179-222): colour half = the
    reference's own nvsquish on the swizzled texels with metric (0, 1, 0);
    alpha half = the QuickCompress alpha block of the red channel."""
    px = blocks.bgra_set(seed=78, count=2000)
    checked = 0
    for i in range(len(px)):
        if np.all(px[i, :, 1] == px[i, 0, 1]):
            continue  # single green: compressDXT1G, checked above
        for wba in (0, 1):
            got = oracle.bc3n_block(px[i], wba)
            assert np.array_equal(got[8:], ref.bc3_color_squish(_bc3n_tile(px[i]), wba, (0, 1, 0))), i
            assert np.array_equal(got[:8], oracle.bc4_channel(px[i, :, 2])), i
        checked += 1
    assert checked > 1500


def test_bc1a_squish_against_compiled_squish(oracle, ref):
    """The squish half of BC1a (CompressorDXT1a, kDxt1): alpha-0 texels leave
    the colour set and take index 3, the three-colour fit, then the
    four-colour fit for blocks without them.  Blocks with no, some and only
    transparent texels, both alpha modes, two metrics; the reference's own
    nvsquish decides every byte."""
    px = blocks.bgra_set(seed=79, count=1800)
    rng = np.random.default_rng(80)
    kinds = 0
    for i in range(len(px)):
        b = px[i].copy()
        mode = i % 3
        if mode == 1:
            b[rng.random(16) < 0.3, 3] = 0
        elif mode == 2:
            b[:, 3] = 0
        kinds += int(np.any(b[:, 3] == 0))
        for wba in (0, 1):
            for cw in ((1, 1, 1), (0.3, 0.59, 0.11)):
                assert np.array_equal(oracle.bc1a_squish(b, wba, cw), ref.bc1a_squish(b, wba, cw)), (i, wba, cw)
    assert kinds > 1000


def test_omatch_alpha_tables_brute_force(oracle):
    """SingleColorLookup.cpp:34-89 with alpha_mode: lerp (max + min) / 2."""
    m5, m6 = oracle.omatch_alpha_tables()
    for bits, tab in ((5, m5), (6, m6)):
        n = 1 << bits
        ex = np.array([(i << 3) | (i >> 2) for i in range(n)] if bits == 5 else [(i << 2) | (i >> 4) for i in range(n)])
        mn, mx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        lerp = (ex[mx] + ex[mn]) // 2
        for v in range(256):
            err = np.abs(lerp - v) * 100 + np.abs(mx - mn) * 3
            k = int(np.argmin(err.reshape(-1)))
            assert (tab[v, 0], tab[v, 1]) == (mx.reshape(-1)[k], mn.reshape(-1)[k])


def test_bc1a_block_paths(oracle, ref):
    """CompressorDXT1a: a block of one RGB with alpha-0 texels decodes to that
    colour's three-colour midpoint with those texels transparent (index 3,
    c0 <= c1); without them it is the four-colour single-colour block; any
    other block is the kDxt1 squish fit."""
    rng = np.random.default_rng(81)
    for _ in range(300):
        c = rng.integers(0, 256, 3)
        b = np.zeros((16, 4), np.uint8)
        b[:, :3] = c
        b[:, 3] = rng.integers(0, 256, 16)
        b[rng.random(16) < 0.4, 3] = 0
        out = oracle.bc1a_block(b)
        c0, c1 = int(out[0]) | int(out[1]) << 8, int(out[2]) | int(out[3]) << 8
        idx = int(out[4]) | int(out[5]) << 8 | int(out[6]) << 16 | int(out[7]) << 24
        if np.any(b[:, 3] == 0):
            assert c0 <= c1
            for i in range(16):
                assert (idx >> (2 * i)) & 3 == (3 if b[i, 3] == 0 else 2)
        dec = oracle.bc1_decode(out).reshape(16, 4)
        opaque = b[:, 3] != 0
        if opaque.any():
            assert np.all(np.abs(dec[opaque, :3].astype(int) - c[::-1].astype(int)) <= 4)
    px = blocks.bgra_set(seed=82, count=300)
    for i in range(len(px)):
        if np.all(px[i, :, :3] == px[i, 0, :3]):
            continue
        assert np.array_equal(oracle.bc1a_block(px[i]), ref.bc1a_squish(px[i], 0, (1, 1, 1))), i
