"""Pins the BC1 part of the oracle (oracle/ntx_oracle.c, restating icbc.h
Level8) against the golden vectors written by the reference's own icbc, the
compiled reference itself (when present), and closed forms."""
import os
import re
from math import comb

import numpy as np
import pytest

import blocks

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF_ICBC = "/root/reference/src/nvtt/icbc.h"


def test_bc1_golden_vectors(oracle):
    g = np.load(os.path.join(GOLD, "bc1_blocks.npz"))
    bad = []
    for i in range(len(g["expected"])):
        out, err, _, _ = oracle.bc1_block(g["colors"][i], g["weights"][i], g["color_weights"][i])
        if not np.array_equal(out, g["expected"][i]) or np.float32(err) != g["error"][i]:
            bad.append(i)
    assert not bad, f"{len(bad)} golden BC1 blocks differ, first {bad[:5]}"


def test_bc1_against_compiled_reference(oracle, ref):
    cols, wts, _ = blocks.bc1_set(seed=99, count=3000)
    for i in range(len(cols)):
        cw = (1, 1, 1) if i % 5 else (0.3, 0.59, 0.11)
        a, ea, _, _ = oracle.bc1_block(cols[i], wts[i], cw)
        b, eb = ref.bc1_block(cols[i], wts[i], cw)
        assert np.array_equal(a, b) and ea == eb, (i, a, b)


@pytest.mark.parametrize("name,quality", [("bc1_level9_blocks.npz", 2), ("bc1_level1_blocks.npz", 0)])
def test_bc1_quality_golden_vectors(oracle, name, quality):
    """Quality_Production = icbc Quality_Level9 (cluster fit + refine_endpoints),
    Quality_Fastest = Quality_Level1 (box fit + least squares): the golden
    vectors written by the reference's icbc."""
    g = np.load(os.path.join(GOLD, name))
    bad = []
    for i in range(len(g["expected"])):
        out, err = oracle.bc1_block_q(g["colors"][i], g["weights"][i], quality, g["color_weights"][i])
        if not np.array_equal(out, g["expected"][i]) or np.float32(err) != g["error"][i]:
            bad.append(i)
    assert not bad, f"{len(bad)} golden blocks of {name} differ, first {bad[:5]}"


@pytest.mark.parametrize("quality", [0, 2, 3])
def test_bc1_qualities_against_compiled_reference(oracle, ref, quality):
    cols, wts, _ = blocks.bc1_set(seed=77 + quality, count=1500)
    for i in range(len(cols)):
        cw = (1, 1, 1) if i % 4 else (0.3, 0.59, 0.11)
        a, ea = oracle.bc1_block_q(cols[i], wts[i], quality, cw)
        b, eb = ref.bc1_block_q(cols[i], wts[i], quality, cw)
        assert np.array_equal(a, b) and ea == eb, (i, a, b)


def test_bc1_level1_is_a_box_fit(oracle):
    """Quality_Fastest: always a four-colour block (c0 > c1 unless both
    endpoints quantise to one colour), whose error is the decoded MSE, and
    never better than the cluster fit by more than float noise on average."""
    cols, wts, _ = blocks.bc1_set(seed=8, count=800)
    worse = 0
    for i in range(len(cols)):
        b1, e1 = oracle.bc1_block_q(cols[i], wts[i], 0)
        b8, e8 = oracle.bc1_block_q(cols[i], wts[i], 1)
        c0, c1 = int(b1[0]) | int(b1[1]) << 8, int(b1[2]) | int(b1[3]) << 8
        assert c0 >= c1
        dec = oracle.bc1_decode(b1).reshape(16, 4)[:, :3].astype(np.float64)
        src = np.asarray(cols[i], np.float32).reshape(16, 4)[:, :3].astype(np.float64) * 255.0
        mse = float(np.sum(np.asarray(wts[i], np.float64)[:, None] * (dec - src) ** 2))
        assert mse == pytest.approx(e1, rel=1e-4, abs=1e-3)
        worse += int(e1 > e8 + 1e-3)
    assert worse > 100  # the box fit is the cheap, worse method


def test_bc1_level9_refines_level8(oracle):
    """refine_endpoints only accepts a lower evaluate_mse, so Level9's error never
    exceeds Level8's, it returns the decoded MSE of its block, and it moves
    a good share of the blocks of a real texture (a dropped step or a wrong
    index rule would change these)."""
    cols, wts, _ = blocks.bc1_set(seed=5, count=1500)
    moved = 0
    for i in range(len(cols)):
        b8, e8, _, _ = oracle.bc1_block(cols[i], wts[i])
        b9, e9 = oracle.bc1_block_q(cols[i], wts[i], 2)
        assert e9 <= e8
        moved += int(not np.array_equal(b8, b9))
        dec = oracle.bc1_decode(b9).reshape(16, 4)[:, :3].astype(np.float64)
        src = np.asarray(cols[i], np.float32).reshape(16, 4)[:, :3].astype(np.float64) * 255.0
        mse = float(np.sum(np.asarray(wts[i], np.float64)[:, None] * (dec - src) ** 2))
        assert mse == pytest.approx(e9, rel=1e-4, abs=1e-3)
    assert moved > 100


def _reference_table_procedure(n_levels, three):
    """The construction loop of init_cluster_tables (icbc.h:1906-1975): at each
    level, add every cumulative partition not present in the previous levels."""
    table, totals = [], []
    for t in range(1, n_levels + 1):
        prev = set(table[: totals[-1]]) if totals else set()
        if three:
            for c0 in range(t + 1):
                for c1 in range(t - c0 + 1):
                    if c0 == 0 and c1 == 0:
                        continue
                    e = (c0, c0 + c1)
                    if t > 1 and e in prev:
                        continue
                    table.append(e)
        else:
            for c0 in range(t + 1):
                for c1 in range(t - c0 + 1):
                    for c2 in range(t - c0 - c1 + 1):
                        if c0 == 0 and c1 == 0 and c2 == 0:
                            continue
                        e = (c0, c0 + c1, c0 + c1 + c2)
                        if t > 1 and e in prev:
                            continue
                        table.append(e)
        totals.append(len(table))
    return table, totals


def test_cluster_tables(oracle):
    t4, tot4 = oracle.cluster_table4()
    t3, tot3 = oracle.cluster_table3()
    r4, rt4 = _reference_table_procedure(16, three=False)
    r3, rt3 = _reference_table_procedure(16, three=True)
    assert list(tot4) == rt4 and list(tot3) == rt3
    assert [tuple(x) for x in t4[:968]] == r4
    assert [tuple(x) for x in t3[:152]] == r3
    # closed forms: N4(n) = C(n+3,3) - 1, N3(n) = C(n+2,2) - 1
    assert all(tot4[n - 1] == comb(n + 3, 3) - 1 for n in range(1, 17))
    assert all(tot3[n - 1] == comb(n + 2, 2) - 1 for n in range(1, 17))
    assert all(tuple(t4[968 + k]) == tuple(t4[967]) for k in range(8))


def test_midpoints(oracle):
    m5, m6 = oracle.midpoints()
    e5 = [((i << 3) | (i >> 2)) / 255 for i in range(32)]
    e6 = [((i << 2) | (i >> 4)) / 255 for i in range(64)]
    for i in range(31):
        assert abs(m5[i] - (e5[i] + e5[i + 1]) / 2) < 6e-7
    for i in range(63):
        assert abs(m6[i] - (e6[i] + e6[i + 1]) / 2) < 6e-7
    assert m5[31] == np.finfo(np.float32).max and m6[63] == np.finfo(np.float32).max
    if os.path.exists(REF_ICBC):
        src = open(REF_ICBC).read()
        for name, arr in (("midpoints5", m5), ("midpoints6", m6)):
            body = re.search(name + r"\[\d+\] = \{(.*?)\};", src, re.S).group(1)
            vals = [v.strip().rstrip("f") for v in body.split(",") if v.strip()]
            tab = np.array([np.finfo(np.float32).max if v == "FLT_MAX" else np.float32(float(v)) for v in vals], np.float32)
            assert np.array_equal(tab, arr), name


def test_single_colour_blocks_match_reference(oracle, ref):
    # Every grey level and a sample of colours through icbc's single-colour path.
    rng = np.random.default_rng(5)
    vals = [np.full(3, k, np.int64) for k in range(256)] + [rng.integers(0, 256, 3) for _ in range(500)]
    for v in vals:
        c = np.zeros((16, 4), np.float32)
        c[:, :3] = v.astype(np.float32) / np.float32(255)
        w = np.ones(16, np.float32)
        a, _, n, _ = oracle.bc1_block(c, w)
        b, _ = ref.bc1_block(c, w)
        assert n == 1 and np.array_equal(a, b), v


def test_single_colour_tables_are_optimal(oracle):
    """icbc.h:3166-3262: each entry minimises max(|AMD - v|, |NV - v|) over all
    endpoint pairs, first in (min, max) loop order."""
    m5, m6 = oracle.match_tables()
    e5 = np.array([(i << 3) | (i >> 2) for i in range(32)])
    e6 = np.array([(i << 2) | (i >> 4) for i in range(64)])
    mn5, mx5 = np.meshgrid(np.arange(32), np.arange(32), indexing="ij")
    amd5 = (43 * e5[mx5] + 21 * e5[mn5] + 32) >> 6
    nv5 = ((2 * mx5 + mn5) * 22) // 8
    mn6, mx6 = np.meshgrid(np.arange(64), np.arange(64), indexing="ij")
    amd6 = (43 * e6[mx6] + 21 * e6[mn6] + 32) >> 6
    d = e6[mx6] - e6[mn6]
    nv6 = (256 * e6[mn6] + np.trunc(d / 4).astype(int) + 128 + d * 80) // 256
    for v in range(256):
        err5 = np.maximum(abs(amd5 - v), abs(nv5 - v))
        k = np.argmin(err5.reshape(-1))  # first minimum in (mn, mx) order
        assert (m5[v, 0], m5[v, 1]) == (mx5.reshape(-1)[k], mn5.reshape(-1)[k])
        err6 = np.maximum(abs(amd6 - v), abs(nv6 - v))
        k = np.argmin(err6.reshape(-1))
        assert (m6[v, 0], m6[v, 1]) == (mx6.reshape(-1)[k], mn6.reshape(-1)[k])


def test_decode_closed_form(oracle):
    """BlockDXT.cpp:43-98 / icbc.h:2557-2585: 565 bit expansion and the D3D10
    palette, for every colour value and a sample of endpoint pairs."""
    rng = np.random.default_rng(11)
    for _ in range(4000):
        c0, c1 = (int(x) for x in rng.integers(0, 65536, 2))
        idx = int(rng.integers(0, 2**32))
        blk = np.frombuffer(np.array([c0, c1], "<u2").tobytes() + np.array([idx], "<u4").tobytes(), np.uint8)
        px = oracle.bc1_decode(blk)

        def ex(c):
            r, g, b = c >> 11, (c >> 5) & 63, c & 31
            return np.array([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)])
        p0, p1 = ex(c0), ex(c1)
        pal = [p0, p1, (2 * p0 + p1) // 3, (p0 + 2 * p1) // 3] if c0 > c1 else [p0, p1, (p0 + p1) // 2, np.zeros(3, int)]
        for t in range(16):
            k = (idx >> (2 * t)) & 3
            assert np.array_equal(px[t, :3], pal[k])
            assert px[t, 3] == (0 if (c0 <= c1 and k == 3) else 255)


def test_error_equals_decoded_mse(oracle):
    """icbc returns the weighted MSE of the block it writes (icbc.h:2751-2760):
    decode the output and recompute it in float64."""
    cols, wts, _ = blocks.bc1_set(seed=3, count=1500)
    for i in range(len(cols)):
        if not np.all((cols[i] >= 0) & (cols[i] <= 1)):
            continue
        out, err, n, _ = oracle.bc1_block(cols[i], wts[i])
        if n == 0:
            assert err == 0 and not out.any()
            continue
        dec = oracle.bc1_decode(out)[:, :3].astype(np.float64) / 255
        mse = float(np.sum(wts[i] * np.sum(((dec - cols[i][:, :3]) * 255) ** 2, axis=1)))
        assert abs(mse - err) <= 1e-3 * max(1.0, mse), (i, mse, err)


def test_indices_pick_a_nearest_palette_entry(oracle):
    """Index selection (icbc.h:2803-2882) assigns every texel a palette entry
    at minimal distance (up to float rounding of the distances)."""
    cols, wts, _ = blocks.bc1_set(seed=4, count=1500)
    for i in range(len(cols)):
        out, _, n, _ = oracle.bc1_block(cols[i], wts[i])
        if n <= 1:
            continue
        c0 = int(out[0]) | int(out[1]) << 8
        c1 = int(out[2]) | int(out[3]) << 8
        dec = oracle.bc1_decode(out)[:, :3].astype(np.float64) / 255
        blk = out.copy()
        pal = []
        for k in range(4):
            blk[4:] = np.frombuffer(np.array([0x55555555 * k], "<u4").tobytes(), np.uint8)
            pal.append(oracle.bc1_decode(blk)[0, :3].astype(np.float64) / 255)
        pal = np.array(pal)
        if c0 > c1 or c0 == c1:
            pal = pal if c0 > c1 else pal[:3]
        d = ((cols[i][:, None, :3].astype(np.float64) - pal[None]) ** 2).sum(-1)
        chosen = ((dec[:, None, :] - cols[i][:, None, :3]) ** 2).sum(-1)[:, 0]
        assert np.all(chosen <= d.min(axis=1) + 1e-6), i
