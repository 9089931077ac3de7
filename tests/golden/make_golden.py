"""Writes the committed golden vectors of tests/golden/ from the reference's
own encoders, compiled from the reference sources by oracle/Makefile into
oracle/_ref/libntxref.so (icbc.h with ICBC_SIMD=2, nvsquish with
SQUISH_USE_SSE=0).  Run in the container that holds the reference tree:

    make -C oracle && python tests/golden/make_golden.py

bc1_blocks.npz        inputs (colors, weights, colour weights) and icbc's 8-byte
                      output + returned error for adversarial BC1 blocks.
bc3_color_blocks.npz  8-bit BGRA blocks and squish's 8-byte colour block.
chain_c1_64.npz       the 64x64 noise+rectangles texture (seed 7), and icbc's
                      BC1 bytes for every level of its Box chain; the levels
                      are built by the reference's FloatImage::fastDownSample.
bc1_level9_blocks.npz adversarial BC1 blocks (another seed) and icbc's output at
                      Quality_Level9 (nvtt's Quality_Production: the cluster
                      fit plus refine_endpoints) + returned error.
bc1_level1_blocks.npz the same for Quality_Level1 (nvtt's Quality_Fastest: box
                      fit + least squares).
bc4_channels.npz      4096 channel blocks (blocks.channel_set) and the reference's
                      QuickCompress::compressDXT5A bytes ("quick") and
                      OptimalCompress::compressDXT5A bytes with its first error
                      seed taken from (max, min) ("optimal").
bc8bit_blocks.npz     1024 BGRA blocks (blocks.bgra_set), their alpha mode, and
                      the bytes of every 8-bit block compressor Context.cpp picks
                      on the path: BC3 / BC3n at Fastest, Normal and Highest,
                      BC1a at Fastest and Normal, BC4 / BC5 at Normal and
                      Production (Production / Highest seeded as above).
mip_chains.npz        a seeded 37x23 random FP32 RGBA image and the CRC-32 of
                      the FP32 bytes of every level of its chain down to 1x1
                      through the reference's FloatImage for the Box (fast and
                      polyphase), Triangle and Kaiser filters x three wrap
                      modes x both alpha modes.

`python tests/golden/make_golden.py levels` writes only the two BC1 level
files, `python tests/golden/make_golden.py nv` only the last three.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "nvidia-texture-tools_amd"))

import zlib  # noqa: E402

import blocks  # noqa: E402
import oracle_lib  # noqa: E402
from ntx_amd import synth  # noqa: E402

META = "reference castano/nvidia-texture-tools 2.1.2; g++ 11.4 -O2 -std=c++11 -msse4.1 -ffp-contract=off -DICBC_SIMD=2 -DSQUISH_USE_SSE=0"


def gather_bc1(planar, w, h):
    """FloatColorCompressorTask (BlockCompressor.cpp:116-166), AlphaMode_None."""
    bw, bh = (w + 3) // 4, (h + 3) // 4
    p = planar.reshape(4, h, w)
    cols = np.zeros((bw * bh, 16, 4), np.float32)
    wts = np.zeros((bw * bh, 16), np.float32)
    for by in range(bh):
        for bx in range(bw):
            i = by * bw + bx
            for y in range(4):
                for x in range(4):
                    X, Y = 4 * bx + x, 4 * by + y
                    if X < w and Y < h:
                        cols[i, 4 * y + x] = p[:, Y, X]
                        wts[i, 4 * y + x] = 1.0
    return cols, wts


def quality_levels(ref):
    for name, quality, seed, level in (("bc1_level9_blocks.npz", 2, 2026, "Quality_Max"),
                                       ("bc1_level1_blocks.npz", 0, 2027, "Quality_Fast")):
        cols, wts, kinds = blocks.bc1_set(seed=seed, count=2048)
        cws = np.ones((len(cols), 3), np.float32)
        cws[2::9] = np.array([0.2126, 0.7152, 0.0722], np.float32)
        out = np.zeros((len(cols), 8), np.uint8)
        err = np.zeros(len(cols), np.float32)
        for i in range(len(cols)):
            out[i], err[i] = ref.bc1_block_q(cols[i], wts[i], quality, cws[i])
        path = os.path.join(HERE, name)
        np.savez_compressed(path, colors=cols, weights=wts, color_weights=cws, kinds=kinds, expected=out, error=err,
                            meta=META + f"; icbc::compress_dxt1({level})")
        print(name, os.path.getsize(path), "bytes")


MIP_FILTERS = ((0, 0.5, None), (0, 1.0, None), (1, 1.0, None), (2, 3.0, (4.0, 1.0)))


def mip_chain_crcs(build, planar, W, H):
    """CRC-32 of the FP32 bytes of every level from W x H down to 1 x 1."""
    crcs = [zlib.crc32(np.ascontiguousarray(planar, np.float32).tobytes())]
    while (W, H) != (1, 1):
        planar = build(planar, W, H)
        W, H = max(1, W // 2), max(1, H // 2)
        crcs.append(zlib.crc32(planar.tobytes()))
    return crcs


def nv_fixtures(ref):
    chans = blocks.channel_set(2031, 4096)
    quick = np.stack([ref.bc4_channel(c) for c in chans])
    optimal = np.stack([ref.bc4_channel_optimal(c) for c in chans])
    np.savez_compressed(os.path.join(HERE, "bc4_channels.npz"), channels=chans, quick=quick, optimal=optimal,
                        meta=META + "; QuickCompress / OptimalCompress::compressDXT5A (seed (max, min))")

    px = blocks.bgra_set(seed=2032, count=1024)
    am = (np.arange(len(px)) % 3 == 2).astype(np.int32)
    out = {}
    for name, fmt, q in (("bc3_fastest", 4, 0), ("bc3_normal", 4, 1), ("bc3_highest", 4, 3),
                         ("bc3n_fastest", 5, 0), ("bc3n_normal", 5, 1), ("bc3n_highest", 5, 3),
                         ("bc1a_fastest", 2, 0), ("bc1a_normal", 2, 1),
                         ("bc4_normal", 6, 1), ("bc4_production", 6, 2), ("bc5_normal", 7, 1), ("bc5_production", 7, 2)):
        out[name] = np.stack([ref.block(fmt, q, px[i], int(am[i])) for i in range(len(px))])
    np.savez_compressed(os.path.join(HERE, "bc8bit_blocks.npz"), bgra=px, alpha_mode=am, meta=META +
                        "; compressBlock of the compressor chooseCpuCompressor picks (Context.cpp:1038-1115)", **out)

    rng = np.random.default_rng(2033)
    W, H = 37, 23
    img = rng.random(4 * W * H).astype(np.float32)
    crcs = {}
    for filt, width, params in MIP_FILTERS:
        for wrap in range(3):
            for alpha_mode in (0, 1):
                key = f"f{filt}_w{width}_wrap{wrap}_a{alpha_mode}"
                crcs[key] = np.array(mip_chain_crcs(
                    lambda p, w, h: ref.build_next_mipmap(filt, width, params, wrap, p, w, h, alpha_mode), img, W, H),
                    np.uint32)
    np.savez_compressed(os.path.join(HERE, "mip_chains.npz"), image=img, width=W, height=H,
                        meta=META + "; Surface::buildNextMipmap steps through FloatImage (fastDownSample / downSample)",
                        **crcs)
    for f in ("bc4_channels.npz", "bc8bit_blocks.npz", "mip_chains.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


def main():
    ref = oracle_lib.Reference.load()
    if ref is None:
        raise SystemExit("oracle/_ref/libntxref.so missing: build it with `make -C oracle` next to the reference tree")
    orc = oracle_lib.Oracle()

    cols, wts, kinds = blocks.bc1_set(seed=2024, count=2048)
    cws = np.ones((len(cols), 3), np.float32)
    cws[1::7] = np.array([0.5, 1.0, 0.25], np.float32)
    cws[3::11] = np.array([0.2126, 0.7152, 0.0722], np.float32)
    out = np.zeros((len(cols), 8), np.uint8)
    err = np.zeros(len(cols), np.float32)
    for i in range(len(cols)):
        out[i], err[i] = ref.bc1_block(cols[i], wts[i], cws[i])
    np.savez_compressed(os.path.join(HERE, "bc1_blocks.npz"), colors=cols, weights=wts, color_weights=cws,
                        kinds=kinds, expected=out, error=err, meta=META)

    px = blocks.bgra_set(seed=2025, count=2048)
    wba = (np.arange(len(px)) % 3 == 2).astype(np.int32)
    out3 = np.zeros((len(px), 8), np.uint8)
    for i in range(len(px)):
        single = np.all(px[i, :, :3] == px[i, 0, :3])
        out3[i] = orc.bc3_color_block(px[i], wba[i]) if single else ref.bc3_color_squish(px[i], wba[i])
    np.savez_compressed(os.path.join(HERE, "bc3_color_blocks.npz"), bgra=px, weight_by_alpha=wba, expected=out3,
                        meta=META + "; single-colour blocks use the oracle's OptimalCompress path")

    img = synth.noise_rect_rgba(64, 64, seed=7)
    planar = synth.to_planar(img).reshape(-1)
    w = h = 64
    levels = {}
    while True:
        c, wt = gather_bc1(planar, w, h)
        b = np.stack([ref.bc1_block(c[i], wt[i])[0] for i in range(len(c))])
        levels[f"bc1_{w}x{h}"] = b
        if w == 1 and h == 1:
            break
        planar = ref.fast_downsample(planar, w, h)
        w, h = max(1, w // 2), max(1, h // 2)
    np.savez_compressed(os.path.join(HERE, "chain_c1_64.npz"), image=img, meta=META, **levels)
    for f in ("bc1_blocks.npz", "bc3_color_blocks.npz", "chain_c1_64.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    if sys.argv[1:] == ["nv"]:
        r = oracle_lib.Reference.load()
        if r is None:
            raise SystemExit("oracle/_ref/libntxref.so missing")
        nv_fixtures(r)
    elif sys.argv[1:] == ["levels"]:
        r = oracle_lib.Reference.load()
        if r is None:
            raise SystemExit("oracle/_ref/libntxref.so missing")
        quality_levels(r)
    else:
        main()
        quality_levels(oracle_lib.Reference.load())
        nv_fixtures(oracle_lib.Reference.load())
