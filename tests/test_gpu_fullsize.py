"""GPU parity at the configurations' full sizes: each workload as the
benchmark runs it, every level and block against the oracle or against the
reference's own code compiled from its sources (oracle/_ref, used when it
was built).  Bit-exact output is the bar."""
import numpy as np
import pytest

import ntx_amd as N
from ntx_amd import synth

pytestmark = pytest.mark.gpu


def _diff_blocks(a, b, bs):
    return int(np.sum(np.any(a.reshape(-1, bs) != b.reshape(-1, bs), axis=1)))


def test_c1_box_chain_bc1(gpu, oracle):
    """1024^2 noise+rectangles, 11-level Box chain, BC1 Normal per level."""
    S = 1024
    planar = synth.to_planar(synth.noise_rect_rgba(S, S, seed=1)).reshape(-1)
    sizes = N.mip_sizes(S, S)
    d_lv = [gpu.alloc(16 * w * h) for w, h in sizes]
    d_lv[0].upload(planar)
    d_out = [gpu.alloc(N.compressed_size(N.FORMAT_BC1, w, h)) for w, h in sizes]
    gpu.compress_chain_device(N.FORMAT_BC1, N.FILTER_BOX, d_lv, S, S, d_out)
    lv = planar
    for k, (w, h) in enumerate(sizes):
        got = d_lv[k].download(np.float32)
        assert np.array_equal(got.view(np.uint32), lv.view(np.uint32)), (k, "mip level")
        exp = oracle.compress(N.FORMAT_BC1, lv, w, h)
        assert _diff_blocks(d_out[k].download(np.uint8), exp, 8) == 0, (k, "blocks")
        if k + 1 < len(sizes):
            lv = oracle.build_next_mipmap(N.FILTER_BOX, 0.5, None, N.WRAP_MIRROR, lv, w, h)
    for d in d_lv + d_out:
        d.free()


def test_c2_srgb_kaiser_bc3_chain(gpu, oracle):
    """4096^2 with 30% transparent alpha: BGRA8 -> sRGB->linear -> 13-level
    Kaiser(3; 4, 1) chain -> BC3 per level, through the chain entry point."""
    S = 4096
    img = synth.noise_rect_rgba(S, S, seed=1, transparent=True)
    bgra = synth.to_bgra(img)
    sizes = N.mip_sizes(S, S)
    d_bgra = gpu.alloc(bgra.nbytes).upload(bgra)
    d_lv = [gpu.alloc(16 * w * h) for w, h in sizes]
    d_out = [gpu.alloc(N.compressed_size(N.FORMAT_BC3, w, h)) for w, h in sizes]
    gpu.unpack_bgra8_srgb_device(d_bgra, S, S, d_lv[0])
    gpu.compress_chain_device(N.FORMAT_BC3, N.FILTER_KAISER, d_lv, S, S, d_out, width=3.0, params=(4.0, 1.0))
    x = synth.to_planar(img).reshape(-1)
    lv = np.concatenate([oracle.srgb_to_linear(x[: 3 * S * S]), x[3 * S * S:]])
    for k, (w, h) in enumerate(sizes):
        got = d_lv[k].download(np.float32)
        assert np.array_equal(got.view(np.uint32), lv.view(np.uint32)), (k, "mip level")
        exp = oracle.compress(N.FORMAT_BC3, lv, w, h, threads=16)
        assert _diff_blocks(d_out[k].download(np.uint8), exp, 16) == 0, (k, "blocks")
        if k + 1 < len(sizes):
            lv = oracle.build_next_mipmap(N.FILTER_KAISER, 3.0, (4.0, 1.0), N.WRAP_MIRROR, lv, w, h)
    for d in [d_bgra] + d_lv + d_out:
        d.free()


@pytest.mark.parametrize("schedule", [0, 1])
def test_c3_batch_matches_per_image_path(gpu, schedule):
    """256 textures of sizes in {64..2048}^2 (BC1 noise+rect, BC5 normal maps,
    BC4 single channel) with full Triangle chains in one batch graph, in both
    schedules (mip stages, or one persistent launch): every level equal to
    the per-image entry points (which the other tests pin to the oracle)."""
    rng = np.random.default_rng(1001)
    choices = [64, 128, 256, 512, 1024, 2048]
    base = {N.FORMAT_BC1: synth.noise_rect_rgba(2048, 2048, seed=2),
            N.FORMAT_BC5: synth.normal_map_rgba(2048, 2048, seed=3),
            N.FORMAT_BC4: synth.single_channel_rgba(2048, 2048, seed=4)}
    kinds = [N.FORMAT_BC1] * 128 + [N.FORMAT_BC5] * 64 + [N.FORMAT_BC4] * 64
    texs = []
    for fmt in kinds:
        w, h = int(rng.choice(choices)), int(rng.choice(choices))
        x0, y0 = int(rng.integers(0, 2048 - w + 1)), int(rng.integers(0, 2048 - h + 1))
        planar = synth.to_planar(base[fmt][y0:y0 + h, x0:x0 + w]).reshape(-1)
        texs.append(dict(d_rgba=gpu.alloc(planar.nbytes).upload(planar), width=w, height=h, format=fmt,
                         filter=N.FILTER_TRIANGLE, wrap=N.WRAP_MIRROR))
    with N.Batch(gpu, texs) as b:
        b.set_tuning(schedule=schedule)
        b.run()
        bad = []
        for i, t in enumerate(texs):
            sz = N.mip_sizes(t["width"], t["height"])
            cur = t["d_rgba"]
            for k, (w, h) in enumerate(sz):
                out = gpu.alloc(N.compressed_size(t["format"], w, h))
                gpu.compress_device(t["format"], cur, w, h, out)
                img = cur.download(np.float32)
                if not (np.array_equal(b.blocks(i, k), out.download(np.uint8))
                        and np.array_equal(b.image(i, k).view(np.uint32), img.view(np.uint32))):
                    bad.append((i, k))
                out.free()
                if k + 1 < len(sz):
                    nxt = gpu.alloc(16 * sz[k + 1][0] * sz[k + 1][1])
                    gpu.downsample_device(N.FILTER_TRIANGLE, cur, w, h, nxt)
                    if cur is not t["d_rgba"]:
                        cur.free()
                    cur = nxt
            if cur is not t["d_rgba"]:
                cur.free()
        assert not bad, f"{len(bad)} levels differ, first {bad[:5]}"
    for t in texs:
        t["d_rgba"].free()


def test_c4_plain_cluster_fit_16384(gpu, oracle):
    """16384^2 (16 tiles of 4 seeded 4096^2 textures in 4 rotations), plain
    BC1 Level8 through the two-pass encoder: every block equal to the
    reference's icbc (compiled from its sources) or, without it, the oracle."""
    import oracle_lib
    S, T = 16384, 4096
    base = [synth.noise_rect_rgba(T, T, seed=3001 + i) for i in range(4)]
    big = np.empty((S, S, 4), np.uint8)
    for k in range(16):
        ty, tx = divmod(k, 4)
        big[ty * T:(ty + 1) * T, tx * T:(tx + 1) * T] = np.rot90(base[k % 4], k // 4)
    del base
    planar = synth.to_planar(big).reshape(-1)
    del big
    d = gpu.alloc(planar.nbytes).upload(planar)
    out = gpu.alloc(N.compressed_size(N.FORMAT_BC1, S, S))
    gpu.compress_device(N.FORMAT_BC1, d, S, S, out)
    got = out.download(np.uint8)
    d.free()
    out.free()
    ref = oracle_lib.Reference.load()
    if ref is not None:
        exp = ref.compress_image(N.FORMAT_BC1, N.QUALITY_NORMAL, planar, S, S, threads=16)
    else:
        exp = oracle.compress(N.FORMAT_BC1, planar, S, S, threads=16)
    assert _diff_blocks(got, exp, 8) == 0


def test_c5_face_chain_bc1(gpu, oracle):
    """One 4096^2 cube face as the cube workload runs it: BGRA8 -> FP32,
    13-level Kaiser(3; 4, 1) chain (ntx_hip_downsample_chain_device), all
    levels compressed to plain BC1 in one ntx_hip_compress_images_device
    call.  Mips against the oracle's buildNextMipmap, blocks against the
    reference's icbc (or the oracle without it)."""
    import oracle_lib
    S = 4096
    img = synth.noise_rect_rgba(S, S, seed=4001)
    bgra = synth.to_bgra(img)
    sizes = N.mip_sizes(S, S)
    d_bgra = gpu.alloc(bgra.nbytes).upload(bgra)
    d_lv = [gpu.alloc(16 * w * h) for w, h in sizes]
    gpu.unpack_bgra8_device(d_bgra, S, S, d_lv[0])
    gpu.downsample_chain_device(N.FILTER_KAISER, d_lv, S, S, 3.0, (4.0, 1.0))
    d_out = [gpu.alloc(N.compressed_size(N.FORMAT_BC1, w, h)) for w, h in sizes]
    gpu.compress_images_device(N.FORMAT_BC1, d_lv, sizes, d_out)
    ref = oracle_lib.Reference.load()
    lv = synth.to_planar(img).reshape(-1)
    for k, (w, h) in enumerate(sizes):
        got = d_lv[k].download(np.float32)
        assert np.array_equal(got.view(np.uint32), lv.view(np.uint32)), (k, "mip level")
        if ref is not None:
            exp = ref.compress_image(N.FORMAT_BC1, N.QUALITY_NORMAL, lv, w, h, threads=16)
        else:
            exp = oracle.compress(N.FORMAT_BC1, lv, w, h, threads=16)
        assert _diff_blocks(d_out[k].download(np.uint8), exp, 8) == 0, (k, "blocks")
        if k + 1 < len(sizes):
            lv = oracle.build_next_mipmap(N.FILTER_KAISER, 3.0, (4.0, 1.0), N.WRAP_MIRROR, lv, w, h)
    for d in [d_bgra] + d_lv + d_out:
        d.free()
