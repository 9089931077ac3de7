#!/usr/bin/env python3
"""Benchmark of the NVTT hot path on one MI355X.

Headline (`value`): BC1 Mtexels/s on the 16384^2 workload (16,777,216 4x4
blocks of RGBA8 noise+rectangles, already resident in HBM as planar FP32) at
PSNR within 0.05 dB of the reference's cluster fit.  Two forms are timed over
the same `--steps`: the plain cluster fit (icbc Level8, byte-identical to the
reference) and the learned form (network proposals from the project's
BF16-MFMA network trained on the device beforehand, with the full cluster fit
as the fallback for blocks whose predicted error exceeds tau_pred, at the
shipped settings ntx_amd.learned.SHIPPED).  `value` is the learned form's
rate when its PSNR is within 0.05 dB of the plain fit's, else the plain fit's.
The CPU baseline is the reference's own BC1 compressor (compiled from the
reference sources into oracle/_ref) on the host's cores over the same image.

Secondary legs: the 4096^2 sRGB->Kaiser->BC3 chain (with the mip chain's HBM
GB/s, the metric's second part), the 1024^2 Box chain (C1), the InputOptions
pipeline, the 256-texture batch (C3) and the cube map (C5).
"""
import argparse
import json
import os
import sys
import time
from math import comb

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nvidia-texture-tools_amd"))

METRIC = "BC1 Mtexels/s at PSNR within 0.05 dB of ref; mip-chain HBM GB/s"
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E

# Algorithmic flops, counted from the reference loops each kernel restates:
#  BC1 icbc Level8 block: 134*N4(n) + 117*N3(m) + 3500 (n >= 2), 200 otherwise,
#  N4(n) = C(n+3,3)-1, N3(m) = C(m+2,2)-1.
#  BC3 colour (squish Compress4): 137 flops per candidate (the loop body of
#  weightedclusterfit.cpp:507-541, every add/mul/div/min/max/floor/compare
#  counted once) x C(k+3,3) candidates for k distinct colours, + 1500 set-up;
#  single-colour blocks take the table path (0).
SQUISH_PER_CANDIDATE = 137
SQUISH_SETUP = 1500
# The proposal form's kernel (bc1_mlp_kernel, no least-squares pass, no
# greedy round), per block: the network forward on the matrix cores,
# 2 * (48*64 + 64*64 + 64*25) = 17,536 flops; on the vector units, the
# features (block mean, largest deviation, scaling: 250), then for each of
# the 4 candidate lanes the colour bounding box (16 texels x 6 min/max = 96),
# the diagonal and the network's offsets (20) and two encodings (5:6:5
# endpoints, D3D10 palette, and for 16 texels 4 squared distances of 8 flops,
# the nearest of 4 and the weighted error: 2 x 650), and the uniform-block
# test (48): 250 + 4 * (96 + 20 + 1300) + 48 = 5,962 flops.
# BC4 / BC5 Normal channel block (QuickCompress::compressDXT5A with its
# refinement, QuickCompressDXT.cpp:505-592): about 3,500 operations.
CHANNEL_BLOCK_OPS = 3500
MLP_FORWARD_FLOPS = 2 * (48 * 64 + 64 * 64 + 64 * 25)
MLP_VALU_FLOPS = 250 + 4 * (96 + 20 + 2 * 650) + 48

# HBM bytes per launch of bc3_search_kernel on the 4096^2 chain (mean of its
# two launches, levels 0 and 1): FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 from the
# rocprofv3 --pmc runs of `bench.py --profile-pass` summarised in
# profiles/bc3_kaiser_chain_4096_kernel_trace.md (the factors come from a
# 256 MiB device copy in the same run).
SEARCH_TRAFFIC_BYTES = {"bc3_search_kernel": round((71.62e6 + 18.55e6 + 16.28e6 + 3.75e6) / 2)}

# The 16384^2 headline kernels, from the rocprofv3 passes of
# `bench.py --profile-pass --no-cpu --steps 3 --warmup 1` summarised in
# profiles/bc1_16384_learned_and_cluster_fit.md: HBM bytes per launch
# (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors from three 256 MiB copies
# in the same run) and VALU instructions issued per launch as a fraction of
# the chip's issue rate (256 CUs x 4 SIMDs x one wave64 instruction per 4
# cycles x 2.4 GHz), from SQ_INSTS_VALU.
HEADLINE_TRAFFIC_BYTES = {"bc1_mlp_kernel": round(3221.6e6 + 136.3e6), "bc1_search_kernel": round(9323.5e6 + 582.4e6)}
HEADLINE_VALU_ISSUE_FRAC = {"bc1_mlp_kernel": 0.884, "bc1_search_kernel": 0.985}
# Algorithmic bytes per 4x4 block on the headline: the R, G and B planes of
# 16 texels as FP32 (AlphaMode_None reads no alpha) in, one 8-byte block out.
HEADLINE_BYTES_PER_BLOCK = 16 * 3 * 4 + 8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bc1_flops(nm: np.ndarray) -> float:
    n = nm[:, 0].astype(np.int64)
    m = nm[:, 1].astype(np.int64)
    n4 = np.where(n >= 2, (n + 3) * (n + 2) * (n + 1) // 6 - 1, 0)
    n3 = np.where(m >= 1, (m + 2) * (m + 1) // 2 - 1, 0)
    f = np.where(n >= 2, 134 * n4 + 117 * n3 + 3500, 200)
    return float(f.sum())


def bc1_levels_flops(ctx, d_levels, sizes) -> float:
    """Counted flops of BC1 Level8 over the given device images, from each
    block's colour counts (ntx_hip_compress_bc1_stats_device)."""
    import ntx_amd as N
    total = 0.0
    for d, (w, h) in zip(d_levels, sizes):
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        d_nm, d_blk = ctx.alloc(4 * nb), ctx.alloc(N.compressed_size(N.FORMAT_BC1, w, h))
        ctx.compress_bc1_stats_device(d, w, h, d_blk, d_nm)
        total += bc1_flops(d_nm.download(np.int16).reshape(-1, 2))
        d_nm.free()
        d_blk.free()
    return total


def kernels_roofline(prof: dict, kernels, work: float, runs: int, bound: str = "valu") -> dict:
    """Counted work (flops) over the summed durations of `kernels` in a
    profiled pass of `runs` runs.  Where the kernels of one run overlap on
    several streams, their summed time exceeds the wall time, so the rate is
    a lower bound of what the kernels sustain."""
    ks = [k for k in kernels if k in prof]
    k_ms = sum(prof[k][0] for k in ks) / runs
    achieved = work / (k_ms / 1e3) / 1e12 if k_ms > 0 else 0.0
    return {"kernels": " + ".join(ks), "bound": bound, "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "algorithmic_flops_per_run": work,
            "kernel_ms_per_run": round(k_ms, 4),
            "kernels_ms_per_run": {k: round(v[0] / runs, 4) for k, v in prof.items()}}


def bc3_color_flops(planar: np.ndarray, w: int, h: int) -> float:
    """Distinct 8-bit RGB colours per block (ColorBlock::init truncation and
    edge repetition, ColourSet's minimal set) -> squish candidate count."""
    p = planar.reshape(4, h, w)[:3]
    u8 = (np.float32(255) * np.clip(p, np.float32(0), np.float32(1))).astype(np.uint8).astype(np.uint32)
    rgb = u8[0] | (u8[1] << 8) | (u8[2] << 16)
    bw, bh = (w + 3) // 4, (h + 3) // 4
    ys = np.array([[4 * by + (i % min(4, h - 4 * by)) for i in range(4)] for by in range(bh)])
    xs = np.array([[4 * bx + (e % min(4, w - 4 * bx)) for e in range(4)] for bx in range(bw)])
    blk = rgb[ys[:, None, :, None], xs[None, :, None, :]].reshape(bh * bw, 16)
    s = np.sort(blk, axis=1)
    k = 1 + np.count_nonzero(np.diff(s, axis=1), axis=1)
    cand = (k + 3) * (k + 2) * (k + 1) // 6
    f = np.where(k >= 2, SQUISH_PER_CANDIDATE * cand + SQUISH_SETUP, 0)
    return float(f.sum())


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    mse = np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2) * 3
    return float("inf") if mse == 0 else float(20 * np.log10(255.0 / np.sqrt(mse)))


def host_threads(cap: int) -> int:
    """Worker threads for the CPU baselines: the CPUs this process may run
    on, at most `cap` (the box's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(cap, n))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096, help="edge of the BC3 chain leg's texture")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-secondary", action="store_true", help="only the headline (C4) leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven-workload leg")
    ap.add_argument("--batch-textures", type=int, default=256)
    ap.add_argument("--learned-size", type=int, default=16384, help="edge of the headline texture")
    ap.add_argument("--profile-pass", action="store_true",
                    help="short run for rocprofv3 counter passes: the headline's two forms, the BC3 chain, the "
                         "Kaiser chain, the BC1 chain (2 steps each) and a 256 MiB device copy used to calibrate "
                         "FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the kernels run on a single MI355X; sharding by texture or block band is a design note only)",
              file=sys.stderr)
        sys.exit(2)

    import ntx_amd as N

    ctx = N.Context(0)
    net, training = train_network(ctx, args)
    result = bench_c4(ctx, args, net, training)
    if args.profile_pass:
        profile_pass(ctx, args)
        net.close()
        ctx.close()
        return
    if not args.no_secondary:
        c2 = bench_c2(ctx, args)
        result["mip_chain_gbs"] = c2.pop("mip_chain_gbs")
        result["mip_roofline"] = c2.pop("mip_roofline")
        result["c2_bc3_chain"] = c2
        result["c1_box_chain_bc1"] = bench_c1(ctx, args, not args.no_cpu)
        result["input_pipeline"] = bench_process(args, not args.no_cpu)
        if not args.no_batch:
            result["batch"] = bench_batch(ctx, args, not args.no_cpu)
        result["cube_map"] = bench_cube(ctx, args, net, not args.no_cpu)
    net.close()
    print(json.dumps(result), flush=True)
    ctx.close()


def train_network(ctx, args):
    """The 48-64-64-25 network (BF16, FP32 accumulation) trained on the
    device with the project's own kernels on the blocks of 8 seeded 2048^2
    noise+rectangle textures, each in 4 rotations with its Kaiser levels 1-3
    (11,141,120 blocks cut; the uniform ones dropped), labelled by the
    cluster fit.  Setup, not timed."""
    import ntx_amd as N
    from ntx_amd import learned, synth

    t0 = time.perf_counter()
    texs = []
    sizes = N.mip_sizes(2048, 2048)[:4]
    for i in range(8):
        img = synth.noise_rect_rgba(2048, 2048, seed=args.seed + 2000 + i)
        for r in range(4):
            # The texture and its first three Kaiser(3; 4, 1) levels: the
            # network also sees the smoother blocks of downsampled levels.
            d_lv = [ctx.alloc(16 * w * h) for w, h in sizes]
            d_lv[0].upload(synth.to_planar(np.ascontiguousarray(np.rot90(img, r))).reshape(-1))
            ctx.downsample_chain_device(N.FILTER_KAISER, d_lv, 2048, 2048, 3.0, (4.0, 1.0))
            for d, (w, h) in zip(d_lv, sizes):
                texs.append((d.download(np.float32), w, h))
                d.free()
    cut = sum(((w + 3) // 4) * ((h + 3) // 4) for _, w, h in texs)
    data = learned.make_dataset(ctx, texs, seed=args.seed)
    del texs
    data_s = time.perf_counter() - t0
    net = N.Mlp(ctx, seed=args.seed)
    log(f"training on {data.n} non-uniform blocks of {cut} cut")
    first, _ = net.train_step(data.feat(0), data.label(0), min(data.n, 65536), 0.0)
    t0 = time.perf_counter()
    epochs, batch = 3, 16384
    losses = learned.train(net, data, epochs=epochs, batch=batch, lr=3e-3)
    ctx.synchronize()
    train_s = time.perf_counter() - t0
    steps = epochs * ((data.n + batch - 1) // batch)
    n = data.n
    data.free()
    return net, {"textures": "8 seeded 2048^2 noise+rectangle textures (seeds %d-%d) x 4 rotations, each with "
                             "its Kaiser(3; 4, 1) levels 1-3" % (args.seed + 2000, args.seed + 2007),
                 "blocks_cut": cut, "blocks_used": n, "epochs": epochs, "batch": batch, "adam_steps": steps,
                 "seconds": round(train_s, 3), "data_seconds": round(data_s, 2), "loss_before": round(first, 5),
                 "loss_per_epoch": [round(x, 5) for x in losses]}


def c4_input(ctx, S: int, seed: int):
    """16 tiles of 4 seeded (S/4)^2 noise+rectangle textures in 4 rotations,
    uploaded as BGRA8 and unpacked on the device to planar FP32."""
    import ntx_amd as N
    from ntx_amd import synth

    T = S // 4
    base = [synth.noise_rect_rgba(T, T, seed=seed + 3000 + i) for i in range(4)]
    big = np.empty((S, S, 4), np.uint8)
    for k in range(16):
        ty, tx = divmod(k, 4)
        big[ty * T:(ty + 1) * T, tx * T:(tx + 1) * T] = np.rot90(base[k % 4], k // 4)
    del base
    bgra = synth.to_bgra(big)
    del big
    d_bgra = ctx.alloc(bgra.nbytes).upload(bgra)
    del bgra
    d = ctx.alloc(16 * S * S)
    ctx.unpack_bgra8_device(d_bgra, S, S, d)
    d_bgra.free()
    return d


def bench_c4(ctx, args, net, training) -> dict:
    """The headline: BC1 of the S^2 texture by the plain cluster fit and by
    the learned form, each timed over args.steps (after args.warmup), PSNR of
    both on the device, the per-kernel breakdown from a separate profiled pass
    of the same length, the rooflines and the CPU reference."""
    import ntx_amd as N
    from ntx_amd import learned

    S = args.learned_size
    texels = S * S
    nb = (S // 4) * (S // 4)
    log(f"headline: {S}x{S} input")
    d = c4_input(ctx, S, args.seed)
    d_plain = ctx.alloc(8 * nb)
    d_mlp = ctx.alloc(8 * nb)
    kw = dict(learned.SHIPPED)

    def plain():
        ctx.compress_device(N.FORMAT_BC1, d, S, S, d_plain)

    def proposals():
        net.compress_bc1_device(d, S, S, d_mlp, count_fallback=False, **kw)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.timer_begin()
        for _ in range(args.steps):
            fn()
        ms = ctx.timer_end()
        wall = time.perf_counter() - t0
        ctx.profile(True)
        ctx.profile_read(reset=True)
        for _ in range(args.steps):
            fn()
        prof = ctx.profile_read(reset=True)
        ctx.profile(False)
        return ms / args.steps, wall * 1e3 / args.steps, {k: (v[0] / args.steps, v[1] / args.steps)
                                                          for k, v in prof.items()}

    ms_plain, wall_plain, prof_plain = timed(plain)
    ms_mlp, wall_mlp, prof_mlp = timed(proposals)
    psnr_plain = ctx.bc1_psnr_device(d, S, S, d_plain)
    psnr_mlp = ctx.bc1_psnr_device(d, S, S, d_mlp)
    fallback = net.compress_bc1_device(d, S, S, d_mlp, **kw)
    delta = psnr_mlp - psnr_plain
    within = delta >= -0.05
    rate_plain = texels / (ms_plain / 1e3) / 1e6
    rate_mlp = texels / (ms_mlp / 1e3) / 1e6
    log(f"  plain {ms_plain:.3f} ms ({rate_plain:.0f} Mtexels/s, {psnr_plain:.3f} dB); learned {ms_mlp:.3f} ms "
        f"({rate_mlp:.0f} Mtexels/s, {delta:+.3f} dB, fallback {fallback / nb:.4f})")

    # Counted flops of the plain fit from its own per-block colour counts.
    d_nm = ctx.alloc(4 * nb)
    ctx.compress_bc1_stats_device(d, S, S, d_plain, d_nm)
    nm = d_nm.download(np.int16).reshape(-1, 2)
    d_nm.free()
    flops_plain = bc1_flops(nm)
    hist = np.bincount(nm[:, 0].clip(0, 16), minlength=17)
    search_ms = prof_plain.get("bc1_search_kernel", (0.0, 0))[0]

    def hbm_share(ms):
        """The launch's algorithmic bytes against the HBM peak."""
        if not ms:
            return None
        gbs = HEADLINE_BYTES_PER_BLOCK * nb / (ms / 1e3) / 1e9
        return {"achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
                "algorithmic_bytes_per_launch": HEADLINE_BYTES_PER_BLOCK * nb}

    search_flops = bc1_flops(nm[nm[:, 0] >= 2])
    plain_roofline = {"kernel": "bc1_search_kernel", "bound": "valu",
                      "achieved": round(search_flops / (search_ms / 1e3) / 1e12, 3) if search_ms else None,
                      "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                      "frac": round(search_flops / (search_ms / 1e3) / 1e12 / PEAK_FP32_TFLOPS, 4) if search_ms else None,
                      "traffic": HEADLINE_TRAFFIC_BYTES["bc1_search_kernel"], "avg_launch_ms": round(search_ms, 4),
                      "algorithmic_flops_per_launch": search_flops,
                      "valu_issue_frac_measured": HEADLINE_VALU_ISSUE_FRAC["bc1_search_kernel"],
                      "hbm_share": hbm_share(search_ms),
                      "note": "icbc Level8 flops counted per block from its colour counts (n, m), blocks of two or "
                              "more colours; the classify pass finishes the others"}
    # The learned step's dominant kernel.
    mlp_ms = prof_mlp.get("bc1_mlp_kernel", (0.0, 0))[0]
    valu = MLP_VALU_FLOPS * nb / (mlp_ms / 1e3) / 1e12
    mfma = MLP_FORWARD_FLOPS * nb / (mlp_ms / 1e3) / 1e12
    mlp_roofline = {"kernel": "bc1_mlp_kernel", "bound": "valu", "achieved": round(valu, 3), "peak": PEAK_FP32_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(valu / PEAK_FP32_TFLOPS, 4),
                    "traffic": HEADLINE_TRAFFIC_BYTES["bc1_mlp_kernel"],
                    "avg_launch_ms": round(mlp_ms, 4), "algorithmic_flops_per_launch": MLP_VALU_FLOPS * nb,
                    "valu_issue_frac_measured": HEADLINE_VALU_ISSUE_FRAC["bc1_mlp_kernel"],
                    "hbm_share": hbm_share(mlp_ms),
                    "mfma_share": {"achieved": round(mfma, 3), "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                                   "frac": round(mfma / PEAK_BF16_TFLOPS, 5),
                                   "algorithmic_flops_per_launch": MLP_FORWARD_FLOPS * nb}}
    head_is_learned = within
    dominant = max(prof_mlp if head_is_learned else prof_plain, key=lambda k: (prof_mlp if head_is_learned else
                                                                                prof_plain)[k][0])
    roofline = mlp_roofline if head_is_learned else plain_roofline
    roofline = dict(roofline, dominant_kernel_of_step=dominant)

    result = {
        "metric": METRIC,
        "value": round(rate_mlp if head_is_learned else rate_plain, 2),
        "unit": "Mtexels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mlp if head_is_learned else ms_plain, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{S}x{S} RGBA8 noise+rectangles (16 tiles: 4 seeded {S // 4}^2 textures x 4 rotations), "
                        f"planar FP32 in HBM -> BC1 ({'learned proposals + cluster-fit fallback' if head_is_learned else 'plain cluster fit'})",
            "texels_per_step": texels,
            "blocks_per_step": nb,
            "form": "learned" if head_is_learned else "plain",
            "input_resident_in_hbm": True,
        },
        "roofline": roofline,
        "psnr_gate": {"plain_db": round(psnr_plain, 4), "learned_db": round(psnr_mlp, 4), "delta_db": round(delta, 4),
                      "limit_db": -0.05, "within": bool(within)},
        "plain_cluster_fit": {"ms_per_step": round(ms_plain, 4), "wall_ms_per_step": round(wall_plain, 4),
                              "mtexels_s": round(rate_plain, 2), "psnr_db": round(psnr_plain, 4),
                              "kernels_ms_per_step": {k: round(v[0], 4) for k, v in prof_plain.items()},
                              "roofline": plain_roofline,
                              "algorithmic_flops_per_step": flops_plain,
                              "colour_count_histogram": hist.tolist()},
        "learned": {"ms_per_step": round(ms_mlp, 4), "wall_ms_per_step": round(wall_mlp, 4),
                    "mtexels_s": round(rate_mlp, 2), "psnr_db": round(psnr_mlp, 4),
                    "speedup_vs_plain": round(ms_plain / ms_mlp, 3), "settings": kw,
                    "fallback_blocks": int(fallback), "fallback_fraction": round(fallback / nb, 5),
                    "kernels_ms_per_step": {k: round(v[0], 4) for k, v in prof_mlp.items()},
                    "roofline": mlp_roofline,
                    "target": "learned >= 3x plain at PSNR within 0.05 dB",
                    "meets_target": bool(ms_plain / ms_mlp >= 3.0 and within),
                    "training": training},
    }

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_reference_bc1(ctx, args, d, S, d_plain, texels)
    for b in (d, d_plain, d_mlp):
        b.free()
    return result


def cpu_reference_bc1(ctx, args, d, S, d_plain, texels) -> dict:
    """The reference's own BC1 compressor (CompressorDXT1 -> icbc Level8,
    compiled from the reference sources into oracle/_ref) over the same image
    on the host, with the reference's task loop (one block per task from an
    atomic counter) on host_threads() workers; its blocks against the
    device's plain fit, and the oracle port's rate beside it on a 2048^2
    crop (the port/reference calibration)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    ref = oracle_lib.Reference.load()
    threads = host_threads(args.cpu_threads)
    planar = d.download(np.float32)
    gpu = d_plain.download(np.uint8)
    info = {"unit": "Mtexels/s", "cores": threads, "cpu_model": cpu_model(), "nproc": os.cpu_count()}
    if ref is None:
        return dict(info, value=None, kind="reference", sample="oracle/_ref not built")
    log(f"CPU reference: BC1 of {S}x{S} on {threads} threads")
    t0 = time.perf_counter()
    out = ref.compress_image(1, 1, planar, S, S, threads=threads)
    sec = time.perf_counter() - t0
    same = float(np.mean(np.all(out.reshape(-1, 8) == gpu.reshape(-1, 8), axis=1)))
    # Port / reference calibration on a 2048^2 crop of the same image.
    C = min(2048, S)
    crop = np.ascontiguousarray(planar.reshape(4, S, S)[:, :C, :C]).reshape(-1)
    orc = oracle_lib.Oracle()
    t0 = time.perf_counter()
    a = ref.compress_image(1, 1, crop, C, C, threads=threads)
    t_ref = time.perf_counter() - t0
    t0 = time.perf_counter()
    b = orc.compress(1, crop, C, C, threads=threads)
    t_port = time.perf_counter() - t0
    return dict(info, value=round(texels / sec / 1e6, 3), kind="reference",
                sample=f"the whole headline image ({texels} texels) through the reference's CompressorDXT1 "
                       f"(icbc Level8, ICBC_SIMD=2 build of the reference sources)",
                seconds=round(sec, 3), blocks_identical_to_device=same,
                port_calibration={"crop": f"{C}x{C}", "reference_mtexels_s": round(C * C / t_ref / 1e6, 3),
                                  "port_mtexels_s": round(C * C / t_port / 1e6, 3),
                                  "port_over_reference_time": round(t_port / t_ref, 3),
                                  "blocks_identical": bool(np.array_equal(a, b))})


def profile_pass(ctx, args):
    """A short run for rocprofv3 counter passes (see --profile-pass)."""
    import ntx_amd as N

    S = args.size
    levels = N.mip_sizes(S, S)
    from ntx_amd import synth
    img = synth.noise_rect_rgba(S, S, seed=args.seed, transparent=True)
    bgra = synth.to_bgra(img)
    d_bgra = ctx.alloc(bgra.nbytes).upload(bgra)
    d_lv = [ctx.alloc(16 * w * h) for w, h in levels]
    d_bc3 = [ctx.alloc(N.compressed_size(N.FORMAT_BC3, w, h)) for w, h in levels]
    d_bc1 = [ctx.alloc(N.compressed_size(N.FORMAT_BC1, w, h)) for w, h in levels]
    kw, kp = 3.0, (4.0, 1.0)
    for _ in range(2):
        ctx.unpack_bgra8_srgb_device(d_bgra, S, S, d_lv[0])
        ctx.compress_chain_device(N.FORMAT_BC3, N.FILTER_KAISER, d_lv, S, S, d_bc3, width=kw, params=kp)
    for _ in range(2):
        ctx.downsample_chain_device(N.FILTER_KAISER, d_lv, S, S, kw, kp)
    for _ in range(2):
        for k, (w, h) in enumerate(levels):
            ctx.compress_device(N.FORMAT_BC1, d_lv[k], w, h, d_bc1[k])
    ctx.synchronize()
    nbytes = 256 << 20
    a, b = ctx.alloc(nbytes), ctx.alloc(nbytes)
    for _ in range(3):
        ctx.lib.ntx_hip_copy_device(ctx.handle, b.ptr, a.ptr, nbytes)
    ctx.synchronize()
    log(f"profile pass done: headline forms, BC3 chain/Kaiser/BC1 x2, 3 copies of {nbytes} bytes")


def bench_c2(ctx, args) -> dict:
    """The 4096^2 sRGB -> Kaiser -> BC3 chain, the Kaiser chain alone (the
    mip-chain HBM GB/s), BC1 over the same levels, and the CPU port."""
    import ntx_amd as N
    from ntx_amd import synth

    S = args.size
    levels = N.mip_sizes(S, S)
    texels = sum(w * h for w, h in levels)
    kw, kp = 3.0, (4.0, 1.0)

    log(f"generating {S}x{S} input (seed {args.seed})")
    img = synth.noise_rect_rgba(S, S, seed=args.seed, transparent=True)
    bgra = synth.to_bgra(img)

    d_bgra = ctx.alloc(bgra.nbytes).upload(bgra)
    d_lv = [ctx.alloc(16 * w * h) for w, h in levels]
    d_bc3 = [ctx.alloc(N.compressed_size(N.FORMAT_BC3, w, h)) for w, h in levels]
    d_bc1 = [ctx.alloc(N.compressed_size(N.FORMAT_BC1, w, h)) for w, h in levels]

    def mips():
        ctx.downsample_chain_device(N.FILTER_KAISER, d_lv, S, S, kw, kp)

    def step_bc3():
        """The workload through the chain entry point: each level's BC3
        compression overlaps the filtering and compression of the others;
        setImage and toLinearFromSrgb as one pass."""
        ctx.unpack_bgra8_srgb_device(d_bgra, S, S, d_lv[0])
        ctx.compress_chain_device(N.FORMAT_BC3, N.FILTER_KAISER, d_lv, S, S, d_bc3, width=kw, params=kp)

    def step_bc3_per_level():
        """The same workload as nvtt's per-level calls, one after another."""
        ctx.unpack_bgra8_device(d_bgra, S, S, d_lv[0])
        ctx.srgb_to_linear_device(d_lv[0], S, S)
        for k, (w, h) in enumerate(levels):
            ctx.compress_device(N.FORMAT_BC3, d_lv[k], w, h, d_bc3[k])
            if k + 1 < len(levels):
                ctx.downsample_device(N.FILTER_KAISER, d_lv[k], w, h, d_lv[k + 1], kw, kp)

    def step_bc1():
        for k, (w, h) in enumerate(levels):
            ctx.compress_device(N.FORMAT_BC1, d_lv[k], w, h, d_bc1[k])

    def step_bc1_images():
        ctx.compress_images_device(N.FORMAT_BC1, d_lv, levels, d_bc1)

    def timed(fn, steps, warmup):
        """Time `steps` runs with nothing else on the streams (the value),
        then the same number of runs with per-kernel event pairs (the
        breakdown and the rooflines; the events themselves cost queue time,
        so they stay out of the timed runs)."""
        for _ in range(warmup):
            fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.timer_begin()
        for _ in range(steps):
            fn()
        ms = ctx.timer_end()
        wall = time.perf_counter() - t0
        ctx.profile(True)
        ctx.profile_read(reset=True)
        for _ in range(steps):
            fn()
        prof = ctx.profile_read(reset=True)
        ctx.profile(False)
        return ms, wall, prof

    log(f"BC3 pipeline: {args.warmup} warm-up + {args.steps} timed steps")
    ms, wall, prof = timed(step_bc3, args.steps, args.warmup)
    value = texels * args.steps / (ms / 1e3) / 1e6
    log(f"  {ms / args.steps:.3f} ms/step (events), wall {wall * 1e3 / args.steps:.3f} ms/step, {value:.1f} Mtexels/s")
    gpu_bc3_chain = [d.download(np.uint8) for d in d_bc3]
    log("BC3 pipeline, per-level calls")
    ms_pl, _, prof_pl = timed(step_bc3_per_level, args.steps, 1)
    same_as_per_level = all(np.array_equal(a, d.download(np.uint8)) for a, d in zip(gpu_bc3_chain, d_bc3))

    # Dominant kernel and its algorithmic flops.  A level runs
    # bc3_search_kernel (one block per lane) when it has at least 2^t blocks,
    # else bc3_search_split_kernel (several lanes per block): t = 18 for
    # per-level calls and for level 0 of a chain, 16 for the other levels of
    # a chain (they run beside other work).  The roofline is taken from the
    # per-level pass, where each kernel has the device to itself; the same
    # kernels' durations inside the chain (sharing the device) are reported
    # beside it.
    lv_host = [d.download(np.float32) for d in d_lv]
    lv_flops = [bc3_color_flops(lv_host[k], w, h) for k, (w, h) in enumerate(levels)]
    nblk = [((w + 3) // 4) * ((h + 3) // 4) for w, h in levels]

    def flops_by_kernel(chain):
        one = [n >= (1 << (16 if chain and k else 18)) for k, n in enumerate(nblk)]
        return {"bc3_search_kernel": sum(f for f, o in zip(lv_flops, one) if o),
                "bc3_search_split_kernel": sum(f for f, o in zip(lv_flops, one) if not o)}

    def valu_roofline(kernel, pr, flops_of):
        k_ms, k_n = pr[kernel]
        achieved = flops_of[kernel] / (k_ms / args.steps / 1e3) / 1e12
        return {"kernel": kernel, "bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": SEARCH_TRAFFIC_BYTES.get(kernel), "avg_launch_ms": round(k_ms / k_n, 5),
                "launches_per_step": k_n / args.steps, "algorithmic_flops_per_step": flops_of[kernel],
                "algorithmic_flops_per_launch": flops_of[kernel] / (k_n / args.steps)}

    dom = max(prof_pl, key=lambda k: prof_pl[k][0])
    fl_pl, fl_chain = flops_by_kernel(False), flops_by_kernel(True)
    if dom in fl_pl:
        roofline = valu_roofline(dom, prof_pl, fl_pl)
        roofline["measured_in"] = "per-level pass (kernel alone on the device)"
        if dom in prof:
            ch = valu_roofline(dom, prof, fl_chain)
            roofline["in_chain"] = {k: ch[k] for k in ("achieved", "frac", "avg_launch_ms", "launches_per_step")}
    else:
        roofline = {"kernel": dom, "bound": "unknown", "avg_launch_ms": prof_pl[dom][0] / prof_pl[dom][1]}
    split_roofline = (valu_roofline("bc3_search_split_kernel", prof_pl, fl_pl)
                      if "bc3_search_split_kernel" in prof_pl else None)

    log("mip chain (Kaiser) alone")
    ms_mip, _, prof_mip = timed(mips, args.steps, 1)
    mip_bytes = sum(16 * (w * h + levels[k + 1][0] * levels[k + 1][1]) for k, (w, h) in enumerate(levels[:-1]))
    mip_gbs = mip_bytes * args.steps / (ms_mip / 1e3) / 1e9

    log("BC1 over the same chain")
    nm = [ctx.alloc(4 * ((w + 3) // 4) * ((h + 3) // 4)) for w, h in levels]
    for k, (w, h) in enumerate(levels):
        ctx.compress_bc1_stats_device(d_lv[k], w, h, d_bc1[k], nm[k])
    ctx.synchronize()
    flops_bc1 = sum(bc1_flops(x.download(np.int16).reshape(-1, 2)) for x in nm)
    ms_bc1, _, prof_bc1 = timed(step_bc1, args.steps, 1)
    bc1_mtex_per_level = texels * args.steps / (ms_bc1 / 1e3) / 1e6
    bc1_kernel_ms = sum(prof_bc1.get(k, (0.0, 0))[0] for k in ("bc1_kernel", "bc1_classify_kernel",
                                                                 "bc1_search_kernel")) / args.steps
    bc1_tflops = flops_bc1 / (bc1_kernel_ms / 1e3) / 1e12
    per_level_bc1 = [d.download(np.uint8) for d in d_bc1]
    ms_bc1i, _, _ = timed(step_bc1_images, args.steps, 1)
    bc1_mtex = texels * args.steps / (ms_bc1i / 1e3) / 1e6
    bc1_images_same = all(np.array_equal(a, d.download(np.uint8)) for a, d in zip(per_level_bc1, d_bc1))
    # The other BC1 qualities over the same chain (icbc Level1 for Fastest,
    # Level9 for Production), and the device decode + rmsError of level 0.
    bc1_q = {}
    for qname, q in (("fastest_level1", N.QUALITY_FASTEST), ("production_level9", N.QUALITY_PRODUCTION)):
        ms_q, _, _ = timed(lambda: ctx.compress_images_device(N.FORMAT_BC1, d_lv, levels, d_bc1, quality=q), args.steps, 1)
        bc1_q[qname] = {"mtexels_s": round(texels * args.steps / (ms_q / 1e3) / 1e6, 1),
                        "level0_psnr_db": round(ctx.bc1_psnr_device(d_lv[0], S, S, d_bc1[0]), 3)}
    d_dec = ctx.alloc(16 * S * S)
    step_bc1_images()  # the Normal-quality blocks again, for the checks below
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.decode_device(N.FORMAT_BC1, d_bc1[0], S, S, d_dec)
        rms = ctx.rms_error_device(N.METRIC_RMS_COLOR, d_lv[0], d_dec, S, S)
    dec_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    d_dec.free()
    bc1_q["normal_level8_psnr_db"] = round(ctx.bc1_psnr_device(d_lv[0], S, S, d_bc1[0]), 3)
    bc1_q["decode_and_rms_level0"] = {"ms": round(dec_ms, 3), "rms_error": round(rms, 6),
                                      "note": "setImage2D(BC1) + rmsError on the device, host-timed incl. the result copy"}

    gpu_bc3 = [d.download(np.uint8) for d in d_bc3]
    gpu_bc1 = [d.download(np.uint8) for d in d_bc1]


    c2 = {
        "workload": f"{S}x{S} RGBA8 noise+rectangles, 30% transparent alpha (seed {args.seed}) -> "
                    f"sRGB->linear FP32 -> {len(levels)}-level Kaiser(3; 4, 1) mip chain -> BC3 per level",
        "texels_per_step": texels,
        "mtexels_s": round(value, 2),
        "ms_per_step": round(ms / args.steps, 4),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "roofline": roofline,
        "split_search_roofline": split_roofline,
        "per_level_calls": {"ms_per_step": round(ms_pl / args.steps, 4),
                            "mtexels_s": round(texels * args.steps / (ms_pl / 1e3) / 1e6, 2),
                            "blocks_identical_to_chain": bool(same_as_per_level),
                            "kernels_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in prof_pl.items()}},
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in prof.items()},
        "mip_chain_gbs": round(mip_gbs, 1),
        "mip_roofline": {"bound": "hbm", "achieved": round(mip_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(mip_gbs / PEAK_HBM_GBS, 4), "algorithmic_bytes_per_chain": mip_bytes,
                         "kernels_ms_per_chain": {k: round(v[0] / args.steps, 4) for k, v in prof_mip.items()}},
        "bc1_mtexels_s": round(bc1_mtex, 2),
        "bc1_levels_together": {"entry_point": "ntx_hip_compress_images_device (the 13 levels scheduled together)",
                                "per_level_calls_mtexels_s": round(bc1_mtex_per_level, 2),
                                "blocks_identical_to_per_level": bool(bc1_images_same)},
        "bc1_qualities": bc1_q,
        "bc1_roofline": {"kernels": "bc1_kernel + bc1_classify_kernel + bc1_search_kernel", "bound": "valu",
                         "achieved": round(bc1_tflops, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(bc1_tflops / PEAK_FP32_TFLOPS, 4), "algorithmic_flops_per_chain": flops_bc1,
                         "kernel_ms_per_chain": round(bc1_kernel_ms, 4),
                         "kernels_ms_per_chain": {k: round(v[0] / args.steps, 4) for k, v in prof_bc1.items()}},
    }

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        threads = host_threads(args.cpu_threads)
        orc = oracle_lib.Oracle()
        log(f"CPU baseline: oracle, {threads} threads, full workload")
        t0 = time.perf_counter()
        x = synth.to_planar(img).reshape(-1)
        x = np.concatenate([orc.srgb_to_linear(x[: 3 * S * S]), x[3 * S * S:]])
        cpu_lv = [x]
        cpu_bc3 = []
        t_mip = 0.0
        for k, (w, h) in enumerate(levels):
            cpu_bc3.append(orc.compress(N.FORMAT_BC3, cpu_lv[k], w, h, threads=threads))
            if k + 1 < len(levels):
                tm = time.perf_counter()
                cpu_lv.append(orc.build_next_mipmap(N.FILTER_KAISER, kw, kp, N.WRAP_MIRROR, cpu_lv[k], w, h))
                t_mip += time.perf_counter() - tm
        cpu_s = time.perf_counter() - t0
        same = [np.array_equal(a.reshape(-1, 16), b.reshape(-1, 16)) for a, b in zip(gpu_bc3, cpu_bc3)]
        ident = sum(int(np.sum(np.all(a.reshape(-1, 16) == b.reshape(-1, 16), axis=1))) for a, b in zip(gpu_bc3, cpu_bc3))
        nblk = sum(len(a) // 16 for a in cpu_bc3)
        mips_equal = all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(lv_host, cpu_lv))
        c2["cpu_port"] = {
            "value": round(texels / cpu_s / 1e6, 3), "unit": "Mtexels/s", "cores": threads, "kind": "port",
            "sample": f"the full workload ({texels} texels, {len(levels)} levels): oracle BC3 with {threads} threads "
                      f"(one block per task, atomic counter), Kaiser mips single-threaded as in the reference "
                      f"({t_mip:.2f} s of {cpu_s:.2f} s)",
            "seconds": round(cpu_s, 3),
        }
        c2["parity"] = {"bc3_identical_blocks": ident / nblk, "bc3_levels_identical": all(same),
                            "mip_levels_bit_identical": bool(mips_equal)}
        ref = oracle_lib.Reference.load()
        if ref is not None:
            # The workload through the reference's own code: Kaiser levels by
            # FloatImage::downSample (single-threaded, as the reference runs
            # it) and BC3 by CompressorDXT5 on the reference's task loop;
            # the sRGB->linear step (Surface.cpp:1562-1592, a per-texel
            # table, not built into oracle/_ref) from the port.
            log(f"CPU reference: BC3 chain (reference sources), {threads} threads")
            t0 = time.perf_counter()
            x = synth.to_planar(img).reshape(-1)
            cur = np.concatenate([orc.srgb_to_linear(x[: 3 * S * S]), x[3 * S * S:]])
            t_srgb = time.perf_counter() - t0
            ref_same = 0
            t_rmip = 0.0
            for k, (w, h) in enumerate(levels):
                blk = ref.compress_image(N.FORMAT_BC3, N.QUALITY_NORMAL, cur, w, h, threads=threads)
                ref_same += int(np.sum(np.all(blk.reshape(-1, 16) == gpu_bc3[k].reshape(-1, 16), axis=1)))
                if k + 1 < len(levels):
                    tm = time.perf_counter()
                    cur = ref.build_next_mipmap(N.FILTER_KAISER, kw, kp, N.WRAP_MIRROR, cur, w, h)
                    t_rmip += time.perf_counter() - tm
            ref_bc3_s = time.perf_counter() - t0
            c2["cpu_reference"] = {
                "value": round(texels / ref_bc3_s / 1e6, 3), "unit": "Mtexels/s", "cores": threads,
                "kind": "reference", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                "sample": f"the full workload ({texels} texels): the reference's CompressorDXT5 on {threads} threads "
                          f"and its FloatImage::downSample Kaiser levels single-threaded ({t_rmip:.2f} s), sRGB->linear "
                          f"by the port ({t_srgb:.2f} s), {ref_bc3_s:.2f} s in all",
                "seconds": round(ref_bc3_s, 3), "bc3_identical_blocks_to_device": round(ref_same / nblk, 6),
                "port_over_reference_time": round(cpu_s / ref_bc3_s, 3)}
            import ctypes
            log(f"CPU reference: icbc (reference sources) BC1 on the chain, {threads} threads")
            fn = ref.lib.ntxref_compress_bc1_image
            fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
            t0 = time.perf_counter()
            ref_bc1 = []
            for k, (w, h) in enumerate(levels):
                out = np.zeros(N.compressed_size(N.FORMAT_BC1, w, h), np.uint8)
                src = np.ascontiguousarray(lv_host[k], np.float32)
                fn(src.ctypes.data, w, h, out.ctypes.data, threads)
                ref_bc1.append(out)
            ref_s = time.perf_counter() - t0
            ident1 = sum(int(np.sum(np.all(a.reshape(-1, 8) == b.reshape(-1, 8), axis=1))) for a, b in zip(gpu_bc1, ref_bc1))
            nb1 = sum(len(a) // 8 for a in ref_bc1)
            # PSNR of both encodings against the FP32 level 0 quantised to 8 bits
            lvl0 = lv_host[0].reshape(4, S, S)[:3]
            src8 = np.floor(np.clip(lvl0, 0, 1) * 255 + 0.5).astype(np.uint8)
            dec = lambda blk: _decode_bc1(blk, S, S)
            p_gpu, p_ref = psnr(dec(gpu_bc1[0]), src8), psnr(dec(ref_bc1[0]), src8)
            c2["cpu_reference_bc1"] = {
                "value": round(texels / ref_s / 1e6, 3), "unit": "Mtexels/s", "cores": threads, "kind": "reference",
                "sample": "icbc Level8 from the reference sources over the same 13 FP32 levels",
                "bc1_identical_blocks": ident1 / nb1, "psnr_level0_gpu_db": round(p_gpu, 4),
                "psnr_level0_ref_db": round(p_ref, 4)}

    for buf in [d_bgra] + d_lv + d_bc3 + d_bc1 + nm:
        buf.free()
    return c2


def bench_batch(ctx, args, cpu: bool = False) -> dict:
    """Batched uneven workload: n textures with power-of-two sizes drawn
    uniformly from {64..2048} x {64..2048}; half RGBA noise+rectangles -> BC1,
    a quarter normal maps (packed XY) -> BC5, a quarter single-channel noise
    -> BC4; full Triangle mip chains.  One HIP graph per run (a launch per mip
    stage, the persistent compression launch, the two-pass BC1 encoder), with
    the autotuned work decomposition.  Each texture is
    cut from one of three 2048x2048 generated images at a seeded offset."""
    import ntx_amd as N
    from ntx_amd import synth

    n = args.batch_textures
    rng = np.random.default_rng(args.seed + 1000)
    sizes = [64, 128, 256, 512, 1024, 2048]
    log(f"batch: generating {n} textures")
    base = {N.FORMAT_BC1: synth.noise_rect_rgba(2048, 2048, seed=args.seed + 1),
            N.FORMAT_BC5: synth.normal_map_rgba(2048, 2048, seed=args.seed + 2),
            N.FORMAT_BC4: synth.single_channel_rgba(2048, 2048, seed=args.seed + 3)}
    kinds = [N.FORMAT_BC1] * (n // 2) + [N.FORMAT_BC5] * (n // 4) + [N.FORMAT_BC4] * (n - n // 2 - n // 4)
    texs = []
    texels = 0
    for i, fmt in enumerate(kinds):
        w, h = int(rng.choice(sizes)), int(rng.choice(sizes))
        x0, y0 = int(rng.integers(0, 2048 - w + 1)), int(rng.integers(0, 2048 - h + 1))
        planar = synth.to_planar(base[fmt][y0:y0 + h, x0:x0 + w]).reshape(-1)
        texs.append(dict(d_rgba=ctx.alloc(planar.nbytes).upload(planar), width=w, height=h, format=fmt,
                         filter=N.FILTER_TRIANGLE, wrap=N.WRAP_MIRROR))
        texels += sum(a * b for a, b in N.mip_sizes(w, h))
    b = N.Batch(ctx, texs)
    log("batch: autotuning")
    t0 = time.perf_counter()
    tuning, tune_ms = b.autotune(reps=2)
    tune_s = time.perf_counter() - t0
    steps = max(3, args.steps // 4)
    b.run()
    ctx.synchronize()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    ctx.timer_begin()
    for _ in range(steps):
        b.submit()
    ms = ctx.timer_end() / steps
    b.wait()
    prof = ctx.profile_read(reset=True)
    ctx.profile(False)

    # The same work through the per-image entry points (one launch per level
    # and kernel), and a byte-for-byte comparison of every level.
    log("batch: per-image path for comparison")
    chains = []
    for t in texs:
        w, h = t["width"], t["height"]
        lv = [t["d_rgba"]] + [ctx.alloc(16 * a * c) for a, c in N.mip_sizes(w, h)[1:]]
        out = [ctx.alloc(N.compressed_size(t["format"], a, c)) for a, c in N.mip_sizes(w, h)]
        chains.append((lv, out))

    def per_image():
        for t, (lv, out) in zip(texs, chains):
            sz = N.mip_sizes(t["width"], t["height"])
            for k, (a, c) in enumerate(sz):
                ctx.compress_device(t["format"], lv[k], a, c, out[k])
                if k + 1 < len(sz):
                    ctx.downsample_device(N.FILTER_TRIANGLE, lv[k], a, c, lv[k + 1])
    per_image()
    ctx.synchronize()
    ctx.timer_begin()
    per_image()
    ms_seq = ctx.timer_end()
    same = 0
    nlev = 0
    for i, (t, (lv, out)) in enumerate(zip(texs, chains)):
        for k in range(b.levels(i)):
            nlev += 1
            same += int(np.array_equal(b.blocks(i, k), out[k].download(np.uint8)))
    b.close()
    # Roofline of the batch run (the mip stages, the persistent compression
    # kernel and the two-pass BC1 launches of one graph): the counted
    # compression work (BC1 by each block's colour counts, BC4 / BC5 per
    # channel block) over its time; the mips (16 B per source and output
    # texel) are part of the same graph.
    work = 0.0
    mip_bytes = 0
    for t, (lv, out) in zip(texs, chains):
        sz = N.mip_sizes(t["width"], t["height"])
        nblk = sum(((a + 3) // 4) * ((c + 3) // 4) for a, c in sz)
        if t["format"] == N.FORMAT_BC1:
            work += bc1_levels_flops(ctx, lv, sz)
        else:
            work += CHANNEL_BLOCK_OPS * nblk * (2 if t["format"] == N.FORMAT_BC5 else 1)
        mip_bytes += sum(16 * (a * c + sz[k + 1][0] * sz[k + 1][1]) for k, (a, c) in enumerate(sz[:-1]))
    roof = kernels_roofline(prof, ("batch_kernel",), work, steps)
    roof["mip_bytes_per_run"] = mip_bytes
    roof["note"] = "compression work only in `achieved`; the same graph also writes every mip level"
    for lv, out in chains:
        for d in lv[1:] + out:
            d.free()
    return {
        "workload": f"{n} textures, sizes uniform over {{64..2048}}^2 (powers of two), {n // 2} BC1 noise+rect, "
                    f"{n // 4} BC5 normal maps, {n - n // 2 - n // 4} BC4 single-channel; full Triangle mip chains; "
                    f"one HIP graph per run (mip stages, compression queue, two-pass BC1 over every BC1 level)",
        "texels_per_run": texels,
        "mtexels_s": round(texels / (ms / 1e3) / 1e6, 2),
        "ms_per_run": round(ms, 4),
        "runs": steps,
        "tuning": tuning,
        "autotune_s": round(tune_s, 3),
        "per_image_path_ms": round(ms_seq, 4),
        "speedup_vs_per_image_path": round(ms_seq / ms, 3),
        "levels_identical_to_per_image_path": f"{same}/{nlev}",
        "kernels_ms_per_run": {k: round(v[0] / steps, 4) for k, v in prof.items()},
        "roofline": roof,
        **({"cpu_reference": cpu_reference_batch(args, texs, texels)} if cpu else {}),
    }


def cpu_reference_batch(args, texs, texels_total) -> dict:
    """The reference's own CPU code over a bounded sample of the batch: the
    first textures of each format (8 BC1, 4 BC5, 4 BC4) in full, each level
    by Surface::buildNextMipmap's Triangle filter (FloatImage::downSample,
    single-threaded as in the reference) and the compressor Context.cpp picks
    (CompressorDXT1 / FastCompressorBC5 / FastCompressorBC4) on
    host_threads() workers, all compiled from the reference sources."""
    import ntx_amd as N
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    ref = oracle_lib.Reference.load()
    threads = host_threads(args.cpu_threads)
    info = {"unit": "Mtexels/s", "cores": threads, "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "kind": "reference"}
    if ref is None:
        return dict(info, value=None, sample="oracle/_ref not built")
    want = {N.FORMAT_BC1: 8, N.FORMAT_BC5: 4, N.FORMAT_BC4: 4}
    sample = []
    for t in texs:
        if want.get(t["format"], 0) > 0:
            want[t["format"]] -= 1
            sample.append(t)
    host = [t["d_rgba"].download(np.float32) for t in sample]
    log(f"CPU reference: batch sample of {len(sample)} textures on {threads} threads")
    t0 = time.perf_counter()
    tex = 0
    for t, planar in zip(sample, host):
        sz = N.mip_sizes(t["width"], t["height"])
        cur = planar
        for k, (w, h) in enumerate(sz):
            ref.compress_image(t["format"], N.QUALITY_NORMAL, cur, w, h, threads=threads)
            tex += w * h
            if k + 1 < len(sz):
                cur = ref.build_next_mipmap(N.FILTER_TRIANGLE, 1.0, None, N.WRAP_MIRROR, cur, w, h)
    sec = time.perf_counter() - t0
    return dict(info, value=round(tex / sec / 1e6, 3), seconds=round(sec, 3),
                sample=f"{len(sample)} of the {len(texs)} textures ({tex} of {texels_total} texels): 8 BC1, 4 BC5, "
                       f"4 BC4, full Triangle chains, through the reference's FloatImage::downSample and compressors")


def _cube_face(face: int, S: int, seed: int) -> np.ndarray:
    """One 4096^2-style cube face: seeded value noise plus smooth gradients of
    the unit direction, which are continuous across face edges."""
    from ntx_amd import synth
    c = (np.arange(S, dtype=np.float32) + 0.5) / S * 2 - 1
    u, v = np.meshgrid(c, c)
    one = np.ones_like(u)
    # (+x, -x, +y, -y, +z, -z) faces, directions as in the D3D cube map layout
    d = [np.stack([one, -v, -u]), np.stack([-one, -v, u]), np.stack([u, one, v]),
         np.stack([u, -one, -v]), np.stack([u, -v, one]), np.stack([-u, -v, -one])][face]
    d = d / np.sqrt((d * d).sum(axis=0, keepdims=True))
    rng = np.random.default_rng(seed + face)
    img = np.empty((S, S, 4), np.uint8)
    axes = ((1.0, 0.3, 0.0), (0.0, 1.0, 0.5), (0.4, 0.0, 1.0))
    for ch in range(3):
        grad = 0.5 + 0.5 * (axes[ch][0] * d[0] + axes[ch][1] * d[1] + axes[ch][2] * d[2]) / np.linalg.norm(axes[ch])
        noise = synth.fbm(rng, S, S)
        img[:, :, ch] = np.floor(np.clip(0.6 * grad + 0.4 * noise, 0, 1) * 255 + 0.5).astype(np.uint8)
    img[:, :, 3] = 255
    return img


def bench_process(args, cpu: bool) -> dict:
    """nvtt::Compressor::process (the InputOptions pipeline) through the nvtt.h
    front end, timed by the C++ driver around whole calls: BGRA8 host input
    uploaded, to linear (gamma 2.2), mips in linear space, the gamma-2.2
    copies, compression, the DDS header and every level delivered to an
    OutputHandler on the host.  Host-to-device and device-to-host copies are
    inside the time.  With cpu, the C1 case through the oracle's restatement
    of the same steps on the host."""
    import struct
    import subprocess
    import tempfile

    import ntx_amd as N
    from ntx_amd import synth

    driver = os.path.join(N.LIB_DIR, "nvtt_api_test")
    out = {}
    cases = [("c1_bc1_box_1024", 1024, N.FORMAT_BC1, N.FILTER_BOX, 20),
             ("c2_bc3_kaiser_4096", 4096, N.FORMAT_BC3, N.FILTER_KAISER, 8)]
    with tempfile.TemporaryDirectory() as tmp:
        for name, S, fmt, filt, reps in cases:
            img = synth.noise_rect_rgba(S, S, seed=args.seed, transparent=(fmt == N.FORMAT_BC3))
            path = os.path.join(tmp, name + ".raw")
            with open(path, "wb") as f:
                f.write(struct.pack("<iiii", S, S, 1, 1))
                f.write(synth.to_bgra(img).tobytes())
            log(f"input pipeline: {name}")
            r = subprocess.run([driver, "process-time", path, str(fmt), str(filt), str(reps)], capture_output=True,
                               text=True, timeout=300)
            if r.returncode != 0:
                out[name] = {"error": r.stderr.strip()[-200:]}
                continue
            fields = r.stdout.split()
            ms = float(fields[fields.index("process_ms") + 1])
            nbytes = int(fields[fields.index("bytes") + 1])
            texels = sum(w * h for w, h in N.mip_sizes(S, S))
            out[name] = {"workload": f"{S}x{S} BGRA8 noise+rectangles -> InputOptions defaults (gamma 2.2 in and "
                                     f"out, {'Box' if filt == N.FILTER_BOX else 'Kaiser(3; 4, 1)'} mips to 1x1) -> "
                                     f"{'BC1' if fmt == N.FORMAT_BC1 else 'BC3'} -> DDS",
                         "ms_per_call": round(ms, 3), "mtexels_s": round(texels / (ms / 1e3) / 1e6, 1),
                         "bytes_delivered": nbytes}
            if cpu and fmt == N.FORMAT_BC1:
                sys.path.insert(0, os.path.join(ROOT, "tests"))
                import oracle_lib
                orc = oracle_lib.Oracle()
                threads = host_threads(args.cpu_threads)
                t0 = time.perf_counter()
                lin = orc.planar_to_linear(synth.to_planar(img).reshape(-1))
                for k, (w, h) in enumerate(N.mip_sizes(S, S)):
                    orc.compress(fmt, orc.planar_to_gamma(lin), w, h, threads=threads)
                    if (w, h) != (1, 1):
                        lin = orc.build_next_mipmap(filt, 0.5, None, N.WRAP_MIRROR, lin, w, h)
                cpu_s = time.perf_counter() - t0
                out[name]["cpu_port"] = {"mtexels_s": round(texels / cpu_s / 1e6, 3), "cores": threads}
    return out


def bench_c1(ctx, args, cpu: bool) -> dict:
    """The reference's CPU-runnable case: one 1024^2 RGBA8 noise+rectangles
    texture (alpha 255), its 11-level Box chain (fastDownSample), BC1 at
    Normal quality per level, through the chain entry point.  With cpu, the
    same chain on the host: the oracle port and the reference's own icbc
    (oracle/_ref), both checked byte for byte against the device output."""
    import ntx_amd as N
    from ntx_amd import synth

    S = 1024
    levels = N.mip_sizes(S, S)
    texels = sum(w * h for w, h in levels)
    img = synth.noise_rect_rgba(S, S, seed=args.seed)
    planar = synth.to_planar(img).reshape(-1)
    d_lv = [ctx.alloc(16 * w * h) for w, h in levels]
    d_lv[0].upload(planar)
    d_out = [ctx.alloc(N.compressed_size(N.FORMAT_BC1, w, h)) for w, h in levels]

    def step():
        ctx.compress_chain_device(N.FORMAT_BC1, N.FILTER_BOX, d_lv, S, S, d_out)

    for _ in range(3):
        step()
    ctx.synchronize()
    ctx.timer_begin()
    for _ in range(args.steps):
        step()
    ms = ctx.timer_end() / args.steps
    gpu_blocks = [d.download(np.uint8) for d in d_out]
    gpu_levels = [d.download(np.float32) for d in d_lv]
    res = {"workload": f"{S}x{S} RGBA8 noise+rectangles (seed {args.seed}), {len(levels)}-level Box chain, BC1 Normal "
                       f"per level (ntx_hip_compress_chain_device)",
           "texels_per_step": texels, "ms_per_step": round(ms, 4), "mtexels_s": round(texels / (ms / 1e3) / 1e6, 1)}
    # Rooflines: BC1's counted flops over its kernels' time, the Box mips'
    # 16 B per source and output texel over box_kernel's.
    flops = bc1_levels_flops(ctx, d_lv, levels)
    ctx.profile(True)
    ctx.profile_read(reset=True)
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    prof = ctx.profile_read(reset=True)
    ctx.profile(False)
    res["roofline"] = kernels_roofline(prof, ("bc1_kernel", "bc1_classify_kernel", "bc1_search_kernel"), flops,
                                       args.steps)
    box_ms = prof.get("box_kernel", (0.0, 0))[0] / args.steps
    box_bytes = sum(16 * (w * h + levels[k + 1][0] * levels[k + 1][1]) for k, (w, h) in enumerate(levels[:-1]))
    if box_ms > 0:
        res["mip_roofline"] = {"kernel": "box_kernel", "bound": "hbm", "achieved": round(box_bytes / (box_ms / 1e3) / 1e9, 1),
                               "peak": PEAK_HBM_GBS, "unit": "GB/s",
                               "frac": round(box_bytes / (box_ms / 1e3) / 1e9 / PEAK_HBM_GBS, 4),
                               "algorithmic_bytes_per_chain": box_bytes, "kernel_ms_per_chain": round(box_ms, 4)}
    res["note"] = ("1.4 M texels per chain: the levels are launch- and latency-bound (the device is far from "
                   "full below 512^2); the kernel times overlap on three streams")
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        threads = host_threads(args.cpu_threads)
        orc = oracle_lib.Oracle()
        t0 = time.perf_counter()
        x = planar
        cpu_blocks = []
        cpu_levels = [x]
        for k, (w, h) in enumerate(levels):
            cpu_blocks.append(orc.compress(N.FORMAT_BC1, cpu_levels[k], w, h, threads=threads))
            if k + 1 < len(levels):
                cpu_levels.append(orc.build_next_mipmap(N.FILTER_BOX, 0.5, None, N.WRAP_MIRROR, cpu_levels[k], w, h))
        cpu_s = time.perf_counter() - t0
        res["cpu_port"] = {"mtexels_s": round(texels / cpu_s / 1e6, 3), "cores": threads,
                           "blocks_identical": all(np.array_equal(a, b) for a, b in zip(gpu_blocks, cpu_blocks)),
                           "levels_bit_identical": all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
                                                       for a, b in zip(gpu_levels, cpu_levels))}
        ref = oracle_lib.Reference.load()
        if ref is not None:
            import ctypes
            fn = ref.lib.ntxref_compress_bc1_image
            fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
            t0 = time.perf_counter()
            ref_blocks = []
            for k, (w, h) in enumerate(levels):
                out = np.zeros(N.compressed_size(N.FORMAT_BC1, w, h), np.uint8)
                src = np.ascontiguousarray(gpu_levels[k], np.float32)
                fn(src.ctypes.data, w, h, out.ctypes.data, threads)
                ref_blocks.append(out)
            ref_s = time.perf_counter() - t0
            res["cpu_reference_icbc"] = {
                "mtexels_s": round(texels / ref_s / 1e6, 3), "cores": threads,
                "blocks_identical": all(np.array_equal(a, b) for a, b in zip(gpu_blocks, ref_blocks))}
    for d in d_lv + d_out:
        d.free()
    return res


def bench_cube(ctx, args, net, cpu: bool):
    """Cube map: 6 faces of 4096^2, full 13-level Kaiser chains per face
    (ntx_hip_downsample_chain_device), then all 78 levels compressed to BC1
    in one call each way: the plain cluster fit (ntx_hip_compress_images_device)
    and the learned proposals (ntx_hip_compress_bc1_mlp_images_device with
    learned.level_settings: the shipped settings for level 0, one least-squares
    pass, one greedy round and the tau_gap check for downsampled levels, and
    the plain fit on a side stream for levels under 4096 blocks).  Per-level
    and overall PSNR, Mtexels/s of both, and the reference's icbc on the host
    over one face's chain."""
    import ntx_amd as N
    from ntx_amd import learned, synth

    S = 4096
    levels = N.mip_sizes(S, S)
    log("cube map: generating 6 faces")
    lv = []
    for f in range(6):
        bgra = synth.to_bgra(_cube_face(f, S, args.seed + 4000))
        d_bgra = ctx.alloc(bgra.nbytes).upload(bgra)
        chain = [ctx.alloc(16 * w * h) for w, h in levels]
        ctx.unpack_bgra8_device(d_bgra, S, S, chain[0])
        d_bgra.free()
        lv.append(chain)
    ctx.synchronize()
    ctx.timer_begin()
    for chain in lv:
        ctx.downsample_chain_device(N.FILTER_KAISER, chain, S, S, 3.0, (4.0, 1.0))
    ms_mips = ctx.timer_end()
    # Every level of every face, largest first (the scheduling order the
    # multi-image entry points take).
    items = sorted(((f, k) for f in range(6) for k in range(len(levels))), key=lambda fk: (fk[1], fk[0]))
    ins = [lv[f][k] for f, k in items]
    sizes = [levels[k] for _, k in items]
    texels = sum(w * h for w, h in sizes)
    res = {}
    outs = {}
    # "learned": learned.level_settings per level (every level within 0.05 dB
    # of the plain fit on held-out chains); "learned_throughput": the shipped
    # settings on every level of 4096 blocks or more.
    fast = [dict(learned.SHIPPED) if ((w + 3) // 4) * ((h + 3) // 4) >= learned.MIN_BLOCKS else None
            for w, h in sizes]
    profs = {}
    for name in ("plain", "learned", "learned_throughput"):
        out = [ctx.alloc(N.compressed_size(N.FORMAT_BC1, w, h)) for w, h in sizes]
        outs[name] = out

        def run():
            if name == "plain":
                ctx.compress_images_device(N.FORMAT_BC1, ins, sizes, out)
            elif name == "learned":
                net.compress_bc1_images_device(ins, sizes, out, levels=[k for _, k in items])
            else:
                net.compress_bc1_images_device(ins, sizes, out, settings=fast)

        run()
        ctx.synchronize()
        reps = 3
        ctx.timer_begin()
        for _ in range(reps):
            run()
        ms = ctx.timer_end() / reps
        ctx.profile(True)
        ctx.profile_read(reset=True)
        for _ in range(reps):
            run()
        ctx.synchronize()
        profs[name] = ctx.profile_read(reset=True)
        ctx.profile(False)
        sse = [ctx.bc1_sse_device(ins[i], w, h, out[i]) for i, (w, h) in enumerate(sizes)]
        per_level = []
        for k, (w, h) in enumerate(levels):
            e = sum(sse[i] for i, (_, kk) in enumerate(items) if kk == k) / (6 * w * h)
            per_level.append(round(10 * np.log10(255.0 ** 2 / max(e, 1e-12)), 3))
        overall = 10 * np.log10(255.0 ** 2 / max(sum(sse) / texels, 1e-12))
        res[name] = {"ms": round(ms, 3), "mtexels_s": round(texels / (ms / 1e3) / 1e6, 1),
                     "psnr_db": round(float(overall), 3), "psnr_db_per_level": per_level}
        log(f"  cube {name}: {res[name]['ms']} ms, {res[name]['mtexels_s']} Mtexels/s, {overall:.3f} dB")
    out = {"workload": "6 faces x 4096^2 RGBA8 (seeded noise + direction gradients continuous across faces), "
                       "13-level Kaiser(3; 4, 1) chains per face, all 78 levels to BC1 in one call: plain cluster fit "
                       "and learned proposals (learned.level_settings: shipped settings at level 0, refined at "
                       "levels 1-4, the plain fit below 4096 blocks; cluster-fit fallback above tau_pred)",
           "texels": texels, "mip_chains_ms": round(ms_mips, 3), **res,
           "speedup": round(res["plain"]["ms"] / res["learned"]["ms"], 3),
           "psnr_delta_db": round(res["learned"]["psnr_db"] - res["plain"]["psnr_db"], 3),
           "worst_level_delta_db": round(min(a - b for a, b in zip(res["learned"]["psnr_db_per_level"],
                                                                   res["plain"]["psnr_db_per_level"])), 3),
           "throughput_speedup": round(res["plain"]["ms"] / res["learned_throughput"]["ms"], 3),
           "throughput_worst_level_delta_db": round(min(a - b for a, b in zip(
               res["learned_throughput"]["psnr_db_per_level"], res["plain"]["psnr_db_per_level"])), 3)}
    # Rooflines: the plain fit's counted flops (every block's colour counts)
    # over its kernels' time, and the network kernel's counted flops per
    # block (the shipped settings' work; the refined levels do more) over
    # bc1_mlp_kernel's time.
    flops_plain = bc1_levels_flops(ctx, ins, sizes)
    out["plain"]["roofline"] = kernels_roofline(profs["plain"], ("bc1_kernel", "bc1_classify_kernel",
                                                                  "bc1_search_kernel"), flops_plain, 3)
    for name, st in (("learned", [learned.level_settings(w, h, k) for (w, h), (_, k) in zip(sizes, items)]),
                     ("learned_throughput", fast)):
        nb = sum(((w + 3) // 4) * ((h + 3) // 4) for (w, h), c in zip(sizes, st) if c is not None)
        out[name]["roofline"] = kernels_roofline(profs[name], ("bc1_mlp_kernel",), MLP_VALU_FLOPS * nb, 3)
    if cpu:
        out["cpu_reference"] = cpu_reference_chain_bc1(args, [d.download(np.float32) for d in lv[0]], levels,
                                                       [outs["plain"][items.index((0, k))].download(np.uint8)
                                                        for k in range(len(levels))], texels)
    for chain in lv + list(outs.values()):
        for b in chain:
            b.free()
    return out


def cpu_reference_chain_bc1(args, host_levels, levels, gpu_blocks, texels_total) -> dict:
    """The reference's icbc Level8 (CompressorDXT1, compiled from the
    reference sources) over the levels of one chain on host_threads()
    workers, with its blocks against the device's plain fit."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    ref = oracle_lib.Reference.load()
    threads = host_threads(args.cpu_threads)
    info = {"unit": "Mtexels/s", "cores": threads, "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "kind": "reference"}
    if ref is None:
        return dict(info, value=None, sample="oracle/_ref not built")
    t0 = time.perf_counter()
    same = 0
    nb = 0
    for k, (w, h) in enumerate(levels):
        blk = ref.compress_image(1, 1, host_levels[k], w, h, threads=threads)
        same += int(np.sum(np.all(blk.reshape(-1, 8) == gpu_blocks[k].reshape(-1, 8), axis=1)))
        nb += blk.size // 8
    sec = time.perf_counter() - t0
    tex = sum(w * h for w, h in levels)
    return dict(info, value=round(tex / sec / 1e6, 3), seconds=round(sec, 3),
                sample=f"one face's {len(levels)}-level chain ({tex} texels of the workload's {texels_total}) through "
                       f"the reference's CompressorDXT1 on the same FP32 levels",
                blocks_identical_to_device=round(same / nb, 6))


def _decode_bc1(blk: np.ndarray, w: int, h: int) -> np.ndarray:
    """D3D10 BC1 decode of a w x h level into (3, h, w) uint8 (BlockDXT.cpp:43-98)."""
    b = blk.reshape(-1, 8)
    c0 = b[:, 0].astype(np.int32) | (b[:, 1].astype(np.int32) << 8)
    c1 = b[:, 2].astype(np.int32) | (b[:, 3].astype(np.int32) << 8)
    idx = b[:, 4:8].copy().view("<u4")[:, 0].astype(np.int64)

    def ex(c):
        r, g, bb = c >> 11, (c >> 5) & 63, c & 31
        return np.stack([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (bb << 3) | (bb >> 2)], 1)
    p0, p1 = ex(c0), ex(c1)
    four = (c0 > c1)[:, None]
    p2 = np.where(four, (2 * p0 + p1) // 3, (p0 + p1) // 2)
    p3 = np.where(four, (p0 + 2 * p1) // 3, 0)
    pal = np.stack([p0, p1, p2, p3], 1)  # (nb, 4, 3)
    sel = (idx[:, None] >> (2 * np.arange(16))[None]) & 3  # (nb, 16)
    px = np.take_along_axis(pal, sel[:, :, None].repeat(3, 2), 1)  # (nb, 16, 3)
    bw, bh = (w + 3) // 4, (h + 3) // 4
    img = px.reshape(bh, bw, 4, 4, 3).transpose(4, 0, 2, 1, 3).reshape(3, 4 * bh, 4 * bw)
    return img[:, :h, :w].astype(np.uint8)


if __name__ == "__main__":
    main()
