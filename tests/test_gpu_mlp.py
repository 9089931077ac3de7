"""GPU tests of the learned BC1 endpoint proposals: the device network's
forward outputs and training gradients against the CPU restatement
(oracle/mlp_ref.py, itself pinned by finite differences), training progress,
and the compressor's contract with the cluster fit it falls back to."""
import os
import sys

import numpy as np
import pytest

import ntx_amd as N
from ntx_amd import learned, synth

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref as R  # noqa: E402

pytestmark = pytest.mark.gpu


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(R.N_PARAMS, np.float32)
    p[R.W1:R.B1] = rng.uniform(-1, 1, R.B1 - R.W1) * np.sqrt(6 / 48)
    p[R.B1:R.W2] = rng.normal(0, 0.1, 64)
    p[R.W2:R.B2] = rng.uniform(-1, 1, R.B2 - R.W2) * np.sqrt(6 / 64)
    p[R.B2:R.W3] = rng.normal(0, 0.1, 64)
    p[R.W3:R.B3] = rng.uniform(-1, 1, R.B3 - R.W3) * np.sqrt(6 / 64)
    p[R.B3:] = rng.normal(0, 0.1, 25)
    return p


def _texture(w, h, seed):
    return synth.to_planar(synth.noise_rect_rgba(w, h, seed=seed, rects=8)).reshape(-1)


def test_features_and_labels(gpu):
    w, h = 37, 21
    planar = _texture(w, h, 3)
    f, lab = learned.block_data(gpu, planar, w, h)
    fr, aux = R.features(planar, w, h)
    assert np.allclose(R.bf16_from_bits(f), R.bf16(fr), atol=2.0 ** -7)
    # labels: the cluster-fit endpoints and RMSE in the block frame
    blocks = gpu.compress(N.FORMAT_BC1, planar, w, h).reshape(-1, 8)
    dec = learned.decode_bc1(blocks, w, h).astype(np.float64) / 255.0
    src = planar.reshape(4, h, w)[:3].astype(np.float64)
    bw = (w + 3) // 4
    for b in range(len(blocks)):
        bx, by = b % bw, b // bw
        c0 = int(blocks[b, 0]) | int(blocks[b, 1]) << 8
        c1 = int(blocks[b, 2]) | int(blocks[b, 3]) << 8
        ends = []
        for c in (c0, c1):
            r5, g6, b5 = c >> 11, (c >> 5) & 63, c & 31
            ends.append(np.array([(r5 << 3 | r5 >> 2), (g6 << 2 | g6 >> 4), (b5 << 3 | b5 >> 2)]) / 255.0)
        m, s = aux[b, :3], aux[b, 3]
        assert np.allclose(lab[b, 0:3], (ends[0] - m) * s, atol=1e-4)
        assert np.allclose(lab[b, 3:6], (ends[1] - m) * s, atol=1e-4)
        d = dec[:, 4 * by:4 * by + 4, 4 * bx:4 * bx + 4] - src[:, 4 * by:4 * by + 4, 4 * bx:4 * bx + 4]
        assert lab[b, 6] == pytest.approx(np.sqrt(np.mean(d ** 2)) * s, rel=1e-4, abs=1e-6)


def test_forward_matches_oracle(gpu):
    rng = np.random.default_rng(5)
    n = 100
    X = R.bf16(rng.uniform(-1, 1, (n, 48)).astype(np.float32))
    bits = (X.view(np.uint32) >> 16).astype(np.uint16)
    p = _random_params(11)
    d = gpu.alloc(bits.nbytes).upload(bits)
    with N.Mlp(gpu, p) as net:
        y = net.forward(d.ptr, n)[:, :25]
    ref, _, _, _ = R.forward(p, X, rounding=True)
    assert np.max(np.abs(y - ref)) <= 1e-2 * (1 + np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [37, 300, 2048])
def test_gradient_matches_oracle(gpu, n):
    rng = np.random.default_rng(n)
    X = R.bf16(rng.uniform(-1, 1, (n, 48)).astype(np.float32))
    bits = (X.view(np.uint32) >> 16).astype(np.uint16)
    L = rng.uniform(-1, 1, (n, 8)).astype(np.float32)
    L[:, 6] = np.abs(L[:, 6])
    p = _random_params(n + 1)
    df, dl = gpu.alloc(bits.nbytes).upload(bits), gpu.alloc(L.nbytes).upload(L)
    with N.Mlp(gpu, p) as net:
        loss, g = net.train_step(df.ptr, dl.ptr, n, 0.0, want_grad=True)
        assert np.array_equal(net.params(), p)  # lr = 0: no update
    ref_loss, ref_g = R.loss_and_grad(p, X, L, rounding=True)
    assert loss == pytest.approx(ref_loss, rel=1e-3)
    err = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    assert err < 1e-2, err
    for lo, hi in ((R.W1, R.B1), (R.B1, R.W2), (R.W2, R.B2), (R.B2, R.W3), (R.W3, R.B3), (R.B3, R.N_PARAMS)):
        e = np.linalg.norm(g[lo:hi] - ref_g[lo:hi]) / max(np.linalg.norm(ref_g[lo:hi]), 1e-12)
        assert e < 2e-2, (lo, hi, e)


def test_training_reduces_loss(gpu):
    data = learned.make_dataset(gpu, [(_texture(256, 256, s), 256, 256) for s in (1, 2)])
    with N.Mlp(gpu, seed=3) as net:
        first, _ = net.train_step(data.feat(0), data.label(0), 4096, 0.0)
        losses = learned.train(net, data, epochs=12, batch=1024, lr=3e-3)
        last, _ = net.train_step(data.feat(0), data.label(0), 4096, 0.0)
    assert last < 0.7 * first, (first, losses, last)
    assert losses[-1] < losses[0]


def test_full_fallback_is_the_cluster_fit(gpu):
    """tau_pred < 0 sends every non-uniform block to the cluster fit, and
    uniform blocks take its single-colour encoding: the output is
    byte-identical to the per-image BC1 encoder (hence to the reference's
    icbc)."""
    w, h = 70, 45
    planar = _texture(w, h, 9)
    d = gpu.alloc(planar.nbytes).upload(planar)
    out = gpu.alloc(N.compressed_size(N.FORMAT_BC1, w, h))
    with N.Mlp(gpu, seed=1) as net:
        fb = net.compress_bc1_device(d, w, h, out, tau_pred=-1.0)
    feats, _ = R.features(planar, w, h)
    assert fb == int(np.sum(np.any(feats != 0, axis=1)))
    assert np.array_equal(out.download(np.uint8), gpu.compress(N.FORMAT_BC1, planar, w, h))


def test_proposals_decode_to_valid_blocks(gpu):
    """No fallback: every block is a 4-colour block (c0 > c1) or the c0 == c1
    case, and its error is finite; the box-candidate mode likewise."""
    w, h = 64, 64
    planar = _texture(w, h, 4)
    d = gpu.alloc(planar.nbytes).upload(planar)
    out = gpu.alloc(N.compressed_size(N.FORMAT_BC1, w, h))
    with N.Mlp(gpu, seed=2) as net:
        for box in (False, True):
            fb = net.compress_bc1_device(d, w, h, out, tau_pred=1e9, box_candidates=box)
            assert fb == 0
            blk = out.download(np.uint8).reshape(-1, 8)
            c0 = blk[:, 0].astype(int) | blk[:, 1].astype(int) << 8
            c1 = blk[:, 2].astype(int) | blk[:, 3].astype(int) << 8
            assert np.all(c0 >= c1)
            assert np.isfinite(learned.bc1_psnr(blk, planar, w, h))


def test_uniform_blocks_match_cluster_fit(gpu):
    """Blocks of one colour take the cluster fit's single-colour encoding
    byte for byte, whatever the network proposes (rectangles of flat colour
    and partial edge blocks included)."""
    w, h = 90, 50
    img = synth.noise_rect_rgba(w, h, seed=12, rects=24)
    img[8:40, 4:60, :3] = (200, 17, 96)
    planar = synth.to_planar(img).reshape(-1)
    d = gpu.alloc(planar.nbytes).upload(planar)
    out = gpu.alloc(N.compressed_size(N.FORMAT_BC1, w, h))
    ref = gpu.compress(N.FORMAT_BC1, planar, w, h).reshape(-1, 8)
    with N.Mlp(gpu, seed=5) as net:
        net.compress_bc1_device(d, w, h, out, tau_pred=1e9)
    got = out.download(np.uint8).reshape(-1, 8)
    p3 = planar.reshape(4, h, w)[:3]
    bw = (w + 3) // 4
    uniform = 0
    for b in range(len(ref)):
        bx, by = b % bw, b // bw
        blk = p3[:, 4 * by:4 * by + 4, 4 * bx:4 * bx + 4].reshape(3, -1)
        if np.all(blk == blk[:, :1]):
            uniform += 1
            assert np.array_equal(got[b], ref[b]), b
    assert uniform > 20


def test_training_improves_the_proposals(gpu):
    """What the trained network gains over the method without it.  Each of a
    block's four lanes encodes one bounding-box diagonal both as it is and
    moved by the network's offsets and keeps the better, so an untrained
    network (zero output layer) gives byte for byte the box-candidate method
    and a trained one can only add.  Measured with the candidates used as
    they are (no least-squares pass, no greedy search) on a held-out
    texture; the benchmark reports the refined settings."""
    data = learned.make_dataset(gpu, [(_texture(512, 512, s), 512, 512) for s in (21, 22, 23)])
    w = h = 256
    planar = _texture(w, h, 40)
    d = gpu.alloc(planar.nbytes).upload(planar)
    out = gpu.alloc(N.compressed_size(N.FORMAT_BC1, w, h))

    def blocks(net, **kw):
        net.compress_bc1_device(d, w, h, out, tau_pred=1e9, refine_iters=0, greedy_rounds=0, **kw)
        return out.download(np.uint8)
    with N.Mlp(gpu, seed=4) as net:
        untrained = blocks(net)
        box = blocks(net, box_candidates=True)
        assert np.array_equal(untrained, box)
        learned.train(net, data, epochs=10, batch=4096, lr=3e-3)
        trained = blocks(net)
    p_box = learned.bc1_psnr(box, planar, w, h)
    p_trained = learned.bc1_psnr(trained, planar, w, h)
    p_ref = learned.bc1_psnr(gpu.compress(N.FORMAT_BC1, planar, w, h), planar, w, h)
    print(f"psnr: box diagonals {p_box:.3f}, trained network {p_trained:.3f}, cluster fit {p_ref:.3f}")
    assert p_trained > p_box


def test_device_psnr_matches_host(gpu):
    w, h = 123, 77
    planar = _texture(w, h, 31)
    d = gpu.alloc(planar.nbytes).upload(planar)
    blocks = gpu.compress(N.FORMAT_BC1, planar, w, h)
    db = gpu.alloc(blocks.nbytes).upload(blocks)
    assert gpu.bc1_psnr_device(d, w, h, db) == pytest.approx(learned.bc1_psnr(blocks, planar, w, h), abs=1e-9)


def test_shipped_settings_contract(gpu):
    """The learned form at its shipped settings (learned.SHIPPED: fallback
    above tau_pred, no least-squares pass, no greedy round) on a held-out
    texture, after training on other seeds: PSNR within 0.05 dB of the
    cluster fit; the trained network at least MARGIN_DB above the box
    diagonals it moves (same candidates, no network); and the fallback is the
    network's own prediction: every block whose predicted cluster-fit RMSE
    is above tau_pred gets exactly the cluster fit's bytes, and the number of
    fallback blocks is the number of such predictions."""
    MARGIN_DB = 0.5
    data = learned.make_dataset(gpu, [(_texture(512, 512, s), 512, 512) for s in (51, 52, 53, 54)])
    w = h = 512
    planar = _texture(w, h, 60)
    n = (w // 4) * (h // 4)
    d = gpu.alloc(planar.nbytes).upload(planar)
    out = gpu.alloc(N.compressed_size(N.FORMAT_BC1, w, h))
    plain = gpu.compress(N.FORMAT_BC1, planar, w, h)
    kw = dict(learned.SHIPPED)
    with N.Mlp(gpu, seed=7) as net:
        learned.train(net, data, epochs=10, batch=4096, lr=3e-3)
        fb = net.compress_bc1_device(d, w, h, out, **kw)
        shipped = out.download(np.uint8)
        net.compress_bc1_device(d, w, h, out, tau_pred=1e9, refine_iters=0, greedy_rounds=0)
        trained = out.download(np.uint8)
        net.compress_bc1_device(d, w, h, out, tau_pred=1e9, refine_iters=0, greedy_rounds=0, box_candidates=True)
        box = out.download(np.uint8)
        # The network's predicted cluster-fit RMSE of every block (8-bit units).
        d_f, d_a = gpu.alloc(learned.FEATURE_BYTES * n), gpu.alloc(16 * n)
        N.mlp_features_device(gpu, d, w, h, d_f, d_a)
        y = net.forward(d_f.ptr, n)
        aux = d_a.download(np.float32).reshape(n, 4)
        feats = d_f.download(np.uint16).reshape(n, 48)
    pred = np.maximum(y[:, 24], 0.0) / aux[:, 3] * 255.0
    uniform = ~np.any(feats != 0, axis=1)
    p_plain = learned.bc1_psnr(plain, planar, w, h)
    p_shipped = learned.bc1_psnr(shipped, planar, w, h)
    p_trained = learned.bc1_psnr(trained, planar, w, h)
    p_box = learned.bc1_psnr(box, planar, w, h)
    tau = kw["tau_pred"]
    above = (pred > tau * 1.001) & ~uniform
    near = (np.abs(pred - tau) <= tau * 0.001) & ~uniform
    print(f"psnr: plain {p_plain:.3f}, shipped {p_shipped:.3f} (fallback {fb} blocks, {int(above.sum())} predicted "
          f"above tau), trained {p_trained:.3f}, box diagonals {p_box:.3f}")
    assert p_shipped >= p_plain - 0.05
    assert p_trained >= p_box + MARGIN_DB
    assert above.sum() > 0
    sb, pb = shipped.reshape(-1, 8), plain.reshape(-1, 8)
    assert np.array_equal(sb[above], pb[above])
    assert int(above.sum()) <= fb <= int(above.sum() + near.sum())


def test_images_entry_matches_per_image_calls(gpu):
    """ntx_hip_compress_bc1_mlp_images_device on a whole mip chain with the
    per-level policy (learned.level_settings) and with explicit settings:
    every level equals the single-image learned call with its settings, or
    compress_device where the settings ask for the plain fit, byte for byte."""
    w, h = 1024, 512
    sizes = N.mip_sizes(w, h)
    d_lv = [gpu.alloc(16 * a * b) for a, b in sizes]
    d_lv[0].upload(_texture(w, h, 21))
    gpu.downsample_chain_device(N.FILTER_KAISER, d_lv, w, h, 3.0, (4.0, 1.0))
    policy = [learned.level_settings(a, b, k) for k, (a, b) in enumerate(sizes)]
    assert policy[0] == learned.SHIPPED and policy[1] == learned.REFINED and policy[-1] is None
    explicit = [dict(tau_pred=6.0, tau_gap=1.0, refine_iters=2, greedy_rounds=0), None,
                dict(learned.SHIPPED), None] + [None] * (len(sizes) - 4)
    with N.Mlp(gpu, params=_random_params(5)) as net:
        for settings in (None, explicit):
            outs = [gpu.alloc(N.compressed_size(N.FORMAT_BC1, a, b)) for a, b in sizes]
            net.compress_bc1_images_device(d_lv, sizes, outs, settings=settings, levels=list(range(len(sizes))))
            got = [o.download(np.uint8) for o in outs]
            used = policy if settings is None else settings
            for k, (a, b) in enumerate(sizes):
                ref = gpu.alloc(N.compressed_size(N.FORMAT_BC1, a, b))
                if used[k] is not None:
                    net.compress_bc1_device(d_lv[k], a, b, ref, count_fallback=False, **used[k])
                else:
                    gpu.compress_device(N.FORMAT_BC1, d_lv[k], a, b, ref)
                assert np.array_equal(got[k], ref.download(np.uint8)), (a, b, used[k])
                ref.free()
            for d in outs:
                d.free()
    for d in d_lv:
        d.free()
