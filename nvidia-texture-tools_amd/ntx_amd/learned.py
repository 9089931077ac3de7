"""Training and evaluation of the learned BC1 endpoint proposals
(ntx_hip_mlp_* in include/ntx_hip.h).

Data: 4x4 blocks cut from seeded synthetic textures; each block's label is
the result of the BC1 cluster fit (icbc Level8, the same encoder as
ntx_hip_compress) computed on the device.  Training: mini-batch Adam with the
device's own forward/backward kernels.  Evaluation: PSNR of BC1 encodings as
the reference's bc1enc measures it (sum of squared R, G, B errors in 8-bit
units after D3D10 decoding, over the texel count).
"""
from __future__ import annotations

import math
from typing import Iterable, List, Optional, Tuple

import numpy as np

from . import FORMAT_BC1, Context, Mlp, mlp_features_device, mlp_labels_device

FEATURE_BYTES = 48 * 2
LABEL_BYTES = 8 * 4

# Shipped settings of the proposal form (Mlp.compress_bc1_device), fixed on a
# held-out 8192^2 texture (16 tiles of 4 seeded 2048^2 textures, seeds
# 5000-5003, none of them in training or evaluation):
#  * tau_pred = 8 (8-bit RMSE units): blocks predicted worse go to the full
#    cluster fit (0.6% of that texture's blocks; +0.11 dB there);
#  * no least-squares pass: one pass on the network's indices lowers PSNR by
#    0.15 dB there once the endpoints are snapped to 5:6:5, and costs 35%
#    more time; with a greedy 5:6:5 round after it, quality rises by 0.4 dB
#    but the form runs at 2.2x the plain fit instead of 3.2x;
#  * no greedy round, tau_gap off.
SHIPPED = dict(tau_pred=8.0, tau_gap=-1.0, refine_iters=0, greedy_rounds=0)

# Downsampled mip levels are smoother than the textures the proposals do
# best on: there the shipped form loses to the cluster fit (on a held-out
# 4096^2 noise+rectangle chain, -0.4 dB at level 1 and -0.75 dB at level 2).
# One least-squares pass, one greedy 5:6:5 round and the tau_gap check (a
# block whose achieved RMSE exceeds the predicted cluster-fit RMSE by 1.0
# goes to the fit) keep levels 1-5 of that chain and of a cube face's chain
# within 0.04 dB of the fit at 0.4-0.45 of the plain fit's time (without
# the gap check two levels fall 0.05-0.11 dB short; with a second
# least-squares pass instead, the same quality, slower on the noise chain).
# Below 4096 blocks (256 x 256 texels) a level's time is launch latency, and
# the plain fit is used.
REFINED = dict(tau_pred=8.0, tau_gap=1.0, refine_iters=1, greedy_rounds=1)
MIN_BLOCKS = 1 << 12


def level_settings(w: int, h: int, level: int = 0) -> Optional[dict]:
    """The proposal settings for mip level `level` (w x h) of a chain:
    SHIPPED for level 0, REFINED for downsampled levels, None (the plain
    cluster fit) for images under MIN_BLOCKS blocks."""
    if ((w + 3) // 4) * ((h + 3) // 4) < MIN_BLOCKS:
        return None
    return dict(SHIPPED) if level == 0 else dict(REFINED)


def decode_bc1(blocks: np.ndarray, w: int, h: int) -> np.ndarray:
    """D3D10 BC1 decode (BlockDXT.cpp:43-98) of a w x h level into (3, h, w) uint8."""
    b = np.asarray(blocks, np.uint8).reshape(-1, 8)
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
    pal = np.stack([p0, p1, p2, p3], 1)
    sel = (idx[:, None] >> (2 * np.arange(16))[None]) & 3
    px = np.take_along_axis(pal, sel[:, :, None].repeat(3, 2), 1)
    bw, bh = (w + 3) // 4, (h + 3) // 4
    img = px.reshape(bh, bw, 4, 4, 3).transpose(4, 0, 2, 1, 3).reshape(3, 4 * bh, 4 * bw)
    return img[:, :h, :w].astype(np.uint8)


def bc1_psnr(blocks: np.ndarray, planar: np.ndarray, w: int, h: int) -> float:
    """PSNR (dB) of a BC1 level against its FP32 source (bc1enc.cpp:73-81, 816-820)."""
    dec = decode_bc1(blocks, w, h).astype(np.float64)
    src = np.asarray(planar, np.float32).reshape(4, h, w)[:3].astype(np.float64) * 255.0
    mse = np.sum((dec - src) ** 2) / (w * h)
    return float("inf") if mse == 0 else 20.0 * math.log10(255.0 / math.sqrt(mse))


def block_data(ctx: Context, planar: np.ndarray, w: int, h: int) -> Tuple[np.ndarray, np.ndarray]:
    """Features (uint16 BF16 bits [n, 48]) and cluster-fit labels (float32 [n, 8])
    of every block of one texture, computed on the device."""
    n = ((w + 3) // 4) * ((h + 3) // 4)
    src = np.ascontiguousarray(planar, np.float32).reshape(-1)
    d = ctx.alloc(src.nbytes).upload(src)
    d_bc1, d_f, d_a, d_l = ctx.alloc(8 * n), ctx.alloc(FEATURE_BYTES * n), ctx.alloc(16 * n), ctx.alloc(LABEL_BYTES * n)
    ctx.compress_device(FORMAT_BC1, d, w, h, d_bc1)
    mlp_features_device(ctx, d, w, h, d_f, d_a)
    mlp_labels_device(ctx, d, w, h, d_bc1, d_a, d_l)
    f = d_f.download(np.uint16, 48 * n).reshape(n, 48)
    lab = d_l.download(np.float32, 8 * n).reshape(n, 8)
    for buf in (d, d_bc1, d_f, d_a, d_l):
        buf.free()
    return f, lab


class Dataset:
    """Shuffled features and labels resident on the device."""

    def __init__(self, ctx: Context, feats: np.ndarray, labels: np.ndarray, seed: int = 1):
        perm = np.random.default_rng(seed).permutation(len(feats))
        f = np.ascontiguousarray(feats[perm])
        lab = np.ascontiguousarray(labels[perm])
        self.n = len(f)
        self.d_feat = ctx.alloc(f.nbytes).upload(f)
        self.d_lab = ctx.alloc(lab.nbytes).upload(lab)

    def feat(self, i: int) -> int:
        return self.d_feat.ptr + FEATURE_BYTES * i

    def label(self, i: int) -> int:
        return self.d_lab.ptr + LABEL_BYTES * i

    def free(self) -> None:
        self.d_feat.free()
        self.d_lab.free()


def make_dataset(ctx: Context, textures: Iterable[Tuple[np.ndarray, int, int]], seed: int = 1) -> Dataset:
    """Blocks of the textures, without uniform blocks (all features zero):
    the compressor gives those the cluster fit's single-colour encoding, and
    their cluster-fit endpoints (chosen from the single-colour tables) are not
    a function the network can or needs to learn."""
    fs, ls = [], []
    for planar, w, h in textures:
        f, lab = block_data(ctx, planar, w, h)
        keep = np.any(f != 0, axis=1)
        fs.append(f[keep])
        ls.append(lab[keep])
    return Dataset(ctx, np.concatenate(fs), np.concatenate(ls), seed)


def train(mlp: Mlp, data: Dataset, epochs: int = 2, batch: int = 16384, lr: float = 3e-3,
          lr_end: Optional[float] = None, log=None) -> List[float]:
    """Adam over the dataset in order (it is shuffled once); cosine decay from
    lr to lr_end.  Returns the mean loss of each epoch."""
    lr_end = lr * 0.05 if lr_end is None else lr_end
    steps = epochs * ((data.n + batch - 1) // batch)
    out, k = [], 0
    for e in range(epochs):
        tot, cnt = 0.0, 0
        for i in range(0, data.n, batch):
            nb = min(batch, data.n - i)
            cur = lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * k / max(1, steps - 1)))
            loss, _ = mlp.train_step(data.feat(i), data.label(i), nb, cur)
            tot += loss * nb
            cnt += nb
            k += 1
        out.append(tot / cnt)
        if log:
            log(f"epoch {e}: mean loss {out[-1]:.5f}")
    return out
