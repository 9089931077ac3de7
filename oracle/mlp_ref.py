"""CPU restatement of the BC1 endpoint-proposal network (test infrastructure
only: imported by tests/ and bench.py, never by the product).

The network has no counterpart in the reference; this file restates, in
numpy, exactly what the gfx950 kernels in nvidia-texture-tools_amd/csrc/
mlp.hip compute, so that their forward outputs, losses and gradients can be
checked:

  features  48 per block: R, G, B planes of the 16 texels minus the block
            mean over texels inside the image, times 1 / max(largest absolute
            deviation, 1/32); outside texels are 0; a block whose texels are
            all equal has exactly zero features.
  network   y = W3 h2 + b3, h2 = relu(W2 h1 + b2), h1 = relu(W1 x + b1);
            with rounding=True, x, the weights and the hidden activations are
            rounded to BF16 (round to nearest even) where the kernels round
            them, and sums are FP32-accurate (here float64).
  loss      per block: candidate k is diagonal k of the bounding box of the
            block's features (k = 0: all channels low -> high, k = 1..3:
            channel k-1 high -> low) plus outputs 6k..6k+5; min over the 4
            candidates of the unordered pair error |a-A|^2+|b-B|^2 (or
            swapped), plus kSpread/4 times every candidate's error, plus
            (y24 - rmse)^2.  Mean over blocks.
  gradient  backpropagation with dy, delta2, delta1 rounded to BF16 (as the
            kernels feed them to the matrix cores), summed over blocks,
            divided by n.

Pinned by tests/test_mlp_oracle.py: with rounding=False its gradient equals
central finite differences of its loss in float64.
"""
import numpy as np

K_IN, K_H, K_OUT = 48, 64, 25
W1, B1 = 0, 64 * 48
W2 = B1 + 64
B2 = W2 + 64 * 64
W3 = B2 + 64
B3 = W3 + 25 * 64
N_PARAMS = B3 + 25
SPREAD = 0.05
MIN_SPREAD = np.float32(1.0 / 32.0)
RMSE_WEIGHT = 1.0


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def bf16_from_bits(bits):
    return (np.asarray(bits, np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[W1:B1].reshape(64, 48), p[B1:W2], p[W2:B2].reshape(64, 64), p[B2:W3], p[W3:B3].reshape(25, 64),
            p[B3:N_PARAMS])


def features(planar, w, h):
    """(features float32 [n, 48] before BF16 rounding, aux float32 [n, 4])."""
    img = np.asarray(planar, np.float32).reshape(4, h, w)[:3]
    bw, bh = (w + 3) // 4, (h + 3) // 4
    n = bw * bh
    feat = np.zeros((n, 48), np.float32)
    aux = np.zeros((n, 4), np.float32)
    for b in range(n):
        bx, by = b % bw, b // bw
        c = np.zeros((3, 16), np.float32)
        inside = np.zeros(16, bool)
        for t in range(16):
            x, y = 4 * bx + (t & 3), 4 * by + (t >> 2)
            if x < w and y < h:
                inside[t] = True
                c[:, t] = img[:, y, x]
        m = c[:, inside].sum(axis=1, dtype=np.float32) / np.float32(inside.sum())
        if np.all(c[:, inside] == c[:, :1]):  # uniform: exactly zero features
            m = c[:, 0].copy()
        d = np.abs(c[:, inside] - m[:, None]).max()
        s = np.float32(1.0) / max(np.float32(d), MIN_SPREAD)
        f = np.where(inside[None, :], (c - m[:, None]) * s, 0.0)
        feat[b] = f.reshape(48)
        aux[b] = (m[0], m[1], m[2], s)
    return feat, aux


def box_diagonals(X):
    """Bounding-box corners of each block's features ([n, 48] -> lo, hi [n, 3])."""
    f = np.asarray(X, np.float64).reshape(-1, 3, 16)
    return f.min(axis=2), f.max(axis=2)


def block_loss(y, L, lo=None, hi=None):
    """Loss of one block (float64) and d loss / d y[0:25]; lo, hi are the
    block's feature bounding box (default: the origin)."""
    y = np.asarray(y, np.float64)
    L = np.asarray(L, np.float64)
    lo = np.zeros(3) if lo is None else np.asarray(lo, np.float64)
    hi = np.zeros(3) if hi is None else np.asarray(hi, np.float64)
    A, B, r = L[0:3], L[3:6], L[6]
    lk, swp, ca = [], [], []
    for k in range(4):
        flip = np.arange(3) == k - 1
        a = np.where(flip, hi, lo) + y[6 * k:6 * k + 3]
        b = np.where(flip, lo, hi) + y[6 * k + 3:6 * k + 6]
        ca.append((a, b))
        ds = np.sum((a - A) ** 2) + np.sum((b - B) ** 2)
        dw = np.sum((a - B) ** 2) + np.sum((b - A) ** 2)
        swp.append(dw < ds)
        lk.append(dw if dw < ds else ds)
    kb = int(np.argmin(lk))  # first minimum
    loss = lk[kb] + SPREAD * sum(lk) / 4
    dy = np.zeros(25)
    for k in range(4):
        coef = (1.0 if k == kb else 0.0) + SPREAD / 4
        ta, tb = (B, A) if swp[k] else (A, B)
        dy[6 * k:6 * k + 3] = coef * 2 * (ca[k][0] - ta)
        dy[6 * k + 3:6 * k + 6] = coef * 2 * (ca[k][1] - tb)
    dy[24] = RMSE_WEIGHT * 2 * (y[24] - r)
    loss += RMSE_WEIGHT * (y[24] - r) ** 2
    return loss, dy


def forward(p, X, rounding=True):
    """Outputs [n, 25] and the hidden activations for inputs X [n, 48]."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    if rounding:
        w1, w2, w3 = (bf16(w).astype(np.float64) for w in (w1, w2, w3))
        X = bf16(X).astype(np.float64)
    r = (lambda v: bf16(v.astype(np.float32)).astype(np.float64)) if rounding else (lambda v: v)
    h1 = r(np.maximum(X @ w1.T + b1, 0.0))
    h2 = r(np.maximum(h1 @ w2.T + b2, 0.0))
    y = h2 @ w3.T + b3
    if rounding:
        y = y.astype(np.float32).astype(np.float64)
    return y, h1, h2, X


def loss_and_grad(p, X, labels, rounding=True):
    """Mean loss and mean gradient (float64 [N_PARAMS]) over the blocks."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    if rounding:
        w2, w3 = bf16(w2).astype(np.float64), bf16(w3).astype(np.float64)
    y, h1, h2, Xr = forward(p, X, rounding)
    lo, hi = box_diagonals(Xr)
    n = len(y)
    r = (lambda v: bf16(v.astype(np.float32)).astype(np.float64)) if rounding else (lambda v: v)
    total = 0.0
    dy = np.zeros((n, 25))
    for i in range(n):
        l, d = block_loss(y[i], labels[i], lo[i], hi[i])
        total += l
        dy[i] = d
    dy = r(dy)
    d2 = r((dy @ w3) * (h2 > 0))
    d1 = r((d2 @ w2) * (h1 > 0))
    g = np.zeros(N_PARAMS)
    g[W1:B1] = (d1.T @ Xr).reshape(-1)
    g[B1:W2] = d1.sum(axis=0)
    g[W2:B2] = (d2.T @ h1).reshape(-1)
    g[B2:W3] = d2.sum(axis=0)
    g[W3:B3] = (dy.T @ h2).reshape(-1)
    g[B3:N_PARAMS] = dy.sum(axis=0)
    return total / n, g / n
