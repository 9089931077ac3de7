"""CPU tests pinning the network restatement (oracle/mlp_ref.py) that the
GPU training kernels are checked against: its backward pass against central
finite differences of its own loss, the loss's symmetry in each endpoint
pair, and the BF16 rounding against hand-worked values."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref as R  # noqa: E402


def _problem(seed, n=12):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, R.N_PARAMS)
    X = rng.uniform(-1, 1, (n, 48))
    L = rng.uniform(-1, 1, (n, 8))
    L[:, 6] = np.abs(L[:, 6])
    return p, X, L


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gradient_matches_finite_differences(seed):
    p, X, L = _problem(seed)
    _, g = R.loss_and_grad(p, X, L, rounding=False)
    rng = np.random.default_rng(100 + seed)
    idx = np.concatenate([rng.choice(np.arange(R.W1, R.B1), 8), rng.choice(np.arange(R.B1, R.W2), 4),
                          rng.choice(np.arange(R.W2, R.B2), 8), rng.choice(np.arange(R.B2, R.W3), 4),
                          rng.choice(np.arange(R.W3, R.B3), 8), np.arange(R.B3, R.N_PARAMS)])
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp, _ = R.loss_and_grad(q, X, L, rounding=False)
        q[i] -= 2 * eps
        lm, _ = R.loss_and_grad(q, X, L, rounding=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_loss_is_symmetric_in_each_pair():
    rng = np.random.default_rng(7)
    y = rng.uniform(-1, 1, 25)
    L = rng.uniform(-1, 1, 8)
    l1, _ = R.block_loss(y, L)
    ys = y.copy()
    for k in range(4):  # swap a and b of every candidate
        ys[6 * k:6 * k + 3], ys[6 * k + 3:6 * k + 6] = y[6 * k + 3:6 * k + 6], y[6 * k:6 * k + 3]
    l2, _ = R.block_loss(ys, L)
    Ls = L.copy()
    Ls[0:3], Ls[3:6] = L[3:6], L[0:3]
    l3, _ = R.block_loss(y, Ls)
    assert l1 == pytest.approx(l2) and l1 == pytest.approx(l3)


def test_perfect_candidate_has_only_spread_and_rmse_loss():
    L = np.array([0.1, -0.2, 0.3, -0.4, 0.5, -0.6, 0.25, 0])
    y = np.zeros(25)
    y[6:12] = L[0:6]  # candidate 1 exact
    y[24] = 0.25
    l, dy = R.block_loss(y, L)
    others = sum(np.sum((y[6 * k:6 * k + 3] - L[0:3]) ** 2) + np.sum((y[6 * k + 3:6 * k + 6] - L[3:6]) ** 2)
                 for k in (0, 2, 3))
    assert l == pytest.approx(R.SPREAD * others / 4)
    assert np.all(dy[6:12] == 0) and dy[24] == 0


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0, np.inf], np.float32)
    got = R.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1.0); 1.005859375 rounds up.
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == np.float32(1.0078125)
    assert got[5] == np.inf
    assert np.all(np.abs(got[:5] - vals[:5]) <= np.abs(vals[:5]) * 2.0 ** -8)


def test_features_are_centred_and_scaled():
    rng = np.random.default_rng(3)
    w, h = 10, 7  # partial blocks on both edges
    planar = rng.uniform(0, 1, (4, h, w)).astype(np.float32)
    f, aux = R.features(planar, w, h)
    assert f.shape == (6, 48)  # 3 x 2 blocks
    assert np.all(np.abs(f) <= 1.0 + 1e-6)
    assert np.allclose(np.max(np.abs(f), axis=1), 1.0, atol=1e-6)
    # the last block holds texels (8..9, 4..6): 6 inside
    inside = [t for t in range(16) if 8 + (t & 3) < w and 4 + (t >> 2) < h]
    assert len(inside) == 6
    blk = f[5].reshape(3, 16)
    assert np.allclose(blk[:, inside].sum(axis=1), 0.0, atol=1e-5)
    assert np.all(blk[:, [t for t in range(16) if t not in inside]] == 0)
