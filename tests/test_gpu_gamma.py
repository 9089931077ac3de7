"""GPU parity of the input pipeline's colour steps against the oracle:
ntx_hip_gamma_device / ntx_hip_gamma_levels_device (FloatImage::toLinear /
toGamma, bit-exact for gamma 2.2; for any other gamma correctly rounded, and
within one ulp of the host C library's powf), ntx_hip_unpack_bgra8_gamma_device (bit-exact for every 8-bit
value) and ntx_hip_normalize_normals_device (bit-exact)."""
import numpy as np
import pytest

import ntx_amd as N

pytestmark = pytest.mark.gpu


def _values(n, seed):
    rng = np.random.default_rng(seed)
    x = np.concatenate([
        rng.random(n, dtype=np.float32) * 1.5 - 0.25,
        np.geomspace(1e-38, 1e30, 512).astype(np.float32),
        np.array([0.0, -0.0, 1.0, -1.0, 1e-45, 1e-40, 3.4e38, np.inf, -np.inf, np.nan, 0.5, 2.2], np.float32),
    ])
    k = (x.size + 3) // 4 * 4  # a whole number of texels
    return np.concatenate([x, np.zeros(k - x.size, np.float32)])


@pytest.mark.parametrize("to_gamma", [False, True])
def test_gamma_22_bit_exact(gpu, oracle, to_gamma):
    x = _values(40000, 1 + int(to_gamma))
    n = x.size // 4
    # planar RGBA of n texels: all four planes hold test values; A must be copied
    d = gpu.alloc(x.nbytes).upload(x)
    o = gpu.alloc(x.nbytes)
    gpu.gamma_device(to_gamma, 2.2, d, n, 1, o)
    got = o.download(np.float32)
    exp = oracle.planar_to_gamma(x) if to_gamma else oracle.planar_to_linear(x)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    # in place gives the same
    gpu.gamma_device(to_gamma, 2.2, d, n, 1)
    assert np.array_equal(d.download(np.float32).view(np.uint32), exp.view(np.uint32))
    d.free(); o.free()


@pytest.mark.parametrize("gamma", [1.8, 2.4, 1.0])
def test_other_gamma_correctly_rounded(gpu, oracle, gamma):
    """Gammas other than 2.2: the correctly rounded powf (float64 power rounded
    once), and within one ulp of the host C library's powf, which the
    reference calls and which is not correctly rounded everywhere."""
    n = 50000
    x = np.random.default_rng(7).random(4 * n, dtype=np.float32) * 1.2 - 0.1
    d = gpu.alloc(x.nbytes).upload(x)
    g32 = np.float32(gamma)
    for to_gamma in (False, True):
        o = gpu.alloc(x.nbytes)
        gpu.gamma_device(to_gamma, gamma, d, n, 1, o)
        got = o.download(np.float32)
        host = oracle.planar_to_gamma(x, gamma) if to_gamma else oracle.planar_to_linear(x, gamma)
        p = float(np.float32(1) / g32) if to_gamma else float(g32)
        cr = np.power(np.maximum(x[:3 * n], 0).astype(np.float64), p).astype(np.float32)
        assert np.mean(got[:3 * n] == cr) >= 0.99999
        ulp = np.spacing(np.abs(host).astype(np.float32))
        assert np.all(np.abs(got - host) <= ulp)
        assert np.mean(got == host) >= 0.998
        assert np.array_equal(got[3 * n:], x[3 * n:])
        o.free()
    d.free()


def test_gamma_levels_one_launch(gpu, oracle):
    sizes = [(37, 23), (18, 11), (9, 5), (4, 2), (2, 1), (1, 1), (64, 64)]
    rng = np.random.default_rng(11)
    xs = [rng.random(4 * w * h, dtype=np.float32) for w, h in sizes]
    d_in = [gpu.alloc(x.nbytes).upload(x) for x in xs]
    d_out = [gpu.alloc(x.nbytes) for x in xs]
    gpu.gamma_levels_device(True, 2.2, d_in, sizes, d_out)
    for x, d in zip(xs, d_out):
        assert np.array_equal(d.download(np.float32).view(np.uint32), oracle.planar_to_gamma(x).view(np.uint32))
    for d in d_in + d_out:
        d.free()


@pytest.mark.parametrize("gamma", [2.2, 1.8, 1.00001, 1.0])
def test_unpack_gamma_every_byte(gpu, oracle, gamma):
    """setImage(BGRA8) + Surface::toLinear(gamma) in one pass, all 256 values per channel."""
    v = np.arange(256, dtype=np.uint8)
    bgra = np.stack([v, v[::-1], np.roll(v, 77), np.roll(v, 5)], axis=1).reshape(-1)  # 256 texels
    d = gpu.alloc(bgra.nbytes).upload(bgra)
    o = gpu.alloc(4 * 256 * 4)
    gpu.unpack_bgra8_gamma_device(d, 256, 1, gamma, o)
    got = o.download(np.float32)
    planar = np.concatenate([bgra[2::4], bgra[1::4], bgra[0::4], bgra[3::4]]).astype(np.float32) / np.float32(255)
    unit = abs(gamma - 1.0) <= 1e-4
    exp = planar if unit else oracle.planar_to_linear(planar, gamma)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    d.free(); o.free()


def test_normalize_normals(gpu, oracle):
    rng = np.random.default_rng(5)
    n = 4099
    x = rng.random(4 * n, dtype=np.float32)
    x[0], x[n], x[2 * n] = 127 / 255, 127 / 255, 127 / 255  # expands to the zero vector
    d = gpu.alloc(x.nbytes).upload(x)
    gpu.normalize_normals_device(d, n, 1)
    got = d.download(np.float32)
    exp = oracle.normalize_normals(x)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    assert got[0] == 0.5 and got[3 * n:].tolist() == x[3 * n:].tolist()
    d.free()
