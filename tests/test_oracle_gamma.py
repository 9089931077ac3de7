"""The oracle's gamma transforms (FloatImage::toLinear / toGamma, the
gamma-2.2 table-and-polynomial powf of Gamma.cpp:311-441) pinned against the
reference's literal exponent tables (digests in tests/golden/gamma_tables.json,
written from the reference source by tests/golden/make_gamma_fixture.py), the
error bounds Gamma.cpp states for its approximations, their special cases, and
the exact powf path of any other gamma."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gamma_tables.json")


@pytest.fixture(scope="module")
def oracle():
    return oracle_lib.Oracle()


def test_tables_match_the_reference_digests(oracle):
    ref = json.load(open(GOLDEN))
    for which, key in ((0, "pow_5_11"), (1, "pow_11_5")):
        t = oracle.gamma_table(which).astype("<f4")
        assert hashlib.sha256(t.tobytes()).hexdigest() == ref[key]["sha256"], key
        assert int(np.count_nonzero(t)) == ref[key]["nonzero"]
        assert int(np.isinf(t).sum()) == ref[key]["inf"]


def test_tables_closed_form_entries(oracle):
    t511, t115 = oracle.gamma_table(0), oracle.gamma_table(1)
    assert t511[127] == 1.0 and t115[127] == 1.0          # 2^0
    assert t511[127 + 11] == np.float32(32.0)              # 2^(11 * 5/11)
    assert t115[127 + 5] == np.float32(2.0 ** 11)          # 2^(5 * 11/5)
    assert t511[0] == 0 and t115[0] == 0 and np.isinf(t511[255]) and np.isinf(t115[255])
    assert not np.any(t511[256:]) and not np.any(t115[256:])


@pytest.mark.parametrize("which,p,bound", [(0, 5.0 / 11.0, 1.2e-5), (1, 11.0 / 5.0, 2.9e-6)])
def test_relative_error_bounds(oracle, which, p, bound):
    """Gamma.cpp:300-340 state rel |error| < 1.2e-5 (5/11) and < 2.9e-6 (11/5);
    a wrong coefficient, exponent or table entry breaks them."""
    x = np.concatenate([np.geomspace(1e-15, 1e15, 20001), np.linspace(0.0005, 1.0, 20001)]).astype(np.float32)
    f = oracle.lib.ntxo_powf_5_11 if which == 0 else oracle.lib.ntxo_powf_11_5
    got = np.array([f(float(v)) for v in x], np.float64)
    exact = np.power(x.astype(np.float64), p)
    rel = np.abs(got - exact) / exact
    assert rel.max() < bound, (rel.max(), x[np.argmax(rel)])


def test_special_values(oracle):
    for f in (oracle.lib.ntxo_powf_5_11, oracle.lib.ntxo_powf_11_5):
        assert f(0.0) == 0.0
        assert f(1e-40) == 0.0                 # subnormal
        assert f(-0.5) == 0.0 and f(-3.0) == 0.0
        assert np.isinf(f(float("inf"))) and np.isinf(f(float("nan")))
        assert f(1.0) == pytest.approx(1.0, rel=1e-5)


def test_round_trip_of_8bit_values(oracle):
    v = np.arange(256, dtype=np.float32) / np.float32(255)
    back = oracle.to_gamma(oracle.to_linear(v))
    assert np.allclose(back, v, rtol=3e-5, atol=1e-7)


def test_other_gammas_are_exponentiation(oracle):
    x = np.array([-1.0, -0.0, 0.0, 1e-3, 0.25, 0.5, 1.0, 1.7], np.float32)
    for g in (1.8, 2.4, 1.0):
        g32 = float(np.float32(g))  # the exponent as the float the transform receives
        lin = oracle.to_linear(x, g)
        assert np.allclose(lin, np.power(np.maximum(x, 0).astype(np.float64), g32), rtol=2e-7, atol=0)
        gam = oracle.to_gamma(x, g)
        assert np.allclose(gam, np.power(np.maximum(x, 0).astype(np.float64), np.float32(1.0) / np.float32(g)),
                           rtol=2e-7, atol=0)


def test_planar_transform_keeps_alpha(oracle):
    rng = np.random.default_rng(3)
    p = rng.random(4 * 37, dtype=np.float32)
    out = oracle.planar_to_linear(p)
    assert np.array_equal(out[3 * 37:], p[3 * 37:])
    assert np.array_equal(out[:3 * 37], oracle.to_linear(p[:3 * 37]))


def test_normal_renormalisation(oracle):
    """expandNormals (2x - 254/255), normalizeSafe, packNormals (x/2 + 1/2):
    the expanded vectors come out of unit length, the zero vector stays zero
    (packed 0.5), alpha is untouched, and the result matches a float64
    evaluation of the same steps to float rounding."""
    rng = np.random.default_rng(9)
    n = 5000
    p = rng.random(4 * n, dtype=np.float32)
    p[0] = p[n] = p[2 * n] = np.float32(127) / np.float32(255)
    out = oracle.normalize_normals(p)
    assert out[0] == out[n] == out[2 * n] == 0.5
    assert np.array_equal(out[3 * n:], p[3 * n:])
    v = (2.0 * p[:3 * n].reshape(3, n).astype(np.float64) - 254.0 / 255.0)
    u = v / np.linalg.norm(v, axis=0)
    exp = (0.5 * u + 0.5)[:, 1:]
    got = out[:3 * n].reshape(3, n)[:, 1:]
    assert np.allclose(got, exp, rtol=0, atol=3e-7)
    e = 2.0 * got.astype(np.float64) - 1.0
    assert np.allclose(np.linalg.norm(e, axis=0), 1.0, atol=2e-6)
