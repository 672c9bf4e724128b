"""Scalar pieces of the oracle against the engine's recorded outputs and the mathematics.

- hllc() and vel(): bit for bit against the reference's own outputs on 4000
  seeded states each (tests/golden/hllc_probe, vel_probe), covering dry, near-dry
  and both dry thresholds the suites use.
- Math.hypot: bit for bit against the engine on 8000 pairs.
- Math.pow(x, 4/3): within 1 ulp of the engine and within 0.9 ulp of the exact
  value (computed with Python integers/fractions), the documented accuracy of the
  fdlibm algorithm the oracle restates.
- Closed forms: HLLC consistency F(U,U) = f(U); the dry-bed Toro speeds; the
  MC/minmod/central limiters; vel = hu/h for deep water and 0 at h = 0.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import goldens
import oracle

G = 9.80665


def test_hllc_bit_exact_against_reference():
    fx = goldens.load("hllc_probe")
    c = goldens.arr(fx["cases"]).reshape(-1, 10)
    g = goldens.hexf(fx["G"])
    for row in c:
        out = oracle.hllc(*row[:7], g=g)
        assert np.array_equal(out.view(np.uint64), row[7:].view(np.uint64)), row


def test_vel_bit_exact_against_reference():
    fx = goldens.load("vel_probe")
    eps4 = goldens.hexf(fx["eps4"])
    c = goldens.arr(fx["cases"]).reshape(-1, 3)
    for h, hq, want in c:
        assert oracle.vel(hq, h, eps4) == want


def test_eps4_matches_engine():
    fx = goldens.load("vel_probe")
    assert oracle.js_pow(1e-3, 4.0) == goldens.hexf(fx["eps4"])


def test_hypot_bit_exact_against_engine():
    c = goldens.arr(goldens.load("math_probe")["cases"]).reshape(-1, 5)
    for x, p, a, b, hy in c:
        assert oracle.js_hypot(a, b) == hy


def _ulp_err(x, y, got):
    """|got - x**y| in ulps of the exact value, with x**y from a 200-bit computation."""
    import decimal
    decimal.getcontext().prec = 60
    exact = decimal.Decimal(x) ** decimal.Decimal(y)
    ulp = Fraction(math.ulp(float(exact)))
    return abs(Fraction(got) - Fraction(exact)) / ulp


def test_pow_within_one_ulp_of_engine_and_accurate():
    c = goldens.arr(goldens.load("math_probe")["cases"]).reshape(-1, 5)
    y = 4.0 / 3.0
    ndiff = 0
    for x, p, *_ in c[:3000]:
        got = oracle.js_pow(x, y)
        d = abs(int(np.float64(got).view(np.int64)) - int(np.float64(p).view(np.int64)))
        assert d <= 1
        ndiff += d
        assert _ulp_err(x, y, got) < 0.9
    assert ndiff < 0.1 * 3000


def test_hllc_consistency_and_dry_speeds():
    rng = np.random.default_rng(1)
    for _ in range(200):
        h = 0.01 + 20 * rng.random()
        u, v = rng.normal(0, 2, 2)
        out = oracle.hllc(h, h * u, h * v, h, h * u, h * v, 1e-3, G)
        # F(U,U) = f(U) up to a few roundings of the HLL average
        f = np.array([h * u, h * u * (h * u / h) + 0.5 * G * h * h, h * v * (h * u / h)])
        assert np.allclose(out, f, rtol=1e-12, atol=1e-12 * (1 + abs(f).max()))
    # dry right state: wave speeds are Toro's, flux of a still left state into a dry bed
    h0 = 2.0
    out = oracle.hllc(h0, 0.0, 0.0, 0.0, 0.0, 0.0, 1e-3, G)
    c = math.sqrt(G * h0)
    sL, sR = -c, 2 * c
    fL = [0.0, 0.5 * G * h0 * h0, 0.0]
    want0 = (sR * fL[0] - sL * 0 + sL * sR * (0 - h0)) / (sR - sL)
    assert out[0] == pytest.approx(want0, rel=1e-14)
    assert want0 == pytest.approx(2 * c * h0 / 3, rel=1e-14)   # the dry-front mass flux 2c h0 / 3
    assert np.all(oracle.hllc(1e-5, 1e-6, 0, 2e-5, 0, 0, 1e-4, G) == 0)


@pytest.mark.parametrize("a,b", [(1.0, 3.0), (3.0, 1.0), (-1.0, -4.0), (2.0, -1.0), (0.0, 1.0), (1e-3, 5e-3)])
def test_limiters_closed_form(a, b):
    mm = 0 if a * b <= 0 else (min(a, b) if a > 0 else max(a, b))
    mc = 0 if a * b <= 0 else math.copysign(min(abs(0.5 * (a + b)), 2 * abs(a), 2 * abs(b)), a)
    assert oracle.limit("minmod", a, b) == mm
    assert oracle.limit("mc", a, b) == mc
    assert oracle.limit("central", a, b) == 0.5 * (a + b)


def test_vel_closed_form():
    eps4 = oracle.js_pow(1e-3, 4.0)
    assert oracle.vel(3.0, 0.0, eps4) == 0.0
    for h in (0.01, 0.5, 3.0, 40.0):
        for u in (-2.0, 0.3, 5.0):
            v = oracle.vel(h * u, h, eps4)
            assert abs(v - u) <= 4 * math.ulp(u)
    # bounded by |hq| / minDepth in the desingularised range
    for h in (1e-4, 5e-4, 1e-3):
        assert abs(oracle.vel(1e-4, h, eps4)) <= 1e-4 / 1e-3 * (1 + 1e-12)
