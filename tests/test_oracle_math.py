"""Pins the oracle's FP32 math library and GPS sampling math against FP64 closed forms."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import ulp_distance


def _vec(fn, xs):
    return np.array([fn(float(x)) for x in xs], np.float32)


def test_exp2_against_fp64(oracle_mod):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-125.0, 0.0, 4000), rng.uniform(-2.0, 2.0, 2000), rng.uniform(0.0, 127.0, 1000)])
    xs = xs.astype(np.float32)
    got = _vec(oracle_mod.exp2f, xs)
    ref = np.exp2(xs.astype(np.float64))
    assert ulp_distance(got, ref).max() <= 2.0
    assert oracle_mod.exp2f(0.0) == 1.0 and oracle_mod.exp2f(-1.0) == 0.5 and oracle_mod.exp2f(3.0) == 8.0
    assert oracle_mod.exp2f(float("nan")) == 0.0 and oracle_mod.exp2f(-126.0) == 0.0
    assert oracle_mod.exp2f(200.0) == float("inf")


def test_exp_against_fp64(oracle_mod):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-86.0, 88.0, 5000), rng.uniform(-1.0, 1.0, 2000)]).astype(np.float32)
    got = _vec(oracle_mod.expf, xs)
    ref = np.exp(xs.astype(np.float64))
    assert ulp_distance(got, ref).max() <= 2.0
    assert oracle_mod.expf(0.0) == 1.0
    assert oracle_mod.expf(-100.0) == 0.0 and oracle_mod.expf(100.0) == float("inf")
    assert math.isnan(oracle_mod.expf(float("nan")))


def test_log_against_fp64(oracle_mod):
    rng = np.random.default_rng(2)
    xs = np.concatenate([10.0 ** rng.uniform(-37, 38, 5000), rng.uniform(0.5, 2.0, 3000),
                         rng.uniform(1e-44, 1e-38, 200)]).astype(np.float32)
    xs = xs[xs > 0]
    got = _vec(oracle_mod.logf, xs)
    ref = np.log(xs.astype(np.float64))
    err = np.abs(got.astype(np.float64) - ref)
    # relative accuracy a few ulps, absolute accuracy near log(1) = 0
    assert np.all((ulp_distance(got, ref) <= 4.0) | (err <= 2.0e-7 * np.abs(ref) + 1.0e-8))
    assert oracle_mod.logf(1.0) == 0.0
    assert oracle_mod.logf(0.0) == float("-inf")
    assert math.isnan(oracle_mod.logf(-1.0))
    assert oracle_mod.logf(float("inf")) == float("inf")


def test_sincos_2pi_against_fp64(oracle_mod):
    rng = np.random.default_rng(3)
    us = np.concatenate([rng.random(6000), [0.0, 0.125, 0.25, 0.5, 0.75, 1.0]]).astype(np.float32)
    s, c = zip(*[oracle_mod.sincos_2pi(float(u)) for u in us])
    ang = 2.0 * np.pi * us.astype(np.float64)
    assert np.abs(np.array(s) - np.sin(ang)).max() <= 4.0e-7
    assert np.abs(np.array(c) - np.cos(ang)).max() <= 4.0e-7
    assert oracle_mod.sincos_2pi(0.0) == (0.0, 1.0)
    assert oracle_mod.sincos_2pi(0.25)[0] == 1.0 and oracle_mod.sincos_2pi(0.5)[1] == -1.0


def _li2(x: float) -> float:
    """FP64 dilogarithm Li2(x) on [0, 1] (series with the reflection formula)."""
    def series(z):
        s, t, k = 0.0, z, 1
        while True:
            term = t / (k * k)
            s += term
            if abs(term) < 1e-18:
                return s
            k += 1
            t *= z
    if x <= 0.5:
        return series(x)
    if x >= 1.0:
        return math.pi ** 2 / 6.0
    return math.pi ** 2 / 6.0 - math.log(x) * math.log1p(-x) - series(1.0 - x)


def test_dilog_against_series(oracle_mod):
    L = oracle_mod.lib()
    xs = np.linspace(0.0, 1.0, 1001, dtype=np.float32)
    got = np.array([L.orc_dilog(float(x)) for x in xs])
    ref = np.array([_li2(float(x)) for x in xs])
    # the reference's polynomial approximation (math_utils.h:14-27) is good to ~6e-5
    assert np.abs(got - ref).max() < 1.0e-4
    assert abs(L.orc_dilog(1.0) - math.pi ** 2 / 6.0) < 1.0e-4
    assert L.orc_dilog(0.0) == 0.0


def test_inv_dilog_round_trip(oracle_mod):
    L = oracle_mod.lib()
    xs = np.linspace(0.004, 1.0, 997, dtype=np.float32)
    back = np.array([L.orc_inv_dilog(L.orc_dilog(float(x))) for x in xs])
    assert np.abs(back - xs).max() < 2.0e-3


@pytest.mark.parametrize("lam", [5.0, 50.0, 500.0])
def test_poisson_moments(oracle_mod, lam):
    n = 20000
    counts = np.array([oracle_mod.poisson(oracle_mod.make_seed(i, 7), lam) for i in range(n)], np.float64)
    assert abs(counts.mean() - lam) < 5.0 * math.sqrt(lam / n) + 0.02 * lam
    assert abs(counts.var() / lam - 1.0) < 0.1


def _ks(samples: np.ndarray, cdf) -> float:
    x = np.sort(samples)
    F = cdf(x)
    n = len(x)
    return max(np.max(np.arange(1, n + 1) / n - F), np.max(F - np.arange(0, n) / n))


def test_box_muller_is_standard_normal(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(4)
    out = np.zeros(2, np.float32)
    xs = []
    for u1, u2 in rng.random((20000, 2)).astype(np.float32):
        L.orc_box_muller(float(u1), float(u2), out)
        xs.append(out[0])
    erf = np.vectorize(math.erf)
    d = _ks(np.array(xs, np.float64), lambda x: 0.5 * (1.0 + erf(x / math.sqrt(2.0))))
    assert d < 1.63 / math.sqrt(20000) * 1.2  # KS 1% critical value (with margin)


@pytest.mark.parametrize("opacity", [0.05, 0.5, 0.95])
def test_corrected_box_muller_radial_law(oracle_mod, opacity):
    """Points of a Gaussian of opacity o have density proportional to -ln(1 - o exp(-r^2/2)),
    so t = exp(-R^2/2) has P(t >= a) = (Li2(o) - Li2(o a)) / Li2(o)."""
    L = oracle_mod.lib()
    rng = np.random.default_rng(5)
    od = L.orc_dilog(opacity)
    o_rcp = 1.0 / opacity
    out = np.zeros(2, np.float32)
    ts = []
    for u1, u2 in rng.random((8000, 2)).astype(np.float32):
        L.orc_corrected_box_muller(float(u1), float(u2), o_rcp, od, out)
        ts.append(math.exp(-0.5 * float(out[0] ** 2 + out[1] ** 2)))
    li2o = _li2(opacity)
    # CDF of t: P(t <= a) = Li2(o a) / Li2(o)
    d = _ks(np.array(ts), np.vectorize(lambda a: _li2(opacity * min(max(a, 0.0), 1.0)) / li2o))
    assert d < 0.03
