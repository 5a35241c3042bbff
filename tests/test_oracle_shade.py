"""Pins eval_sh (gs_utils.h:197-243) against the real spherical harmonics built from
scipy's associated Legendre functions, and the resolve helpers (resolve.cu:146-162)."""
from __future__ import annotations

import math

import numpy as np
from scipy.special import lpmv


def real_sh(l: int, m: int, d: np.ndarray) -> float:
    """Real SH with the Condon-Shortley phase (in lpmv), basis order m = -l..l."""
    x, y, z = d
    theta = math.acos(max(-1.0, min(1.0, z)))
    phi = math.atan2(y, x)
    am = abs(m)
    K = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    P = lpmv(am, l, math.cos(theta))
    if m == 0:
        return K * P
    if m > 0:
        return math.sqrt(2.0) * K * math.cos(m * phi) * P
    return math.sqrt(2.0) * K * math.sin(am * phi) * P


def test_eval_sh_basis_functions(oracle_mod):
    rng = np.random.default_rng(0)
    dirs = rng.standard_normal((40, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    k = 0
    for l in range(4):
        for m in range(-l, l + 1):
            for d in dirs:
                c = np.zeros((16, 3), np.float32)
                c[k] = [0.1, -0.05, 0.02]
                got = oracle_mod.eval_sh(c, d.astype(np.float32), 16)
                y = real_sh(l, m, d)
                ref = np.maximum(0.5 + np.array([0.1, -0.05, 0.02]) * y, 0.0)
                assert np.abs(got - ref).max() < 2e-6, (l, m)
            k += 1


def test_eval_sh_degree_truncation(oracle_mod):
    rng = np.random.default_rng(1)
    c = (0.1 * rng.standard_normal((16, 3))).astype(np.float32)
    d = np.array([0.36, -0.48, 0.8], np.float32)
    for bases in (1, 4, 9):
        full = oracle_mod.eval_sh(np.ascontiguousarray(c[:bases]), d, bases)
        masked = c.copy()
        masked[bases:] = 0.0
        assert np.array_equal(full, oracle_mod.eval_sh(masked, d, 16))
    # degree 0: 0.5 + C0 * c0, clamped at zero
    c0 = np.array([[0.2, -3.0, 0.0]], np.float32)
    out = oracle_mod.eval_sh(c0, d, 1)
    assert np.allclose(out, [0.5 + 0.28209479177387814 * 0.2, 0.0, 0.5], atol=1e-7)


def test_catmull_rom_partition_and_linear_reproduction(oracle_mod):
    out = np.zeros(4, np.float32)
    for t in np.linspace(0.0, 1.0, 33, dtype=np.float32):
        oracle_mod.lib().orc_catmull_rom_weights(float(t), out)
        assert abs(out.sum() - 1.0) < 1e-6
        # taps at -1, 0, 1, 2 reproduce a linear ramp exactly
        assert abs(float(np.dot(out, [-1.0, 0.0, 1.0, 2.0])) - float(t)) < 1e-6
    oracle_mod.lib().orc_catmull_rom_weights(0.0, out)
    assert np.array_equal(out, [0.0, 1.0, 0.0, 0.0])


def test_ycocg_is_the_lossless_transform(oracle_mod):
    rng = np.random.default_rng(2)
    out = np.zeros(3, np.float32)
    for rgb in rng.random((50, 3)).astype(np.float32):
        oracle_mod.lib().orc_rgb_to_ycocg(rgb, out)
        y, co, cg = out.astype(np.float64)
        tmp = y - cg
        back = np.array([tmp + co, y + cg, tmp - co])
        assert np.abs(back - rgb).max() < 1e-6
