"""Pins the oracle's counter-based RNG against an independent restatement in Python
integers (arbitrary precision, explicit 32-bit masking) of math_utils.h:45-79, 113-136,
plus the uniformity the renderer relies on."""
from __future__ import annotations

import struct

import numpy as np

M = 0xFFFFFFFF


def hash32(x):
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M
    x ^= x >> 15
    x = (x * 0x846CA68B) & M
    x ^= x >> 16
    return x


def make_seed(idx, frame):
    h1 = hash32(idx ^ ((frame * 0x9E3779B9) & M))
    h2 = hash32(((idx * 0xBB67AE85) & M) ^ frame)
    h1 ^= (h2 * 0x3C6EF372) & M
    h2 ^= (h1 * 0xA54FF53A) & M
    return h1, h2


def pcg2d(x, y):
    x = (x * 1664525 + 1013904223) & M
    y = (y * 1664525 + 1013904223) & M
    x = (x + y * 1664525) & M
    y = (y + x * 1664525) & M
    x ^= x >> 16
    y ^= y >> 16
    x = (x + y * 1664525) & M
    y = (y + x * 1664525) & M
    x ^= x >> 16
    y ^= y >> 16
    return x, y


def pcg_hash(seed):
    state = (seed * 747796405 + 2891336453) & M
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M
    return (word >> 22) ^ word


def rand(frame, instance, x, y):
    seed = instance ^ ((frame * 1973) & M) ^ ((x * 9277) & M) ^ ((y * 2699) & M)
    bits = 0x3F800000 | (pcg_hash(seed) >> 9)
    return np.float32(struct.unpack("<f", struct.pack("<I", bits))[0]) - np.float32(1.0)


def blue_noise(slice_, rotation, instance, x, y):
    ox = ((instance * 0xC13FA9A9) & M) >> 25
    oy = ((instance * 0x91E10DA5) & M) >> 25
    texel = ((y + oy) % 128) * 128 + (x + ox) % 128
    value = ((int(slice_[texel]) << 24) + (1 << 23) + rotation) & M
    return np.float32(value >> 8) * np.float32(2.0 ** -24)


def test_hash_and_seed_known_answers(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(0)
    for a, b in rng.integers(0, 2**32, size=(500, 2), dtype=np.uint64):
        a, b = int(a), int(b)
        assert L.orc_hash32(a) == hash32(a)
        assert L.orc_pcg_hash(a) == pcg_hash(a)
        assert oracle_mod.make_seed(a, b) == make_seed(a, b)
        assert oracle_mod.pcg2d_uint(a, b) == pcg2d(a, b)
    assert L.orc_hash32(0) == 0


def test_rand_known_answers(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(1)
    for f, i, x, y in rng.integers(0, 2**32, size=(500, 4), dtype=np.uint64):
        f, i = int(f), int(i)
        x, y = int(x) % 65536, int(y) % 65536
        assert np.float32(L.orc_rand(f, i, x, y)) == rand(f, i, x, y)


def test_pcg2d_float_and_point_seed(oracle_mod):
    L = oracle_mod.lib()
    st = np.array([12345, 678910], np.uint32)
    out = np.zeros(2, np.float32)
    L.orc_pcg2d(st, out)
    x, y = pcg2d(12345, 678910)
    assert tuple(int(v) for v in st) == (x, y)
    assert out[0] == np.float32(np.float32(x) * np.float32(2.0 ** -32))
    seed = np.zeros(2, np.uint32)
    L.orc_gps_point_seed(77, 5, 9, seed)
    s0, s1 = make_seed(77, 5)
    assert (int(seed[0]), int(seed[1])) == (s0 ^ ((10 * 0x9E3779B9) & M), s1)


def test_blue_noise_known_answers(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(2)
    vol = rng.integers(0, 256, size=128 * 128, dtype=np.uint8)
    for rot, inst, x, y in rng.integers(0, 2**32, size=(300, 4), dtype=np.uint64):
        rot, inst = int(rot), int(inst)
        x, y = int(x) % 4096, int(y) % 4096
        assert np.float32(L.orc_blue_noise(vol.ctypes.data, rot, inst, x, y)) == blue_noise(vol, rot, inst, x, y)


def test_rand_is_uniform(oracle_mod):
    L = oracle_mod.lib()
    n = 64 * 1024
    us = np.array([L.orc_rand(3, 1000 + k // 256, k % 256, (k // 256) % 97) for k in range(n)])
    assert us.min() >= 0.0 and us.max() < 1.0
    hist, _ = np.histogram(us, bins=64, range=(0.0, 1.0))
    chi2 = ((hist - n / 64) ** 2 / (n / 64)).sum()
    assert chi2 < 110.0  # chi^2 with 63 dof: p ~ 1e-4
