"""Pins the oracle's sample/resolve (resolve.cu:27-349) with closed forms: box filter,
STATIC running mean, REPROJECT under identity motion, spatial-reuse invariants."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import oracle_scene, small_camera, small_scene


def _samples(oracle_mod, sc, cam, fb):
    W, H = oracle_mod.fb_dims(cam)
    samples = np.zeros((H * W, 4), np.float32)
    oracle_mod.lib().orc_sample(C.byref(sc.c), C.byref(oracle_mod.make_camera(cam)), np.ascontiguousarray(fb.reshape(-1)),
                                samples, None, None, None)
    return samples.reshape(H, W, 4)


@pytest.mark.parametrize("ss", [1, 2, 3])
def test_reset_is_the_subpixel_box_filter(oracle_mod, ss):
    sc = oracle_scene(small_scene(400, seed=1, log_scale_mean=-3.3))
    cam = small_camera(30, 22, supersampling_factor=ss)
    img, dep, alp, wm, fb, _ = oracle_mod.render(sc, cam, 4)
    s = _samples(oracle_mod, sc, cam, fb).astype(np.float64)
    H, W = 22, 30
    blocks = s.reshape(H, ss, W, ss, 4)
    assert np.allclose(img, blocks[..., :3].mean(axis=(1, 3)), atol=1e-6)
    assert np.allclose(dep[..., 0], blocks[..., 3].mean(axis=(1, 3)), atol=1e-5)
    assert np.allclose(alp[..., 0], (blocks[..., 3] > 0).mean(axis=(1, 3)), atol=1e-6)
    assert np.all(wm == 1.0)


def test_static_is_the_running_mean(oracle_mod):
    sc = oracle_scene(small_scene(300, seed=2, log_scale_mean=-3.3))
    cam = small_camera(24, 20)
    singles = [oracle_mod.render(sc, cam, f)[:3] for f in range(10, 16)]
    hist = None
    for k, f in enumerate(range(10, 16)):
        out = oracle_mod.render(sc, cam, f, history=hist, temporal_mode=0 if k == 0 else 1)
        hist = out[:4]
    for i in range(3):
        mean = np.mean([s[i].astype(np.float64) for s in singles], axis=0)
        assert np.allclose(hist[i], mean, atol=2e-6)
    assert np.all(hist[3] == 6.0)


def test_reproject_identity_motion(oracle_mod):
    """Identity view_to_previous_view samples the history at the same pixel (Catmull-Rom
    at integer texels), then blends with weight 0.7 (or 0.7/distance when rejected)."""
    sc = oracle_scene(small_scene(300, seed=3, log_scale_mean=-3.2))
    cam = small_camera(32, 24)
    H, W = 24, 32
    rng = np.random.default_rng(0)
    hist = (rng.random((H, W, 3)).astype(np.float32), rng.random((H, W, 1)).astype(np.float32),
            rng.random((H, W, 1)).astype(np.float32), np.ones((H, W, 1), np.float32))
    cur = oracle_mod.render(sc, cam, 9)
    out = oracle_mod.render(sc, cam, 9, history=hist, temporal_mode=2, view_to_previous_view=np.eye(4))
    interior = np.zeros((H, W), bool)
    interior[2:H - 2, 2:W - 2] = True  # pixel centres within the 2 px border margin
    S, Hh, O = cur[0].astype(np.float64), hist[0].astype(np.float64), out[0].astype(np.float64)
    # discount d satisfies O = S + d (H - S) per pixel, d in (0, 0.7]
    num = ((O - S) * (Hh - S)).sum(-1)
    den = ((Hh - S) ** 2).sum(-1)
    ok = interior & (den > 1e-4)
    d = num[ok] / den[ok]
    assert np.all(d > 0.0) and np.all(d <= 0.7 + 1e-5)
    assert np.allclose(O[ok], S[ok] + d[:, None] * (Hh[ok] - S[ok]), atol=1e-5)
    wm = out[3][..., 0][ok]
    assert np.allclose(wm, 1.0 / (1.0 - d), rtol=1e-5)
    # near the border the history is not used
    assert np.array_equal(out[0][0], cur[0][0]) and np.array_equal(out[3][0], cur[3][0])


def test_spatial_reuse_invariants(oracle_mod):
    sc = oracle_scene(small_scene(300, seed=4, log_scale_mean=-3.0))
    cam = small_camera(32, 24, spatial_reuse=True)
    kb = oracle_mod.empty_k_buffer(cam)
    img, dep, alp, wm, fb, _ = oracle_mod.render(sc, cam, 1, k_buffer=kb)
    covered = fb.reshape(-1) != np.iinfo(np.uint64).max
    kp = kb[0]
    assert np.array_equal(kp[0, covered, 0].astype(np.uint32), (fb.reshape(-1)[covered] & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    assert np.all(kp[1:] == -1)                     # one entry after one sample
    assert np.all(kp[0, ~covered] == -1)
    assert np.all(wm >= 1.0 - 1e-6)                 # own weight 1 plus non-negative neighbour weights
    assert np.all(alp >= 0.0) and np.all(alp <= 1.0 + 1e-6)
    # a second sample with the same winners does not insert duplicates
    oracle_mod.render(sc, cam, 1, k_buffer=kb)
    assert np.all(kb[0][1] == -1)


def test_invalid_temporal_mode(oracle_mod):
    sc = oracle_scene(small_scene(10))
    with pytest.raises(ValueError):
        oracle_mod.render(sc, small_camera(8, 8), 0, temporal_mode=3)
