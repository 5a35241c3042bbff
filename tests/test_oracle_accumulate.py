"""Pins the oracle's accumulation (accumulate_scatter.cu, scatter_utils.cuh):
AABB scatter == pixel-major brute force, thread-count and crop invariance, GPS point
statistics, and convergence of the SPP average to the FP64 sorted alpha blend."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import oracle_scene, small_camera, small_scene

EMPTY = np.iinfo(np.uint64).max


@pytest.mark.parametrize("ss,aa,blue", [(1, False, False), (2, False, False), (1, True, False), (1, False, True)])
def test_scatter_equals_bruteforce(oracle_mod, ss, aa, blue):
    sc = oracle_scene(small_scene(250, seed=11, log_scale_mean=-3.2, log_scale_std=0.6))
    cam = small_camera(48, 40, supersampling_factor=ss, proper_antialiasing=aa)
    bn, rot = None, 0
    if blue:
        bn = np.random.default_rng(5).integers(0, 256, 128 * 128, dtype=np.uint8)
        rot = 0x9E3779B9
    for frame in (0, 7):
        fb, st = oracle_mod.accumulate(sc, cam, frame, blue_noise_slice=bn, blue_noise_rotation=rot)
        bf = oracle_mod.accumulate_bruteforce(sc, cam, frame, blue_noise_slice=bn, blue_noise_rotation=rot)
        assert np.array_equal(fb, bf)
        assert (fb != EMPTY).mean() > 0.1
        assert st["n_accepted"] > 0 and st["n_tested"] >= st["n_accepted"]


def test_thread_count_and_crop_invariance(oracle_mod):
    sc = oracle_scene(small_scene(3000, seed=2, log_scale_mean=-3.8))
    cam = small_camera(160, 120)
    fb1, st1 = oracle_mod.accumulate(sc, cam, 3, n_threads=1)
    fb4, st4 = oracle_mod.accumulate(sc, cam, 3, n_threads=4)
    assert np.array_equal(fb1, fb4) and st1 == st4
    crop = (37, 21, 101, 90)
    fbc, _ = oracle_mod.accumulate(sc, cam, 3, crop=crop)
    assert np.array_equal(fbc, fb1[crop[1]:crop[3], crop[0]:crop[2]])
    for flags in ({"use_gps": True}, {"use_gps": True, "max_gps_points": 64}):
        camg = dict(cam, **flags)
        g1, _ = oracle_mod.accumulate(sc, camg, 3, n_threads=1)
        g4, _ = oracle_mod.accumulate(sc, camg, 3, n_threads=4)
        gc, _ = oracle_mod.accumulate(sc, camg, 3, crop=crop)
        assert np.array_equal(g1, g4) and np.array_equal(gc, g1[crop[1]:crop[3], crop[0]:crop[2]])


def test_winner_encoding(oracle_mod):
    arrays = small_scene(400, seed=4, log_scale_mean=-3.5)
    sc = oracle_scene(arrays)
    cam = small_camera(64, 48)
    fb, _ = oracle_mod.accumulate(sc, cam, 1)
    won = fb[fb != EMPTY]
    idx = (won & np.uint64(0xFFFFFFFF)).astype(np.int64)
    depth = (won >> np.uint64(32)).astype(np.uint32).view(np.float32)
    for g, d in zip(idx[:200], depth[:200]):
        p = oracle_mod.project(sc, int(g), cam)
        assert p.active and np.float32(p.depth) == d and d >= cam["near_plane"]


def test_gps_counts_follow_importance(oracle_mod):
    sc = oracle_scene(small_scene(2000, seed=6, log_scale_mean=-3.0, log_scale_std=0.3))
    cam = small_camera(128, 96, use_gps=True, max_gps_points=1 << 30)
    total_c, total_l = 0.0, 0.0
    for frame in range(4):
        capped, true, lam = oracle_mod.gps_counts(sc, cam, frame)
        big = lam > 20.0
        total_c += true[big].sum()
        total_l += lam[big].astype(np.float64).sum()
        assert np.array_equal(capped, true)
    assert abs(total_c / total_l - 1.0) < 0.02
    # the reference's cap rn(W*H / (2 ss^2)) (accumulate_scatter.cu:128)
    capped, true, _ = oracle_mod.gps_counts(sc, dict(cam, max_gps_points=0), 0)
    assert capped.max() <= 128 * 96 // 2 and np.array_equal(capped, np.minimum(true, 128 * 96 // 2))


def test_gps_winners_are_projected_and_in_image(oracle_mod):
    sc = oracle_scene(small_scene(600, seed=8, log_scale_mean=-3.2))
    cam = small_camera(64, 48, use_gps=True, supersampling_factor=2)
    fb, st = oracle_mod.accumulate(sc, cam, 5)
    assert fb.shape == (96, 128)
    won = fb[fb != EMPTY]
    assert len(won) > 100 and st["n_accepted"] >= len(won)
    for g in np.unique((won & np.uint64(0xFFFFFFFF)).astype(np.int64))[:100]:
        assert oracle_mod.project(sc, int(g), cam).active


def _render_mean(oracle_mod, sc, cam, spp, first_frame=0):
    """SPP loop of Renderer.py:85-87: RESET then STATIC, returns the final history."""
    hist = None
    for s in range(spp):
        img, dep, alp, wm, _, _ = oracle_mod.render(sc, cam, first_frame + s, history=hist,
                                                    temporal_mode=0 if s == 0 else 1)
        hist = (img, dep, alp, wm)
    return hist


def test_spp_average_converges_to_fp64_blend(oracle_mod):
    """The expectation of one ST sample is the sorted front-to-back blend; the SPP average
    approaches it with RMSE ~ 1/sqrt(spp)."""
    sc = oracle_scene(small_scene(500, seed=9, log_scale_mean=-3.0, log_scale_std=0.5, opacity_std=1.0))
    cam = small_camera(40, 32)
    ref_img, ref_dep, ref_alp = oracle_mod.blend_fp64(sc, cam)
    rmse = {}
    for spp in (4, 16, 64):
        img, _, alp, wm = _render_mean(oracle_mod, sc, cam, spp, first_frame=1000)
        assert np.allclose(wm, spp)  # weight mass counts samples in STATIC mode
        rmse[spp] = math.sqrt(float(np.mean((img.astype(np.float64) - ref_img) ** 2)))
        rmse_a = math.sqrt(float(np.mean((alp[..., 0].astype(np.float64) - ref_alp) ** 2)))
        assert rmse_a < 1.2 / math.sqrt(spp)
    assert 1.5 < rmse[4] / rmse[16] < 2.7 and 1.5 < rmse[16] / rmse[64] < 2.7
    assert rmse[64] < 0.05
