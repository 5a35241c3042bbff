"""GPU parity at the benchmark configurations' full size: the whole framebuffer is rendered
on the GPU and a 256x256 crop of it is compared bit for bit with the oracle, which
evaluates only the fragments landing in the crop. Size-independent properties cover the
rest of the frame (determinism across schedules)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import oracle_scene
from test_gpu_parity import gpu_fb, print_case, to_device

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from StochasticGSCudaBackend import _native
    return _native


def centre_crop(W, H, size=256):
    x0, y0 = (W - size) // 2, (H - size) // 2
    return (x0, y0, x0 + size, y0 + size)


def test_c2_scatter_and_hybrid_crop_bit_exact(gpu, oracle_mod):
    """1M Gaussians (3-component mixture), 1920x1080, ST: scatter vs hybrid vs oracle."""
    from StochasticGSCudaBackend.synthetic import camera_c2, make_scene, scene_c2
    arrays = make_scene(scene_c2())
    sc, t = oracle_scene(arrays), to_device(arrays)
    threads = min(16, oracle_mod.n_threads_available())
    for frame in (0, 15):
        cam = camera_c2()
        crop = centre_crop(1920, 1080)
        print_case(test="c2", frame=frame)
        ref, st = oracle_mod.accumulate(sc, cam, frame, crop=crop, n_threads=threads)
        scatter = gpu_fb(gpu, t, cam, frame)
        hybrid = gpu_fb(gpu, t, dict(cam, hybrid_scheduling=True), frame)
        assert np.array_equal(scatter, hybrid), "hybrid differs from scatter on the full frame"
        x0, y0, x1, y1 = crop
        assert np.array_equal(scatter[y0:y1, x0:x1], ref)
        assert (ref != np.iinfo(np.uint64).max).mean() > 0.3 and st["n_accepted"] > 0


def test_c3_gps_crop_bit_exact(gpu, oracle_mod):
    """3M heavy-tailed Gaussians, 1920x1080 with 2x2 supersampling, GPS capped at 4096
    points: scatter vs hybrid (point chunks) vs oracle."""
    from StochasticGSCudaBackend.synthetic import camera_c3, make_scene, scene_c3
    arrays = make_scene(scene_c3())
    sc, t = oracle_scene(arrays), to_device(arrays)
    threads = min(16, oracle_mod.n_threads_available())
    cam = camera_c3(use_gps=True, max_gps_points=4096)
    tuning = gpu.Tuning(max_gps_points=4096)
    crop = centre_crop(3840, 2160)
    for frame in (3,):
        print_case(test="c3", frame=frame)
        ref, st = oracle_mod.accumulate(sc, cam, frame, crop=crop, n_threads=threads)
        scatter = gpu_fb(gpu, t, cam, frame, tuning=tuning)
        hybrid = gpu_fb(gpu, t, dict(cam, hybrid_scheduling=True), frame, tuning=tuning)
        assert np.array_equal(scatter, hybrid), "hybrid GPS differs from scatter GPS on the full frame"
        x0, y0, x1, y1 = crop
        assert np.array_equal(scatter[y0:y1, x0:x1], ref)
        assert (ref != np.iinfo(np.uint64).max).mean() > 0.3


def test_c3_st_hybrid_full_frame(gpu, oracle_mod):
    """Heavy-tailed scene with ST: hybrid tiles (16/32/64) equal scatter on the full
    3840x2160 framebuffer; the crop equals the oracle."""
    from StochasticGSCudaBackend.synthetic import camera_c3, make_scene, scene_c3
    arrays = make_scene(scene_c3())
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = camera_c3()
    frame = 4
    print_case(test="c3_st")
    scatter = gpu_fb(gpu, t, cam, frame)
    for tile in (16, 32, 64):
        print_case(test="c3_st_hybrid", tile=tile)
        hybrid = gpu_fb(gpu, t, dict(cam, hybrid_scheduling=True), frame, tuning=gpu.Tuning(tile_size=tile))
        assert np.array_equal(scatter, hybrid), f"tile {tile}: hybrid differs from scatter"
    crop = centre_crop(3840, 2160)
    ref, _ = oracle_mod.accumulate(sc, cam, frame, crop=crop, n_threads=min(16, oracle_mod.n_threads_available()))
    x0, y0, x1, y1 = crop
    assert np.array_equal(scatter[y0:y1, x0:x1], ref)
