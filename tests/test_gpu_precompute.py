"""Load-time precompute (sgs_precompute_cov3d, README.md:139-142) and the workspace cap.

* The cache holds each Gaussian's sigmoid opacity, 3D covariance R diag(exp(2 s)) R^T of the
  normalised quaternion and log(255 opacity) (checked in FP64), opacity -1 for Gaussians culled
  in every view.
* Renders with the cache equal renders without it bit for bit, in every schedule and mode.
* A workspace cap below the tile schedule's need switches the view to the scatter schedule:
  same framebuffer."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import small_camera, small_scene
from test_gpu_batch import assert_close, settings_of
from test_gpu_parity import gpu_args, print_case, to_device

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from StochasticGSCudaBackend import _native
    return _native


def test_cache_contents_fp64(gpu):
    arrays = small_scene(2000, seed=81, log_scale_mean=-3.5)
    arrays["opacities"][:10] = -9.0                      # sigmoid < 1/255: culled everywhere
    arrays["rotations"][10:12] = 0.0                     # degenerate quaternion: culled
    t = to_device(arrays)
    cache = gpu.precompute_cov3d(t["means"], t["scales"], t["rotations"], t["opacities"]).cpu().numpy()
    assert cache.shape == (2000, 8)
    q = arrays["rotations"].astype(np.float64)
    ok = np.arange(2000) >= 12
    q = q[ok] / np.linalg.norm(q[ok], axis=1, keepdims=True)
    r, i, j, k = q.T
    R = np.stack([np.stack([1 - 2 * (j * j + k * k), 2 * (i * j - r * k), 2 * (i * k + r * j)], -1),
                  np.stack([2 * (i * j + r * k), 1 - 2 * (i * i + k * k), 2 * (j * k - r * i)], -1),
                  np.stack([2 * (i * k - r * j), 2 * (j * k + r * i), 1 - 2 * (i * i + j * j)], -1)], -2)
    S = np.exp(2.0 * arrays["scales"][ok].astype(np.float64))
    cov = np.einsum("nab,nb,ncb->nac", R, S, R)
    got = cache[ok]
    sym = np.stack([cov[:, 0, 0], cov[:, 0, 1], cov[:, 0, 2], cov[:, 1, 1], cov[:, 1, 2], cov[:, 2, 2]], -1)
    scale = np.abs(sym).max(axis=1, keepdims=True)
    assert np.all(np.abs(got[:, :6] - sym) <= 1e-5 * scale)
    o = 1.0 / (1.0 + np.exp(-arrays["opacities"][ok].astype(np.float64)))
    live = o >= 1.0 / 255.0
    assert np.allclose(got[live, 6], o[live], rtol=1e-6)
    assert np.allclose(got[live, 7], np.log(255.0 * o[live]), atol=2e-6)
    assert np.all(cache[:12, 6] == -1.0)


@pytest.mark.parametrize("flags,tuning", [
    (dict(), None), (dict(), dict(schedule=2)), (dict(proper_antialiasing=True), None),
    (dict(use_gps=True), None), (dict(hybrid_scheduling=True), None),
    (dict(use_gps=True, hybrid_scheduling=True, supersampling_factor=2), None),
    (dict(spatial_reuse=True), None), (dict(), dict(bin_size=32)),
])
def test_render_with_cache_is_bit_identical(gpu, flags, tuning):
    import torch
    arrays = small_scene(3000, seed=82, log_scale_mean=-3.4, log_scale_std=0.8)
    t = to_device(arrays)
    cache = gpu.precompute_cov3d(t["means"], t["scales"], t["rotations"], t["opacities"])
    cam = small_camera(128, 96, **flags)
    cam2 = small_camera(128, 96, angle=0.03, **flags)
    v2p = torch.from_numpy((np.asarray(cam["w2c"], np.float64) @ np.linalg.inv(np.asarray(cam2["w2c"], np.float64)))
                           .astype(np.float32)).cuda()
    tun = gpu.Tuning(**tuning) if tuning else None
    ss = max(1, cam.get("supersampling_factor", 1))
    P = 128 * 96 * ss * ss
    res = []
    for use in (False, True):
        print_case(test="cache", flags=flags, tuning=tuning, cache=use)
        kb = (torch.full((4, P, 2), -1, dtype=torch.int32, device="cuda"), torch.zeros((4, P, 4), device="cuda")) \
            if flags.get("spatial_reuse") else (torch.empty((0, 0, 2), dtype=torch.int32, device="cuda"),
                                                torch.empty((0, 0, 4), device="cuda"))
        hist = tuple(torch.empty(0) for _ in range(4))
        outs = []
        for step, (c, mode) in enumerate(((cam, 0), (cam, 1), (cam2, 2))):
            o = gpu.render(*gpu_args(t, c), torch.empty(0, dtype=torch.uint8, device="cuda"), kb[0], kb[1], *hist,
                           v2p, mode, frame_idx=50 + step, tuning=tun, return_framebuffer=True,
                           cov3d_cache=cache if use else None)
            hist = o[:4]
            outs.append([x.clone() for x in o])
        torch.cuda.synchronize()
        res.append((outs, [k.clone() for k in kb]))
    for a, b in zip(res[0][0], res[1][0]):
        for x, y in zip(a, b):
            assert torch.equal(x, y)
    for x, y in zip(res[0][1], res[1][1]):
        assert torch.equal(x, y)


def test_batch_and_serve_with_cache(gpu):
    import torch
    from StochasticGSCudaBackend.batch import BatchView, render_batch
    arrays = small_scene(3000, seed=83, log_scale_mean=-3.3, log_scale_std=0.8)
    t = to_device(arrays)
    cache = gpu.precompute_cov3d(t["means"], t["scales"], t["rotations"], t["opacities"])
    cams = [small_camera(96, 64, angle=0.2), small_camera(200, 150, angle=1.0)]
    views = [BatchView(settings_of(c), spp=s, frame_idx=f, bin_size=b)
             for c, s, f, b in zip(cams, (3, 5), (7, 900), (16, 32))]
    a = render_batch(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views)
    b = render_batch(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views, cov3d_cache=cache)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert_close(x, y)


def test_workspace_cap_switches_to_scatter(gpu):
    """With a cap below the tile schedule's workspace the view scatters: the workspace size
    drops to the scatter schedule's and the framebuffer is the same."""
    import ctypes as C
    import torch
    arrays = small_scene(20000, seed=84, log_scale_mean=-3.5)
    t = to_device(arrays)
    cam = small_camera(320, 240)
    keep = []
    s = gpu.settings_struct(cam["w2c"], cam["cam_position"], 320, 240, cam["focal_x"], cam["focal_y"], cam["center_x"],
                            cam["center_y"], cam["near_plane"], cam["far_plane"], False, 1, False, False, False,
                            torch.device("cuda"), keep)
    lib = gpu.library()
    size = lambda tun: lib.sgs_render_workspace_size(C.byref(s), 20000, C.byref(tun.c()))
    tiles, capped, scatter = size(gpu.Tuning(max_workspace_mib=1 << 20)), size(gpu.Tuning(max_workspace_mib=1)), \
        size(gpu.Tuning(schedule=2))
    assert tiles > (1 << 20) and capped == scatter < tiles
    print_case(test="workspace_cap")
    fa = gpu.accumulate(*gpu_args(t, cam), frame_idx=3, tuning=gpu.Tuning(max_workspace_mib=1))
    fb = gpu.accumulate(*gpu_args(t, cam), frame_idx=3)
    fc = gpu.accumulate(*gpu_args(t, cam), frame_idx=3, tuning=gpu.Tuning(schedule=2))
    torch.cuda.synchronize()
    assert torch.equal(fa, fb) and torch.equal(fa, fc)
