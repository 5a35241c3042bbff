"""GPU parity: the HIP kernels of libsgs.so against the CPU oracle on the same seeded
inputs. Framebuffer winners and GPS point sets are compared bit for bit; resolved
colour, depth, alpha and weight mass within 1e-5 absolute (FP32)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import oracle_scene, small_camera, small_scene

pytestmark = pytest.mark.gpu
RESOLVE_ATOL = 1e-5
EMPTY = np.iinfo(np.uint64).max


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from StochasticGSCudaBackend import _native
    return _native


def to_device(arrays):
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in arrays.items()}


def gpu_args(t, cam):
    import torch
    return (t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"],
            torch.from_numpy(np.asarray(cam["w2c"], np.float32)).cuda(),
            torch.from_numpy(np.asarray(cam["cam_position"], np.float32)).cuda(),
            cam["width"], cam["height"], cam["focal_x"], cam["focal_y"], cam["center_x"], cam["center_y"],
            cam["near_plane"], cam["far_plane"], cam.get("proper_antialiasing", False),
            cam.get("supersampling_factor", 1), cam.get("use_gps", False), cam.get("hybrid_scheduling", False),
            cam.get("spatial_reuse", False))


def gpu_fb(gpu, t, cam, frame, tuning=None, blue=None, counters=None):
    import torch
    fb = gpu.accumulate(*gpu_args(t, cam), blue_noise=blue, frame_idx=frame, tuning=tuning, counters=counters)
    torch.cuda.synchronize()
    return fb.cpu().numpy().view(np.uint64)


def print_case(**kw):
    print("case:", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


SCATTER, TILES = 2, 1  # sgs_tuning.schedule


@pytest.mark.parametrize("ss,aa,seq,schedule,bin_size", [
    (1, False, 8, SCATTER, 0), (2, False, 8, SCATTER, 0), (1, True, 8, SCATTER, 0), (1, False, 1, SCATTER, 0),
    (1, False, 64, SCATTER, 0),
    (1, False, 0, TILES, 0), (2, False, 0, TILES, 0), (1, True, 0, TILES, 0),
    (1, False, 0, TILES, 32), (2, False, 0, TILES, 64),   # bins of 2x2 and 4x4 raster tiles
])
def test_st_framebuffer_bit_exact(gpu, oracle_mod, ss, aa, seq, schedule, bin_size):
    arrays = small_scene(3000, seed=21, log_scale_mean=-3.6, log_scale_std=0.8)
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = small_camera(160, 120, supersampling_factor=ss, proper_antialiasing=aa)
    for frame in (0, 5, 1 << 20):
        print_case(test="st", ss=ss, aa=aa, seq=seq, schedule=schedule, bin_size=bin_size, frame=frame)
        ref, st = oracle_mod.accumulate(sc, cam, frame)
        got = gpu_fb(gpu, t, cam, frame,
                     tuning=gpu.Tuning(sequential_threshold=seq, schedule=schedule, bin_size=bin_size))
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} of {ref.size} pixels differ"


def test_st_tiles_overflow_falls_back_to_scatter(gpu, oracle_mod):
    # 400 large Gaussians over a 512x384 view bin into far more than the single-view
    # instance capacity (2^16 entries for so few Gaussians): the gated scatter runs instead
    import math
    arrays = small_scene(400, seed=23, log_scale_mean=-1.2, log_scale_std=0.3)
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = small_camera(512, 384, distance=2.2)
    # the (Gaussian, tile) instances the binning would need, from the oracle's AABBs
    need = 0
    for i in range(400):
        p = oracle_mod.project(sc, i, cam)
        if p.active:
            x0, x1 = max(math.floor(p.bounds[0]), 0), min(math.ceil(p.bounds[1]), 512)
            y0, y1 = max(math.floor(p.bounds[2]), 0), min(math.ceil(p.bounds[3]), 384)
            if x1 > x0 and y1 > y0:
                need += ((x1 - 1) // 16 - x0 // 16 + 1) * ((y1 - 1) // 16 - y0 // 16 + 1)
    assert need > (1 << 16), need
    for frame in (1, 9):
        print_case(test="st_tiles_overflow", frame=frame)
        ref, _ = oracle_mod.accumulate(sc, cam, frame)
        got = gpu_fb(gpu, t, cam, frame, tuning=gpu.Tuning(schedule=TILES))
        assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} of {ref.size} pixels differ"


def test_st_blue_noise_bit_exact(gpu, oracle_mod):
    import torch
    arrays = small_scene(2000, seed=22, log_scale_mean=-3.5)
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = small_camera(128, 96)
    vol = np.random.default_rng(3).integers(0, 256, (64, 128, 128), dtype=np.uint8)
    for frame in (3, 64 + 9):
        print_case(test="st_blue", frame=frame)
        ref, _ = oracle_mod.accumulate(sc, cam, frame, blue_noise_slice=vol[frame % 64].reshape(-1),
                                       blue_noise_rotation=((frame // 64) * 0x9E3779B9) & 0xFFFFFFFF)
        got = gpu_fb(gpu, t, cam, frame, blue=torch.from_numpy(vol).cuda())
        assert np.array_equal(got, ref)


def test_st_counters_match_oracle(gpu, oracle_mod):
    import torch
    arrays = small_scene(2500, seed=23, log_scale_mean=-3.7)
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = small_camera(128, 96)
    counters = torch.zeros(4, dtype=torch.int64, device="cuda")
    gpu_fb(gpu, t, cam, 2, counters=counters)
    _, st = oracle_mod.accumulate(sc, cam, 2)
    c = counters.cpu().numpy()
    assert (c[0], c[1], c[2]) == (st["n_tested"], st["n_accepted"], st["n_projected"])


@pytest.mark.parametrize("ss,max_points,seq", [(1, 0, 8), (2, 0, 8), (1, 64, 4), (1, 0, 200)])
def test_gps_framebuffer_bit_exact(gpu, oracle_mod, ss, max_points, seq):
    arrays = small_scene(2000, seed=24, log_scale_mean=-3.6, log_scale_std=0.8)
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = small_camera(128, 96, supersampling_factor=ss, use_gps=True, max_gps_points=max_points)
    for frame in (0, 11):
        print_case(test="gps", ss=ss, max_points=max_points, seq=seq, frame=frame)
        ref, _ = oracle_mod.accumulate(sc, cam, frame)
        got = gpu_fb(gpu, t, cam, frame, tuning=gpu.Tuning(sequential_threshold=seq, max_gps_points=max_points))
        assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} of {ref.size} pixels differ"


def _gpu_render(gpu, t, cam, frame, hist=None, mode=0, v2p=None, kb=None):
    import torch
    H, W = cam["height"], cam["width"]
    if hist is None:
        hist = tuple(torch.empty(0) for _ in range(4))
    else:
        hist = tuple(torch.from_numpy(np.ascontiguousarray(h, np.float32)).cuda() for h in hist)
    v2p_t = torch.from_numpy(np.asarray(np.eye(4) if v2p is None else v2p, np.float32)).cuda()
    if kb is None:
        kb = (torch.empty((0, 0, 2), dtype=torch.int32, device="cuda"), torch.empty((0, 0, 4), device="cuda"))
    out = gpu.render(*gpu_args(t, cam), torch.empty(0, dtype=torch.uint8, device="cuda"), kb[0], kb[1], *hist, v2p_t,
                     mode, frame_idx=frame, return_framebuffer=True)
    torch.cuda.synchronize()
    ss = max(1, int(cam.get("supersampling_factor", 1)))
    fb = out[4].cpu().numpy().view(np.uint64).reshape(H * ss, W * ss)
    return [o.cpu().numpy() for o in out[:4]], fb


def _close(got, ref):
    for g, r, name in zip(got, ref, ("image", "depth", "alpha", "weight_mass")):
        assert g.shape == r.shape, name
        err = np.abs(g.astype(np.float64) - r.astype(np.float64)).max() if g.size else 0.0
        assert err <= RESOLVE_ATOL, f"{name}: max abs error {err}"


@pytest.mark.parametrize("ss,gps", [(1, False), (2, False), (1, True)])
def test_render_temporal_modes_match_oracle(gpu, oracle_mod, ss, gps):
    arrays = small_scene(2000, seed=25, log_scale_mean=-3.6)
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = small_camera(96, 72, supersampling_factor=ss, use_gps=gps)
    hist_o = hist_g = None
    for k, frame in enumerate((40, 41, 42)):
        mode = 0 if k == 0 else 1
        print_case(test="render", ss=ss, gps=gps, mode=mode, frame=frame)
        ref = oracle_mod.render(sc, cam, frame, history=hist_o, temporal_mode=mode)
        got, fb = _gpu_render(gpu, t, cam, frame, hist=hist_g, mode=mode)
        assert np.array_equal(fb, ref[4])
        _close(got, ref[:4])
        hist_o, hist_g = ref[:4], ref[:4]
    # camera moves: REPROJECT against the accumulated history
    cam2 = small_camera(96, 72, angle=0.02, supersampling_factor=ss, use_gps=gps)
    w2c_prev = np.asarray(cam["w2c"], np.float64)
    c2w_new = np.linalg.inv(np.asarray(cam2["w2c"], np.float64))
    v2p = (w2c_prev @ c2w_new).astype(np.float32)
    print_case(test="render_reproject", ss=ss, gps=gps)
    ref = oracle_mod.render(sc, cam2, 43, history=hist_o, temporal_mode=2, view_to_previous_view=v2p)
    got, fb = _gpu_render(gpu, t, cam2, 43, hist=hist_g, mode=2, v2p=v2p)
    assert np.array_equal(fb, ref[4])
    _close(got, ref[:4])


@pytest.mark.parametrize("sh_degree", [0, 1, 2])
def test_render_sh_degrees_match_oracle(gpu, oracle_mod, sh_degree):
    """Shading with SH degrees 0-2 (total_sh_bases 1, 4, 9; gs_utils.h:212-243,
    render.cu:150): the fused shade+resolve (RESET, STATIC), the separate sample + resolve
    (REPROJECT) and the spatial-reuse sample kernel, against the oracle."""
    import torch
    arrays = small_scene(1500, seed=29, log_scale_mean=-3.5, sh_degree=sh_degree)
    assert arrays["sh"].shape[1] == (sh_degree + 1) ** 2
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = small_camera(80, 60, supersampling_factor=2 if sh_degree == 1 else 1)
    hist = None
    for k, frame in enumerate((3, 4)):
        print_case(test="sh_degree", sh_degree=sh_degree, mode=min(k, 1), frame=frame)
        ref = oracle_mod.render(sc, cam, frame, history=hist, temporal_mode=min(k, 1))
        got, fb = _gpu_render(gpu, t, cam, frame, hist=hist, mode=min(k, 1))
        assert np.array_equal(fb, ref[4])
        _close(got, ref[:4])
        hist = ref[:4]
    cam2 = small_camera(80, 60, angle=0.03, supersampling_factor=2 if sh_degree == 1 else 1)
    v2p = (np.asarray(cam["w2c"], np.float64) @ np.linalg.inv(np.asarray(cam2["w2c"], np.float64))).astype(np.float32)
    print_case(test="sh_degree_reproject", sh_degree=sh_degree)
    ref = oracle_mod.render(sc, cam2, 5, history=hist, temporal_mode=2, view_to_previous_view=v2p)
    got, fb = _gpu_render(gpu, t, cam2, 5, hist=hist, mode=2, v2p=v2p)
    assert np.array_equal(fb, ref[4])
    _close(got, ref[:4])
    cam3 = small_camera(64, 48, spatial_reuse=True)
    kb_o = oracle_mod.empty_k_buffer(cam3)
    kb_g = tuple(torch.from_numpy(k.copy()).cuda() for k in kb_o)
    hist = None
    for k, frame in enumerate((6, 7)):
        print_case(test="sh_degree_reuse", sh_degree=sh_degree, frame=frame)
        ref = oracle_mod.render(sc, cam3, frame, history=hist, temporal_mode=min(k, 1), k_buffer=kb_o)
        got, fb = _gpu_render(gpu, t, cam3, frame, hist=hist, mode=min(k, 1), kb=kb_g)
        assert np.array_equal(fb, ref[4])
        _close(got, ref[:4])
        assert np.array_equal(kb_g[0].cpu().numpy(), kb_o[0])
        hist = ref[:4]


def test_spatial_reuse_matches_oracle(gpu, oracle_mod):
    import torch
    arrays = small_scene(1500, seed=26, log_scale_mean=-3.4)
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = small_camera(80, 60, spatial_reuse=True)
    kb_o = oracle_mod.empty_k_buffer(cam)
    kb_g = tuple(torch.from_numpy(k.copy()).cuda() for k in kb_o)
    hist_o = None
    for k, frame in enumerate((7, 8, 9, 10)):
        mode = 0 if k == 0 else 1
        print_case(test="spatial_reuse", mode=mode, frame=frame)
        ref = oracle_mod.render(sc, cam, frame, history=hist_o, temporal_mode=mode, k_buffer=kb_o)
        got, fb = _gpu_render(gpu, t, cam, frame, hist=hist_o, mode=mode, kb=kb_g)
        assert np.array_equal(fb, ref[4])
        _close(got, ref[:4])
        assert np.array_equal(kb_g[0].cpu().numpy(), kb_o[0])
        assert np.abs(kb_g[1].cpu().numpy() - kb_o[1]).max() <= RESOLVE_ATOL
        hist_o = ref[:4]


def test_determinism_and_empty_scene(gpu):
    arrays = small_scene(3000, seed=27, log_scale_mean=-3.4)
    t = to_device(arrays)
    cam = small_camera(128, 96)
    a = gpu_fb(gpu, t, cam, 9)
    b = gpu_fb(gpu, t, cam, 9)
    assert np.array_equal(a, b)
    empty = to_device({k: v[:0] for k, v in arrays.items()})
    assert np.all(gpu_fb(gpu, empty, cam, 0) == EMPTY)
    got, fb = _gpu_render(gpu, empty, cam, 0)
    assert np.all(fb == EMPTY) and np.all(got[0] == 0.0) and np.all(got[2] == 0.0) and np.all(got[3] == 1.0)


def test_invalid_arguments_raise(gpu):
    import torch
    arrays = small_scene(10)
    t = to_device(arrays)
    cam = small_camera(16, 16)
    with pytest.raises(ValueError):
        _gpu_render(gpu, t, cam, 0, mode=3)
    with pytest.raises(TypeError):
        bad = dict(t, means=t["means"].double())
        gpu_fb(gpu, bad, cam, 0)
    with pytest.raises(ValueError):
        gpu.accumulate(*gpu_args(t, dict(cam, spatial_reuse=True)), frame_idx=0, tuning=gpu.Tuning(tile_size=48))
    torch.cuda.synchronize()


def test_public_render_api(gpu):
    """The drop-in API of render.py:72-95 with the SPP loop of Renderer.py:85-87."""
    import torch
    from StochasticGSCudaBackend import History, Settings, TemporalMode, create_k_buffer, render
    arrays = small_scene(1000, seed=28, log_scale_mean=-3.5)
    t = to_device(arrays)
    cam = small_camera(64, 48, spatial_reuse=True)
    settings = Settings(torch.from_numpy(cam["w2c"]).cuda(), torch.from_numpy(cam["cam_position"]).cuda(), 64, 48,
                        cam["focal_x"], cam["focal_y"], cam["center_x"], cam["center_y"], 0.2, 1000.0, False, 1,
                        False, False, True)
    k_buffer = create_k_buffer(settings)
    assert k_buffer[0].shape == (4, 64 * 48, 2) and k_buffer[1].shape == (4, 64 * 48, 4)
    history = History(*(torch.empty(0) for _ in range(4)))
    mode = TemporalMode.RESET
    for _ in range(4):
        history = render(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], settings, k_buffer, history,
                         torch.eye(4, device="cuda"), mode)
        mode = TemporalMode.STATIC
    torch.cuda.synchronize()
    assert history.image.shape == (48, 64, 3) and history.alpha.shape == (48, 64, 1)
    assert float(history.alpha.max()) <= 1.0 + 1e-5 and float(history.alpha.max()) > 0.1


# ---------------------------------------------------------------------------------------------
# hybrid scheduling: must equal the scatter path (and hence the oracle) bit for bit
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("tile,threshold,max_instances,max_binned,blue", [
    (32, 64, 0, 0, False), (16, 32, 0, 0, False), (64, 128, 0, 0, True),
    (32, 64, 1, 0, False),      # instance list overflow: binned Gaussians go to the fallback scatter
    (32, 16, 0, 7, False),      # binning capacity overflow: the rest is scattered directly
])
def test_hybrid_st_equals_scatter(gpu, oracle_mod, tile, threshold, max_instances, max_binned, blue):
    import torch
    arrays = small_scene(3000, seed=31, log_scale_mean=-3.2, log_scale_std=0.9)
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = small_camera(200, 150, hybrid_scheduling=True)
    vol = np.random.default_rng(4).integers(0, 256, (64, 128, 128), dtype=np.uint8) if blue else None
    tuning = gpu.Tuning(tile_size=tile, binning_pixel_threshold=threshold, max_tile_instances=max_instances,
                        max_binned_gaussians=max_binned)
    counters = torch.zeros(4, dtype=torch.int64, device="cuda")
    for frame in (2, 70):
        print_case(test="hybrid_st", tile=tile, threshold=threshold, max_instances=max_instances,
                   max_binned=max_binned, blue=blue, frame=frame)
        if blue:
            ref, _ = oracle_mod.accumulate(sc, cam, frame, blue_noise_slice=vol[frame % 64].reshape(-1),
                                           blue_noise_rotation=((frame // 64) * 0x9E3779B9) & 0xFFFFFFFF)
        else:
            ref, _ = oracle_mod.accumulate(sc, cam, frame)
        got = gpu_fb(gpu, t, cam, frame, tuning=tuning, counters=counters,
                     blue=torch.from_numpy(vol).cuda() if blue else None)
        assert int(counters[3]) > (0 if max_binned == 0 else 6), "nothing was binned"
        assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} of {ref.size} pixels differ"


@pytest.mark.parametrize("ss,threshold,max_binned", [(1, 32, 0), (2, 64, 0), (1, 16, 5)])
def test_hybrid_gps_equals_scatter(gpu, oracle_mod, ss, threshold, max_binned):
    import torch
    arrays = small_scene(2000, seed=32, log_scale_mean=-3.3, log_scale_std=0.8)
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = small_camera(160, 120, supersampling_factor=ss, use_gps=True, hybrid_scheduling=True)
    tuning = gpu.Tuning(binning_point_threshold=threshold, max_binned_gaussians=max_binned)
    counters = torch.zeros(4, dtype=torch.int64, device="cuda")
    for frame in (1, 9):
        print_case(test="hybrid_gps", ss=ss, threshold=threshold, max_binned=max_binned, frame=frame)
        ref, _ = oracle_mod.accumulate(sc, cam, frame)
        got = gpu_fb(gpu, t, cam, frame, tuning=tuning, counters=counters)
        assert int(counters[3]) > 0, "nothing was binned"
        assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} of {ref.size} pixels differ"


def test_hybrid_render_matches_oracle(gpu, oracle_mod):
    arrays = small_scene(2000, seed=33, log_scale_mean=-3.3)
    sc, t = oracle_scene(arrays), to_device(arrays)
    for gps in (False, True):
        cam = small_camera(96, 72, hybrid_scheduling=True, use_gps=gps)
        print_case(test="hybrid_render", gps=gps)
        ref = oracle_mod.render(sc, cam, 5)
        got, fb = _gpu_render(gpu, t, cam, 5)
        assert np.array_equal(fb, ref[4])
        _close(got, ref[:4])
