"""GPU parity of the batched renderer: every view equals spp successive single-view
render() calls (RESET/STATIC/REPROJECT first sample, then STATIC) and the oracle's SPP
loop; the tile and scatter schedules, their overflow paths and graph replays give the
same bits."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import oracle_scene, small_camera, small_scene
from test_gpu_parity import print_case, to_device

pytestmark = pytest.mark.gpu
RESOLVE_ATOL = 1e-5


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from StochasticGSCudaBackend import _native
    return _native


def settings_of(cam):
    import torch
    from StochasticGSCudaBackend import Settings
    return Settings(torch.from_numpy(np.asarray(cam["w2c"], np.float32)),
                    torch.from_numpy(np.asarray(cam["cam_position"], np.float32)), cam["width"], cam["height"],
                    cam["focal_x"], cam["focal_y"], cam["center_x"], cam["center_y"], cam["near_plane"],
                    cam["far_plane"], cam.get("proper_antialiasing", False), cam.get("supersampling_factor", 1),
                    False, False, False)


def plain_spp(gpu, t, cam, frame0, spp, mode=0, history=None, v2p=None):
    """spp render() calls of the single-view path, like the reference's SPP loop."""
    import torch
    s = settings_of(cam)
    eye = torch.from_numpy(np.asarray(np.eye(4) if v2p is None else v2p, np.float32)).cuda()
    hist = tuple(torch.empty(0) for _ in range(4)) if history is None else history
    kb = (torch.empty((0, 0, 2), dtype=torch.int32, device="cuda"), torch.empty((0, 0, 4), device="cuda"))
    for k in range(spp):
        hist = gpu.render(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"],
                          s.w2c.cuda(), s.cam_position.cuda(), *s.as_tuple()[2:], torch.empty(0, dtype=torch.uint8,
                                                                                              device="cuda"),
                          kb[0], kb[1], *hist, eye, mode if k == 0 else 1, frame_idx=frame0 + k)
    return hist


def assert_close(got, ref, exact=True):
    for g, r, name in zip(got, ref, ("image", "depth", "alpha", "weight_mass")):
        g = g.cpu().numpy() if hasattr(g, "cpu") else g
        r = r.cpu().numpy() if hasattr(r, "cpu") else r
        err = float(np.abs(g.astype(np.float64) - r.astype(np.float64)).max())
        assert err <= (0.0 if exact else RESOLVE_ATOL), f"{name}: max abs error {err}"


def make_views():
    cams = [
        small_camera(64, 48, angle=0.0),
        small_camera(100, 80, angle=0.7, supersampling_factor=2),
        small_camera(37, 29, angle=-1.2, proper_antialiasing=True),
        small_camera(128, 96, angle=2.0, distance=2.2),
        small_camera(50, 50, angle=3.0),
    ]
    spps = [1, 3, 5, 2, 4]
    frames = [0, 100, 2000, 77777, 5]
    return cams, spps, frames


@pytest.mark.parametrize("tuning", [
    dict(),                                                   # defaults (tile schedule)
    dict(max_tile_instances=16),                              # tile list overflow: scatter fallback
    dict(tile_samples=8),                                     # 8 samples per LDS pass
    dict(spps=[16, 9, 8, 1, 12]),                             # 8 per pass up to 8, 16 per half tile above
    dict(spps=[16, 9, 8, 1, 12], bins=[32, 16, 64, 32, 16]),
    dict(bins=[16, 16, 16, 16, 16]),                          # one raster tile per bin
    dict(bins=[32, 64, 16, 64, 32], tile_samples=16),         # bin sizes mixed in one launch
    dict(bins=[64, 64, 64, 64, 64]),                          # 4x4 raster tiles per bin
    dict(schedule=2),                                         # scatter schedule
    dict(schedule=2, big_threshold=64, entry_fragments=64),   # most Gaussians through the work queue
    dict(schedule=2, big_threshold=64, entry_fragments=64, queue_capacity=3),  # queue overflow: in place
    dict(schedule=2, view_major=1, sequential_threshold=1, blocks_per_cu=1),
])
def test_batch_equals_plain_spp_loop(gpu, tuning):
    import torch
    from StochasticGSCudaBackend.batch import BatchRenderer, BatchTuning, BatchView
    arrays = small_scene(3000, seed=41, log_scale_mean=-3.3, log_scale_std=0.8)
    t = to_device(arrays)
    cams, spps, frames = make_views()
    tuning = dict(tuning)
    bins = tuning.pop("bins", [0] * len(cams))
    spps = tuning.pop("spps", spps)
    views = [BatchView(settings_of(c), spp=n, frame_idx=f, bin_size=b) for c, n, f, b in zip(cams, spps, frames, bins)]
    print_case(test="batch_vs_plain", bins=bins, spps=spps, **tuning)
    r = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views, BatchTuning(**tuning))
    outs = r.render()
    torch.cuda.synchronize()
    for v, (c, n, f) in enumerate(zip(cams, spps, frames)):
        ref = plain_spp(gpu, t, c, f, n)
        torch.cuda.synchronize()
        assert_close(outs[v], ref)


@pytest.mark.parametrize("case", ["many_samples", "many_samples_8", "wide_views", "wide_views_16",
                                  "very_wide", "shade_table_overflow"])
def test_batch_tile_schedule_edges(gpu, case, oracle_mod):
    """More than 16 samples (several LDS passes per tile), views wider than a workgroup's
    1024-tile binning window with large Gaussians (direct binning), a framebuffer wider than
    4080 pixels (the 8-byte pair records of the binning passes, in the batch and in the plain
    calls), and a resolve tile with more distinct winners than the batched resolve's
    2048-slot shading table (the winners it cannot take are shaded by their pixel), each
    against the scatter schedule and the plain SPP loop."""
    import torch
    from StochasticGSCudaBackend.batch import BatchRenderer, BatchTuning, BatchView
    if case.startswith("many_samples"):
        arrays = small_scene(2000, seed=45, log_scale_mean=-3.3, log_scale_std=0.8)
        cams, spps, frames = [small_camera(72, 40, angle=0.4), small_camera(33, 47, angle=1.9)], [20, 17], [9, 4000]
    elif case == "shade_table_overflow":
        # one 16x16 image tile, 4x4 subpixels, 40 samples of 100k tiny Gaussians
        arrays = small_scene(100000, seed=47, log_scale_mean=-5.0, log_scale_std=0.5)
        cams, spps, frames = [small_camera(16, 16, angle=0.3, supersampling_factor=4)], [40], [500]
        sc = oracle_scene(arrays)
        fb = np.stack([oracle_mod.accumulate(sc, cams[0], 500 + k, n_threads=oracle_mod.n_threads_available())[0]
                       for k in range(40)])
        won = fb[(fb >> np.uint64(32)) != np.uint64(0xFFFFFFFF)] & np.uint64(0xFFFFFFFF)
        assert len(np.unique(won)) > 2 * 2048  # the oracle's winners: 4198 distinct
    elif case == "very_wide":
        # a 4200 x 120 framebuffer: 263 columns of 16x16 tiles
        arrays = small_scene(3000, seed=48, log_scale_mean=-3.3, log_scale_std=0.8)
        cams = [small_camera(1400, 40, angle=0.2, supersampling_factor=3), small_camera(64, 48, angle=1.1)]
        spps, frames = [2, 3], [70, 80]
    else:
        arrays = small_scene(1500, seed=46, log_scale_mean=-2.6, log_scale_std=0.6)
        cams = [small_camera(640, 520, angle=0.2), small_camera(300, 200, angle=2.5, supersampling_factor=3)]
        spps, frames = [2, 1], [50, 60]
    t = to_device(arrays)
    bin_size = 16 if case.endswith("_16") else 32 if case == "wide_views" else 0
    views = [BatchView(settings_of(c), spp=n, frame_idx=f, bin_size=bin_size) for c, n, f in zip(cams, spps, frames)]
    print_case(test="batch_tile_edges", case=case)
    tiles = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views,
                          BatchTuning(tile_samples=8 if case.endswith("_8") else 0))
    scatter = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views,
                            BatchTuning(schedule=2))
    tiles.render()
    scatter.render()
    torch.cuda.synchronize()
    for v, (c, n, f) in enumerate(zip(cams, spps, frames)):
        assert_close(tiles.outputs[v], scatter.outputs[v])
        assert_close(tiles.outputs[v], plain_spp(gpu, t, c, f, n))


def test_batch_temporal_modes(gpu):
    """STATIC and REPROJECT first samples against a given history."""
    import torch
    from StochasticGSCudaBackend import History, TemporalMode
    from StochasticGSCudaBackend.batch import BatchView, render_batch
    arrays = small_scene(2500, seed=42, log_scale_mean=-3.4)
    t = to_device(arrays)
    cam0 = small_camera(80, 60, angle=0.3)
    hist = plain_spp(gpu, t, cam0, 10, 3)
    cam1 = small_camera(80, 60, angle=0.32)
    v2p = (np.asarray(cam0["w2c"], np.float64) @ np.linalg.inv(np.asarray(cam1["w2c"], np.float64))).astype(np.float32)
    views = [BatchView(settings_of(cam0), spp=2, frame_idx=20, temporal_mode=TemporalMode.STATIC, history=History(*hist)),
             BatchView(settings_of(cam1), spp=3, frame_idx=30, temporal_mode=TemporalMode.REPROJECT,
                       history=History(*hist), view_to_previous_view=v2p)]
    print_case(test="batch_temporal")
    outs = render_batch(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views)
    torch.cuda.synchronize()
    assert_close(outs[0], plain_spp(gpu, t, cam0, 20, 2, mode=1, history=hist))
    assert_close(outs[1], plain_spp(gpu, t, cam1, 30, 3, mode=2, history=hist, v2p=v2p))


def test_batch_history_chain(gpu):
    """Consecutive frames of one camera in one batch: view k reads view k-1's outputs as its
    history (resolved after it), equal to rendering the frames one after another."""
    import torch
    from StochasticGSCudaBackend import _native as N
    arrays = small_scene(2500, seed=47, log_scale_mean=-3.3)
    t = to_device(arrays)
    cams = [small_camera(72, 52, angle=0.3 + 0.03 * k) for k in range(3)]
    spps, frames, modes = [2, 1, 3], [40, 42, 43], [0, 2, 2]
    opts = dict(dtype=torch.float32, device="cuda")
    outs = [tuple(torch.empty((52, 72, c), **opts) for c in (3, 1, 1, 1)) for _ in cams]
    views = (N.sgs_view * 3)()
    v2ps = [None]
    for k in range(1, 3):
        v2ps.append((np.asarray(cams[k - 1]["w2c"], np.float64) @ np.linalg.inv(
            np.asarray(cams[k]["w2c"], np.float64))).astype(np.float32))
    for k, c in enumerate(cams):
        v = views[k]
        v.w2c[:] = np.asarray(c["w2c"], np.float32).reshape(16).tolist()
        v.cam_position[:] = np.asarray(c["cam_position"], np.float32).tolist()
        v.width, v.height = c["width"], c["height"]
        v.focal_x, v.focal_y, v.center_x, v.center_y = c["focal_x"], c["focal_y"], c["center_x"], c["center_y"]
        v.near_plane, v.far_plane = c["near_plane"], c["far_plane"]
        v.supersampling_factor = 1
        v.spp, v.frame_idx, v.temporal_mode = spps[k], frames[k], modes[k]
        v.view_to_previous_view[:] = (np.eye(4, dtype=np.float32) if v2ps[k] is None else v2ps[k]).reshape(16).tolist()
        v.out = N.sgs_history(*[o.data_ptr() for o in outs[k]])
        if k:
            v.history = N.sgs_history(*[o.data_ptr() for o in outs[k - 1]])
    keep = []
    g = N.gaussians_struct(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], keep)
    ws = N._workspace(N._lib.sgs_render_batch_workspace_size(views, 3, g.n_gaussians, None), torch.device("cuda"))
    print_case(test="batch_history_chain")
    rc = N._lib.sgs_render_batch(C.byref(g), views, 3, None, None, ws.data_ptr(), ws.numel(), None,
                                 torch.cuda.current_stream().cuda_stream)
    N._check(rc, "sgs_render_batch")
    torch.cuda.synchronize()
    hist = None
    for k, c in enumerate(cams):
        hist = plain_spp(gpu, t, c, frames[k], spps[k], mode=modes[k], history=hist, v2p=v2ps[k])
        assert_close(outs[k], hist)


@pytest.mark.parametrize("sh_degree", [0, 1, 2, 3])
def test_batch_matches_oracle(gpu, oracle_mod, sh_degree):
    """The batched resolve shades with every SH degree the reference evaluates
    (total_sh_bases 1, 4, 9, 16: gs_utils.h:212-243)."""
    import torch
    from StochasticGSCudaBackend.batch import BatchView, render_batch
    arrays = small_scene(1500, seed=43, log_scale_mean=-3.4, sh_degree=sh_degree)
    assert arrays["sh"].shape[1] == (sh_degree + 1) ** 2
    print_case(test="batch_oracle", sh_degree=sh_degree)
    sc, t = oracle_scene(arrays), to_device(arrays)
    cams = [small_camera(48, 40, angle=0.1), small_camera(40, 30, angle=1.0, supersampling_factor=2)]
    views = [BatchView(settings_of(c), spp=3, frame_idx=300 + 10 * i) for i, c in enumerate(cams)]
    outs = render_batch(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views)
    torch.cuda.synchronize()
    for i, c in enumerate(cams):
        hist = None
        for k in range(3):
            ref = oracle_mod.render(sc, c, 300 + 10 * i + k, history=hist, temporal_mode=0 if k == 0 else 1)
            hist = ref[:4]
        assert_close(outs[i], hist, exact=False)


def test_batch_graph_replay(gpu):
    """A captured batch replays with new frames identically to eager renders, with eager
    work and synchronisation between replays."""
    import torch
    from StochasticGSCudaBackend.batch import BatchRenderer, BatchView
    arrays = small_scene(2000, seed=44, log_scale_mean=-3.4)
    t = to_device(arrays)
    cams, spps, frames = make_views()
    views = [BatchView(settings_of(c), spp=n, frame_idx=f) for c, n, f in zip(cams, spps, frames)]
    graphed = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views)
    eager = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views)
    graph = graphed.capture()
    for step in range(3):
        new = [f + 1000 * (step + 1) for f in frames]
        graphed.set_frames(new)
        graphed.upload_table()
        graph.replay()
        eager.set_frames(new)
        eager.upload_table()
        eager.render()
        torch.cuda.synchronize()
        _ = (t["means"] * 2.0).sum().item()  # unrelated eager work between replays
        print_case(test="batch_graph", step=step)
        for a, b in zip(graphed.outputs, eager.outputs):
            assert_close(a, b)


@pytest.mark.parametrize("streams", [2, 3])
def test_streamed_batches_equal_one_batch(gpu, streams):
    """Views split into groups rendered by concurrent batches on their own streams give one
    batch's bits, eagerly and replayed from a graph (with eager work between replays), and
    frames taken from the process counter are drawn in view order as by one batch."""
    import torch
    from StochasticGSCudaBackend.batch import BatchRenderer, BatchView, StreamedBatchRenderer, split_views
    arrays = small_scene(3000, seed=48, log_scale_mean=-3.3, log_scale_std=0.8)
    t = to_device(arrays)
    cams, spps, frames = make_views()
    views = [BatchView(settings_of(c), spp=n, frame_idx=f) for c, n, f in zip(cams, spps, frames)]
    groups = split_views(views, streams)
    assert [i for g in groups for i in g] == list(range(len(views))) and len(groups) == streams
    print_case(test="streamed_batches", streams=streams, groups=[len(g) for g in groups])
    one = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views)
    split = StreamedBatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views,
                                  streams=streams)
    one.render()
    split.render()
    torch.cuda.synchronize()
    for a, b in zip(split.outputs, one.outputs):
        assert_close(a, b)
    graph = split.capture()
    for step in range(3):
        new = [f + 1000 * (step + 1) for f in frames]
        split.set_frames(new)
        split.upload_table()
        graph.replay()
        one.set_frames(new)
        one.upload_table()
        one.render()
        torch.cuda.synchronize()
        _ = (t["means"] * 2.0).sum().item()  # unrelated eager work between replays
        for a, b in zip(split.outputs, one.outputs):
            assert_close(a, b)
    # frames from the counter: consecutive spp-long runs in view order, as for one batch
    auto = [BatchView(settings_of(c), spp=n) for c, n in zip(cams, spps)]
    drawn_one = [int(v.frame_idx) for v in BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"],
                                                         t["sh"], auto).c_views[:len(auto)]]
    r = StreamedBatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], auto,
                              streams=streams)
    drawn = [0] * len(auto)
    for g, part in zip(r.groups, r.parts):
        for j, i in enumerate(g):
            drawn[i] = int(part.c_views[j].frame_idx)
    steps = [drawn_one[i + 1] - drawn_one[i] for i in range(len(auto) - 1)]
    assert steps == spps[:-1]
    assert [drawn[i + 1] - drawn[i] for i in range(len(auto) - 1)] == steps


def test_batch_rejects_incomplete_feature_views(gpu):
    """Spatial reuse without the view's k-buffer, a k-buffer of the wrong size and a blue-noise
    volume of the wrong size are refused before anything runs."""
    import torch
    from StochasticGSCudaBackend.batch import BatchView, render_batch
    arrays = small_scene(10)
    t = to_device(arrays)
    s = settings_of(small_camera(16, 16))._replace(spatial_reuse=True)
    args = (t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"])
    with pytest.raises(ValueError, match="k-buffer"):
        render_batch(*args, [BatchView(s)])
    small = (torch.full((4, 10, 2), -1, dtype=torch.int32, device="cuda"), torch.zeros((4, 10, 4), device="cuda"))
    with pytest.raises(ValueError, match="k-buffer"):
        render_batch(*args, [BatchView(s, k_buffer=small)])
    with pytest.raises(ValueError, match="blue_noise"):
        render_batch(*args, [BatchView(settings_of(small_camera(16, 16)),
                                       blue_noise=torch.zeros(100, dtype=torch.uint8, device="cuda"))])


@pytest.mark.parametrize("case", ["no_gaussians", "all_culled", "ragged_views"])
def test_batch_degenerate_inputs(gpu, case):
    """A scene without Gaussians, views in which every Gaussian is culled by the near plane (no
    instances at all), and views smaller than one raster tile or one pixel wide or high: the tile
    schedule equals the scatter schedule and the plain SPP loop."""
    import torch
    from StochasticGSCudaBackend.batch import BatchRenderer, BatchTuning, BatchView
    arrays = small_scene(2500, seed=50, log_scale_mean=-3.3, log_scale_std=0.8)
    if case == "no_gaussians":
        arrays = {k: v[:0] for k, v in arrays.items()}
    t = to_device(arrays)
    if case == "ragged_views":
        cams = [small_camera(1, 1, angle=0.3), small_camera(17, 1, angle=1.0), small_camera(1, 33, angle=1.7),
                small_camera(15, 15, angle=2.4, supersampling_factor=2), small_camera(47, 5, angle=3.1)]
        spps, frames = [1, 3, 16, 2, 9], [5, 6, 7, 8, 9]
    else:
        cams = [small_camera(40, 30, angle=0.6), small_camera(33, 47, angle=2.0)]
        spps, frames = [4, 16], [31, 32]
        if case == "all_culled":
            for c in cams:
                c["near_plane"], c["far_plane"] = 50.0, 100.0  # the scene lies within a few units of the origin
    views = [BatchView(settings_of(c), spp=k, frame_idx=f) for c, k, f in zip(cams, spps, frames)]
    print_case(test="batch_degenerate", case=case)
    tiles = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views)
    scatter = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views,
                            BatchTuning(schedule=2))
    tiles.render()
    scatter.render()
    torch.cuda.synchronize()
    n_inst, pairs, overflow = tiles.tile_stats()
    assert not overflow
    if case != "ragged_views":
        assert n_inst == 0 and pairs == 0
    for v, (c, k, f) in enumerate(zip(cams, spps, frames)):
        assert_close(tiles.outputs[v], scatter.outputs[v])
        assert_close(tiles.outputs[v], plain_spp(gpu, t, c, f, k))


@pytest.mark.parametrize("n", [6000, 15000, 45000])
def test_batch_sort_long_bins(gpu, n):
    """Views of one 16-pixel bin, so that the bin's list holds about every Gaussian: the
    per-bin sort keeps up to 8192 entries in registers, the buckets of the next 12288 in LDS and
    reads those of any further entries twice; lists within each range (one- and 16-sample
    views, whole and half raster tiles) equal the scatter schedule and the plain SPP loop."""
    import torch
    from StochasticGSCudaBackend.batch import BatchRenderer, BatchTuning, BatchView
    arrays = small_scene(n, seed=49, log_scale_mean=-4.5, log_scale_std=0.4)
    t = to_device(arrays)
    cams, spps, frames = [small_camera(16, 16, angle=0.5), small_camera(16, 16, angle=2.0)], [16, 1], [21, 77]
    views = [BatchView(settings_of(c), spp=k, frame_idx=f, bin_size=16) for c, k, f in zip(cams, spps, frames)]
    print_case(test="batch_sort_long_bins", n=n)
    tiles = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views)
    scatter = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views,
                            BatchTuning(schedule=2))
    tiles.render()
    scatter.render()
    torch.cuda.synchronize()
    n_inst, pairs, overflow = tiles.tile_stats()
    assert not overflow and n_inst == pairs
    # per view about n entries: below 8192, within the LDS stage, and beyond it
    lo, hi = {6000: (1, 8192), 15000: (8193, 20480), 45000: (20481, 1 << 30)}[n]
    assert 2 * lo <= n_inst <= 2 * hi
    for v, (c, k, f) in enumerate(zip(cams, spps, frames)):
        assert_close(tiles.outputs[v], scatter.outputs[v])
        assert_close(tiles.outputs[v], plain_spp(gpu, t, c, f, k))


def test_batch_bins_without_subrectangles(gpu):
    """More Gaussians than a 64-pixel bin's instance word leaves index bits for (2^24): the
    fill stores no tile sub-rectangles and every tile of a bin takes every instance; the
    result still equals the scatter schedule and the plain SPP loop (16-pixel bins, no
    sub-rectangles either way)."""
    import torch
    from StochasticGSCudaBackend.batch import BatchRenderer, BatchTuning, BatchView
    n = (1 << 24) + 1
    arrays = small_scene(n, seed=48, log_scale_mean=-6.5, log_scale_std=0.5, sh_degree=0)
    t = to_device(arrays)
    del arrays
    cams, spps, frames = [small_camera(80, 60, angle=0.7), small_camera(40, 30, angle=2.2)], [3, 1], [11, 90]
    views = [BatchView(settings_of(c), spp=k, frame_idx=f, bin_size=64) for c, k, f in zip(cams, spps, frames)]
    print_case(test="batch_no_subrect", n=n)
    tiles = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views)
    scatter = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views,
                            BatchTuning(schedule=2))
    tiles.render()
    scatter.render()
    torch.cuda.synchronize()
    n_inst, pairs, overflow = tiles.tile_stats()
    assert not overflow and n_inst > 0
    for v, (c, k, f) in enumerate(zip(cams, spps, frames)):
        assert_close(tiles.outputs[v], scatter.outputs[v])
        assert_close(tiles.outputs[v], plain_spp(gpu, t, c, f, k))
    del tiles, scatter, t
    torch.cuda.empty_cache()
