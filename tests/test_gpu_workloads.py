"""GPU parity at the full size of the benchmark workloads.

* 100k Gaussians in U[-1,1]^3, 800x600, 4 samples per pixel, stochastic transparency: every
  sample's full framebuffer bit-exact against the oracle (tile and scatter schedules), the
  RESET + STATIC accumulation within 1e-5.
* 6M Gaussians (8-component mixture, Morton order), 64 views of 256x256..2048x1152 with
  spp in {1, 4, 16}, rendered as two concurrent view groups: every view bit-equal to the plain
  SPP loop of render() calls (Renderer.py:85-87); for a few views 256x256 framebuffer crops
  bit-exact against the oracle and the resolve of the full framebuffer within 1e-5.
* Serving over the same scene: 32 camera streams at 720p/1080p/1440p, spp in {1, 2, 8},
  Poisson arrivals, the spatio-temporal resolve (spatial reuse over each stream's k-buffer,
  REPROJECT against its history), two batches in flight: every frame of a sample of streams
  equals rendering that stream alone, outputs and k-buffers bit for bit.
"""
from __future__ import annotations

import math
import time

import numpy as np
import pytest

from conftest import oracle_scene
from test_gpu_batch import assert_close, plain_spp, settings_of
from test_gpu_batch_features import empty_k_buffer, plain_loop
from test_gpu_parity import _gpu_render, gpu_fb, print_case, to_device

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


@pytest.fixture(scope="module")
def scene6m(gpu):
    """The 6M-Gaussian scene on the host (for the oracle) and on the device."""
    import torch
    from StochasticGSCudaBackend.synthetic import make_scene, morton_order, reorder, scene_c4
    t0 = time.perf_counter()
    arrays = make_scene(scene_c4())
    arrays = reorder(arrays, morton_order(arrays["means"]))
    dev = to_device(arrays)
    torch.cuda.synchronize()
    print(f"6M scene ready in {time.perf_counter() - t0:.1f} s", flush=True)
    yield arrays, dev
    del dev
    torch.cuda.empty_cache()


def _threads(oracle_mod):
    return min(16, oracle_mod.n_threads_available())


def test_c1_full_frame(gpu, oracle_mod):
    """100k Gaussians, 800x600, SPP 4: full framebuffers bit-exact, resolve within 1e-5."""
    from StochasticGSCudaBackend.synthetic import camera_c1, make_scene, scene_c1
    arrays = make_scene(scene_c1())
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = camera_c1()
    hist_o = None
    for k, frame in enumerate(range(20, 24)):
        mode = 0 if k == 0 else 1
        print_case(test="c1", spp_index=k, frame=frame)
        ref = oracle_mod.render(sc, cam, frame, history=hist_o, temporal_mode=mode, n_threads=_threads(oracle_mod))
        got, fb = _gpu_render(gpu, t, cam, frame, hist=hist_o, mode=mode)
        assert np.array_equal(fb, ref[4]), f"{np.count_nonzero(fb != ref[4])} of {fb.size} pixels differ"
        scatter = gpu_fb(gpu, t, cam, frame, tuning=gpu.Tuning(schedule=2))
        assert np.array_equal(scatter, ref[4])
        for g, r, name in zip(got, ref[:4], ("image", "depth", "alpha", "weight_mass")):
            err = float(np.abs(g.astype(np.float64) - r.astype(np.float64)).max())
            assert err <= RESOLVE_ATOL, f"{name}: {err}"
        hist_o = ref[:4]
    assert (fb != EMPTY).mean() > 0.2


def test_c4_batched_full_size(gpu, oracle_mod, scene6m):
    """6M Gaussians x 64 views in two concurrent view groups: equal to the plain SPP loop view
    by view; oracle crops and full-frame resolves for a few views."""
    import torch
    from StochasticGSCudaBackend.batch import BatchView, StreamedBatchRenderer
    from StochasticGSCudaBackend.synthetic import views_c4
    arrays, t = scene6m
    cams = views_c4(64)
    frames, f = [], 1000
    for c in cams:
        frames.append(f)
        f += c["spp"]
    views = [BatchView(settings_of(c), spp=c["spp"], frame_idx=fr) for c, fr in zip(cams, frames)]
    print_case(test="c4_batched", views=len(views), pixel_samples=sum(c["width"] * c["height"] * c["spp"] for c in cams))
    r = StreamedBatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views, streams=2)
    outs = r.render()
    torch.cuda.synchronize()
    n_inst, pairs, overflow = r.tile_stats()
    print(f"instances {n_inst}, visible pairs {pairs}, overflow {overflow}", flush=True)
    for v, (c, fr) in enumerate(zip(cams, frames)):
        ref = plain_spp(gpu, t, c, fr, c["spp"])
        torch.cuda.synchronize()
        assert_close(outs[v], ref)
    print("c4: all 64 views equal the plain SPP loop", flush=True)
    sc = oracle_scene(arrays)
    spp1 = [v for v, c in enumerate(cams) if c["spp"] == 1]
    big = max(range(64), key=lambda v: cams[v]["width"] * cams[v]["height"])
    checked = 0
    for v in ([spp1[0]] if spp1 else []) + [big]:
        c = cams[v]
        W, H = c["width"], c["height"]
        x0, y0 = (W - 256) // 2, (H - 256) // 2
        crop = (x0, y0, x0 + 256, y0 + 256)
        for fr in sorted({frames[v], frames[v] + c["spp"] - 1}):
            print_case(test="c4_oracle_crop", view=v, width=W, height=H, frame=fr)
            fb = gpu_fb(gpu, t, c, fr)
            ref, _ = oracle_mod.accumulate(sc, c, fr, crop=crop, n_threads=_threads(oracle_mod))
            assert np.array_equal(fb[y0:y0 + 256, x0:x0 + 256], ref)
            assert (ref != EMPTY).mean() > 0.3
            checked += 1
        if c["spp"] == 1:
            # the whole view: the oracle shades and resolves the (crop-pinned) framebuffer
            print_case(test="c4_oracle_resolve", view=v)
            fb = gpu_fb(gpu, t, c, frames[v])
            ref = oracle_mod.shade_resolve(sc, c, fb)
            for g, rr, name in zip(outs[v], ref, ("image", "depth", "alpha", "weight_mass")):
                err = float(np.abs(g.cpu().numpy().astype(np.float64) - rr.astype(np.float64)).max())
                assert err <= RESOLVE_ATOL, f"view {v} {name}: {err}"
    assert checked >= 2


def test_c5_serving_full_size(gpu, scene6m):
    """32 streams over the 6M scene, a few hundred Poisson requests, two batches in flight:
    each sampled stream's frames equal the stream rendered alone with spatial reuse over its
    own k-buffer (RESET, then REPROJECT against its previous frame), and the served k-buffer
    equals the one the alone loop leaves."""
    import torch
    from StochasticGSCudaBackend.serve import Request, Server, StreamSpec
    from StochasticGSCudaBackend.synthetic import camera, orbit_eye
    _, t = scene6m
    rng = np.random.default_rng(5)
    sizes = [(1280, 720), (1920, 1080), (2560, 1440)]
    n_streams, n_req = 32, 256
    specs, paths = [], []
    for s in range(n_streams):
        w, h = sizes[int(rng.integers(3))]
        c = camera(orbit_eye(0.0, 3.0), w, h)
        specs.append(StreamSpec(w, h, c["focal_x"], c["focal_y"], c["center_x"], c["center_y"], c["near_plane"],
                                c["far_plane"], spatial_reuse=True))
        paths.append((rng.uniform(0, 2 * math.pi), rng.uniform(0.2, 0.6) * rng.choice([-1, 1]),
                      rng.uniform(2.6, 3.4)))
    times = np.cumsum(rng.exponential(1.0 / 300.0, n_req))
    reqs, cams = [], {}
    for tt in times:
        s = int(rng.integers(n_streams))
        phase, omega, radius = paths[s]
        c = camera(orbit_eye(phase + omega * tt, radius, 0.1 * math.sin(tt)), specs[s].width, specs[s].height,
                   spatial_reuse=True)
        reqs.append(Request(stream=s, w2c=c["w2c"], cam_position=c["cam_position"], spp=int(rng.choice([1, 2, 8])),
                            arrival=float(tt)))
        cams[id(reqs[-1])] = c
    print_case(test="c5_serving", streams=n_streams, requests=n_req)
    server = Server(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], specs, max_views=32, inflight=2,
                    chain=3, max_wait=0.01, keep_outputs=True)
    done = server.run(reqs)
    torch.cuda.synchronize()
    assert len(done) == n_req and all(r.output is not None for r in done)
    assert server.overflowed_batches == 0
    print(f"served {len(done)} requests in {server.batches} batches (max {max(server.batch_sizes)} views)", flush=True)
    assert max(server.batch_sizes) > 1
    by_stream: dict[int, list] = {}
    for r in done:
        by_stream.setdefault(r.stream, []).append(r)
    sample = sorted(by_stream, key=lambda s: -len(by_stream[s]))[:4]
    for s in sample:
        frames = sorted(by_stream[s], key=lambda r: r.frame)
        hist, prev = None, None
        kb = empty_k_buffer(cams[id(frames[0])])
        for r in frames:
            c = cams[id(r)]
            v2p = None
            if prev is not None:
                v2p = (np.asarray(prev["w2c"], np.float64) @ np.linalg.inv(np.asarray(c["w2c"], np.float64))).astype(
                    np.float32)
            print_case(test="c5_alone", stream=s, frame=r.frame, spp=r.spp, width=c["width"])
            ref = plain_loop(gpu, t, c, r.frame_idx, r.spp, mode=0 if r.frame == 0 else 2, history=hist, v2p=v2p,
                             kb=kb)
            torch.cuda.synchronize()
            assert_close(r.output, ref)
            hist, prev = ref, c
        k_served = server.streams[s].k_buffer
        assert torch.equal(k_served[0], kb[0]) and torch.equal(k_served[1], kb[1])
    del server
    torch.cuda.empty_cache()
