"""The batched renderer and the server with every single-view feature: GPS (with and
without a point cap, supersampling), blue noise, hybrid scheduling, spatial reuse (its
k-buffer updated sample by sample), proper antialiasing, temporal chains. Every view equals
spp successive render() calls of the plain path with the same frames (Renderer.py:85-87),
outputs and k-buffers bit for bit; served streams equal the stream rendered alone."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import small_camera, small_scene
from test_gpu_batch import assert_close
from test_gpu_parity import gpu_args, print_case, to_device

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from StochasticGSCudaBackend import _native
    return _native


def full_settings(cam):
    import torch
    from StochasticGSCudaBackend import Settings
    return Settings(torch.from_numpy(np.asarray(cam["w2c"], np.float32)),
                    torch.from_numpy(np.asarray(cam["cam_position"], np.float32)), cam["width"], cam["height"],
                    cam["focal_x"], cam["focal_y"], cam["center_x"], cam["center_y"], cam["near_plane"],
                    cam["far_plane"], cam.get("proper_antialiasing", False), cam.get("supersampling_factor", 1),
                    cam.get("use_gps", False), cam.get("hybrid_scheduling", False), cam.get("spatial_reuse", False))


def empty_k_buffer(cam):
    import torch
    ss = max(1, int(cam.get("supersampling_factor", 1)))
    P = cam["width"] * cam["height"] * ss * ss
    return (torch.full((4, P, 2), -1, dtype=torch.int32, device="cuda"), torch.zeros((4, P, 4), device="cuda"))


def plain_loop(gpu, t, cam, frame0, spp, mode=0, history=None, v2p=None, kb=None, blue=None, max_points=0):
    """spp render() calls of the single-view path with every flag of ``cam``."""
    import torch
    eye = torch.from_numpy(np.asarray(np.eye(4) if v2p is None else v2p, np.float32)).cuda()
    hist = tuple(torch.empty(0) for _ in range(4)) if history is None else history
    if kb is None:
        kb = (torch.empty((0, 0, 2), dtype=torch.int32, device="cuda"), torch.empty((0, 0, 4), device="cuda"))
    bn = torch.empty(0, dtype=torch.uint8, device="cuda") if blue is None else blue
    tun = gpu.Tuning(max_gps_points=max_points) if max_points else None
    for k in range(spp):
        hist = gpu.render(*gpu_args(t, cam), bn, kb[0], kb[1], *hist, eye, mode if k == 0 else 1,
                          frame_idx=frame0 + k, tuning=tun)
    return hist


FEATURES = {
    "gps": dict(use_gps=True),
    "gps_ss2_cap": dict(use_gps=True, supersampling_factor=2, max_gps_points=32),
    "blue": dict(blue=True),
    "hybrid_st": dict(hybrid_scheduling=True),
    "hybrid_gps": dict(use_gps=True, hybrid_scheduling=True),
    "reuse": dict(spatial_reuse=True),
    "reuse_aa_ss2": dict(spatial_reuse=True, proper_antialiasing=True, supersampling_factor=2),
    "reuse_blue_gps": dict(spatial_reuse=True, use_gps=True),
    "reuse_blue": dict(spatial_reuse=True, blue=True),
}


@pytest.mark.parametrize("name", sorted(FEATURES))
def test_batch_feature_equals_plain_loop(gpu, name):
    """A batch of a plain view and two views with the feature (different sizes, spp and
    frames, the second starting from a history) equals the plain path view by view."""
    import torch
    from StochasticGSCudaBackend.batch import BatchRenderer, BatchView
    feat = dict(FEATURES[name])
    blue = feat.pop("blue", False)
    max_points = feat.pop("max_gps_points", 0)
    arrays = small_scene(2500, seed=91, log_scale_mean=-3.3, log_scale_std=0.8)
    t = to_device(arrays)
    vol = torch.from_numpy(np.random.default_rng(5).integers(0, 256, (64, 128, 128), dtype=np.uint8)).cuda() \
        if blue else None
    cams = [small_camera(64, 48, angle=0.1), small_camera(96, 72, angle=0.8, **feat),
            small_camera(80, 50, angle=-0.6, **feat)]
    spps, frames = [3, 4, 2], [11, 5000, 70]
    # view 2 continues from a history (STATIC) rendered by the plain path
    start = plain_loop(gpu, t, cams[2], 900, 2, kb=empty_k_buffer(cams[2]) if feat.get("spatial_reuse") else None,
                       blue=vol, max_points=max_points)
    torch.cuda.synchronize()
    modes, hists = [0, 0, 1], [None, None, start]
    kbs = [empty_k_buffer(c) if c.get("spatial_reuse") else None for c in cams]
    kbs_plain = [None if k is None else tuple(x.clone() for x in k) for k in kbs]
    views = [BatchView(full_settings(c), spp=n, frame_idx=f, temporal_mode=m, history=h, k_buffer=kb,
                       blue_noise=vol if c is not cams[0] else None, max_gps_points=max_points if c is not cams[0] else 0)
             for c, n, f, m, h, kb in zip(cams, spps, frames, modes, hists, kbs)]
    print_case(test="batch_feature", name=name)
    r = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views)
    outs = r.render()
    torch.cuda.synchronize()
    for v, c in enumerate(cams):
        ref = plain_loop(gpu, t, c, frames[v], spps[v], mode=modes[v], history=hists[v], kb=kbs_plain[v],
                         blue=vol if v else None, max_points=max_points if v else 0)
        torch.cuda.synchronize()
        assert_close(outs[v], ref)
        if kbs[v] is not None:
            assert torch.equal(kbs[v][0], kbs_plain[v][0]) and torch.equal(kbs[v][1], kbs_plain[v][1])


@pytest.mark.parametrize("gps", [False, True])
def test_reuse_stream_chain_in_one_batch(gpu, gps):
    """Three consecutive frames of one stream in one batch (RESET, then REPROJECT against the
    previous frame's outputs), sharing the stream's k-buffer: as the plain path frame by frame."""
    import torch
    from StochasticGSCudaBackend.batch import BatchRenderer, BatchView
    arrays = small_scene(2500, seed=92, log_scale_mean=-3.3)
    t = to_device(arrays)
    cams = [small_camera(72, 56, angle=0.05 * k, spatial_reuse=True, use_gps=gps) for k in range(3)]
    v2ps = [None] + [(np.asarray(cams[k - 1]["w2c"], np.float64) @ np.linalg.inv(np.asarray(cams[k]["w2c"],
                                                                                               np.float64))).astype(
        np.float32) for k in range(1, 3)]
    kb = empty_k_buffer(cams[0])
    kb_plain = tuple(x.clone() for x in kb)
    r_views, outs = [], []
    opts = dict(dtype=torch.float32, device="cuda")
    bufs = [tuple(torch.empty((56, 72, c), **opts) for c in (3, 1, 1, 1)) for _ in range(3)]
    from StochasticGSCudaBackend import History
    for k in range(3):
        r_views.append(BatchView(full_settings(cams[k]), spp=2 + k, frame_idx=100 * k, temporal_mode=0 if k == 0 else 2,
                                 history=None if k == 0 else History(*bufs[k - 1]), view_to_previous_view=v2ps[k],
                                 k_buffer=kb))
    print_case(test="reuse_chain", gps=gps)
    r = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], r_views)
    # outputs of frame k are the history of frame k + 1: point them at the shared buffers
    import ctypes as C
    for k in range(3):
        r.outputs[k] = History(*bufs[k])
        r.c_views[k].out = gpu.sgs_history(*[b.data_ptr() for b in bufs[k]])
        if k:
            r.c_views[k].history = gpu.sgs_history(*[b.data_ptr() for b in bufs[k - 1]])
    r.upload_table()
    r.render()
    torch.cuda.synchronize()
    hist = None
    for k in range(3):
        hist = plain_loop(gpu, t, cams[k], 100 * k, 2 + k, mode=0 if k == 0 else 2, history=hist, v2p=v2ps[k],
                          kb=kb_plain)
        torch.cuda.synchronize()
        assert_close(History(*bufs[k]), hist)
    assert torch.equal(kb[0], kb_plain[0]) and torch.equal(kb[1], kb_plain[1])


def test_served_reuse_streams_equal_alone(gpu):
    """Streams with the spatio-temporal resolve (spatial reuse + REPROJECT) served in a mix
    with continuous batching equal each stream rendered alone, k-buffers included."""
    import torch
    from StochasticGSCudaBackend.serve import Request, Server, StreamSpec
    from StochasticGSCudaBackend.synthetic import camera, orbit_eye
    arrays = small_scene(2500, seed=93, log_scale_mean=-3.3, log_scale_std=0.7)
    t = to_device(arrays)
    sizes = [(64, 48), (80, 60), (56, 40)]
    n_frames = 4
    rng = np.random.default_rng(8)
    cams = {(s, f): camera(orbit_eye(0.9 * s + 0.04 * f, 3.0), *sizes[s], spatial_reuse=True, use_gps=(s == 2))
            for s in range(3) for f in range(n_frames)}
    spps = {k: int(rng.choice([1, 2, 3])) for k in cams}
    order = sorted(cams, key=lambda k: (k[1], k[0]))
    reqs = [Request(stream=s, w2c=cams[(s, f)]["w2c"], cam_position=cams[(s, f)]["cam_position"], spp=spps[(s, f)],
                    arrival=0.0) for s, f in order]
    specs = [StreamSpec(*sizes[s], cams[(s, 0)]["focal_x"], cams[(s, 0)]["focal_y"], cams[(s, 0)]["center_x"],
                        cams[(s, 0)]["center_y"], cams[(s, 0)]["near_plane"], cams[(s, 0)]["far_plane"],
                        spatial_reuse=True, use_gps=(s == 2)) for s in range(3)]
    print_case(test="serve_reuse")
    server = Server(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], specs, max_views=6,
                    inflight=2, chain=2, arena_bytes=256 << 20, keep_outputs=True)
    done = server.run(reqs)
    by_key = {(r.stream, r.frame): r for r in done}
    for s in range(3):
        kb = empty_k_buffer(cams[(s, 0)])
        hist, prev = None, None
        for f in range(n_frames):
            r = by_key[(s, f)]
            c = cams[(s, f)]
            v2p = None if prev is None else (np.asarray(prev["w2c"], np.float64) @ np.linalg.inv(
                np.asarray(c["w2c"], np.float64))).astype(np.float32)
            ref = plain_loop(gpu, t, c, r.frame_idx, r.spp, mode=0 if f == 0 else 2, history=hist, v2p=v2p, kb=kb)
            torch.cuda.synchronize()
            assert_close(r.output, ref)
            hist, prev = ref, c
        assert torch.equal(server.streams[s].k_buffer[0], kb[0]) and torch.equal(server.streams[s].k_buffer[1], kb[1])
