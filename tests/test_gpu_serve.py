"""GPU tests of the serving loop: every request's output, rendered in a mix of streams with
continuous batching and two batches in flight, equals rendering its stream alone frame by
frame (plain path, RESET then REPROJECT against the previous frame), bit for bit."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import small_scene
from test_gpu_batch import assert_close, plain_spp
from test_gpu_parity import print_case, to_device

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from StochasticGSCudaBackend import _native
    return _native


def stream_camera(stream: int, frame: int, size):
    from StochasticGSCudaBackend.synthetic import camera, orbit_eye
    angle = 0.9 * stream + 0.04 * frame  # smooth trajectory per stream
    eye = orbit_eye(angle, 3.0 + 0.2 * math.sin(0.3 * frame + stream), 0.1 * math.cos(0.2 * frame))
    return camera(eye, size[0], size[1])


def spec_of(cam):
    from StochasticGSCudaBackend.serve import StreamSpec
    return StreamSpec(cam["width"], cam["height"], cam["focal_x"], cam["focal_y"], cam["center_x"], cam["center_y"],
                      cam["near_plane"], cam["far_plane"])


@pytest.mark.parametrize("max_views,inflight,chain,burst", [(3, 2, 3, False), (8, 1, 3, True), (1, 2, 1, False),
                                                             (16, 2, 2, True)])
def test_mix_equals_alone(gpu, max_views, inflight, chain, burst):
    from StochasticGSCudaBackend.serve import Request, Server
    arrays = small_scene(2500, seed=51, log_scale_mean=-3.3, log_scale_std=0.7)
    t = to_device(arrays)
    sizes = [(64, 48), (80, 60), (50, 40), (72, 56)]
    n_frames = 4
    rng = np.random.default_rng(7)
    cams = {(s, f): stream_camera(s, f, sizes[s]) for s in range(len(sizes)) for f in range(n_frames)}
    spps = {(s, f): int(rng.choice([1, 2, 3])) for s in range(len(sizes)) for f in range(n_frames)}
    # burst: everything queued at once, so consecutive frames of a stream share batches
    arrivals = np.zeros(len(cams)) if burst else np.sort(rng.uniform(0.0, 0.01, len(cams)))
    order = sorted(cams, key=lambda k: (k[1], k[0]))  # frame-major so every stream is in flight
    reqs = []
    for (s, f), a in zip(order, arrivals):
        c = cams[(s, f)]
        reqs.append(Request(stream=s, w2c=c["w2c"], cam_position=c["cam_position"], spp=spps[(s, f)], arrival=float(a)))
    print_case(test="serve_mix", max_views=max_views, inflight=inflight, chain=chain, burst=burst)
    server = Server(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"],
                    [spec_of(cams[(s, 0)]) for s in range(len(sizes))], max_views=max_views, inflight=inflight,
                    chain=chain, arena_bytes=256 << 20, keep_outputs=True)
    done = server.run(reqs)
    assert len(done) == len(reqs) and all(r.output is not None for r in done)
    if max_views > 1 and burst:
        # with every request queued before the first dispatch, batches of several form for
        # certain; with spread arrivals that depends on render speed against the wall clock
        assert max(server.batch_sizes) > 1, "no batching happened"
    by_key = {(r.stream, r.frame): r for r in done}
    for s in range(len(sizes)):
        hist, prev = None, None
        used = 0
        for f in range(n_frames):
            r = by_key[(s, f)]
            c = cams[(s, f)]
            assert r.frame_idx == (s << 24) + used
            v2p = None
            if prev is not None:
                v2p = (np.asarray(prev["w2c"], np.float64) @ np.linalg.inv(np.asarray(c["w2c"], np.float64))).astype(
                    np.float32)
            ref = plain_spp(gpu, t, c, r.frame_idx, r.spp, mode=0 if f == 0 else 2, history=hist, v2p=v2p)
            assert_close(r.output, ref)
            hist, prev = ref, c
            used += r.spp
