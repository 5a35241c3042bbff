"""GPU tests of the autotuner: the tuned configuration renders the same bits as the
defaults (the searched knobs are scheduling only), and every trial is timed."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import small_camera, small_scene
from test_gpu_batch import assert_close, settings_of
from test_gpu_parity import print_case, to_device

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from StochasticGSCudaBackend import _native
    return _native


@pytest.mark.parametrize("gps", [False, True])
def test_tuned_render_is_bit_identical(gpu, gps):
    import torch
    from StochasticGSCudaBackend.autotune import tune_render
    arrays = small_scene(4000, seed=61, log_scale_mean=-3.0, log_scale_std=0.9)
    t = to_device(arrays)
    cam = small_camera(160, 120, angle=0.4)
    s = settings_of(cam)._replace(hybrid_scheduling=True, use_gps=gps)
    print_case(test="autotune_render", gps=gps)
    res = tune_render(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], s, frame_idx=5, repeats=1,
                      passes=1)
    assert res.trials and all(ms > 0 for _, ms in res.trials)
    args = (t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], s.w2c.cuda(), s.cam_position.cuda(),
            *s.as_tuple()[2:])
    a = gpu.accumulate(*args, frame_idx=9).clone()
    b = gpu.accumulate(*args, frame_idx=9, tuning=res.best)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_tuned_batch_is_bit_identical(gpu):
    from StochasticGSCudaBackend.autotune import tune_batch
    from StochasticGSCudaBackend.batch import BatchRenderer, BatchView
    arrays = small_scene(3000, seed=62, log_scale_mean=-3.3, log_scale_std=0.8)
    t = to_device(arrays)
    cams = [small_camera(96, 64, angle=0.1), small_camera(64, 64, angle=1.4)]
    views = [BatchView(settings_of(c), spp=s, frame_idx=f) for c, s, f in zip(cams, (3, 9), (11, 500))]
    print_case(test="autotune_batch")
    res = tune_batch(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views, repeats=1)
    assert len(res.trials) >= 3
    tuned = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views, res.best).render()
    ref = BatchRenderer(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views).render()
    for a, b in zip(tuned, ref):
        assert_close(a, b)
