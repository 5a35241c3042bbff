"""Convergence of the GPU path's SPP average to the FP64 sorted front-to-back blend.

One stochastic-transparency sample is an unbiased estimate of the blend of every fragment
it can accept, sorted by (depth, index) (the oracle's orc_blend_fp64, FP64); the average of
spp samples (the batched renderer's RESET + STATIC accumulation, Renderer.py:85-87) must
approach it with RMSE proportional to 1/sqrt(spp). The curve is printed (-s) for the record."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import oracle_scene, small_camera, small_scene
from test_gpu_batch import settings_of
from test_gpu_parity import print_case, to_device

pytestmark = pytest.mark.gpu
SPPS = (1, 4, 16, 64, 256)


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from StochasticGSCudaBackend import _native
    return _native


@pytest.mark.parametrize("ss", [1, 2])
def test_spp_average_converges_to_fp64_blend(gpu, oracle_mod, ss):
    import torch
    from StochasticGSCudaBackend.batch import BatchView, render_batch
    arrays = small_scene(3000, seed=71, log_scale_mean=-3.4, log_scale_std=0.6, opacity_std=1.0)
    sc, t = oracle_scene(arrays), to_device(arrays)
    cam = small_camera(96, 72, supersampling_factor=ss)
    ref_img, ref_dep, ref_alp = oracle_mod.blend_fp64(sc, cam)
    assert ref_alp.mean() > 0.2
    views = [BatchView(settings_of(cam), spp=n, frame_idx=100_000 * (k + 1)) for k, n in enumerate(SPPS)]
    outs = render_batch(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], views)
    torch.cuda.synchronize()
    rows = []
    for n, o in zip(SPPS, outs):
        img = o.image.cpu().numpy().astype(np.float64)
        alp = o.alpha.cpu().numpy()[..., 0].astype(np.float64)
        assert np.allclose(o.weight_mass.cpu().numpy(), n)
        e_img = math.sqrt(float(np.mean((img - ref_img) ** 2)))
        e_alp = math.sqrt(float(np.mean((alp - ref_alp) ** 2)))
        rows.append((n, e_img, e_alp))
    print_case(test="convergence", ss=ss)
    for n, e_img, e_alp in rows:
        print(f"spp {n:4d}: rmse image {e_img:.5f} (x sqrt(spp) {e_img * math.sqrt(n):.4f}), "
              f"alpha {e_alp:.5f} (x sqrt(spp) {e_alp * math.sqrt(n):.4f})", flush=True)
    spp = np.log([r[0] for r in rows])
    for col in (1, 2):
        err = np.log([r[col] for r in rows])
        slope = np.polyfit(spp, err, 1)[0]
        assert -0.6 < slope < -0.4, f"log-log slope {slope}"
        scaled = np.array([r[col] * math.sqrt(r[0]) for r in rows])
        assert scaled.max() / scaled.min() < 1.5
    assert rows[-1][1] < 0.02
