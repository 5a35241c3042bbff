"""sgs_render_noalloc (the allocation-free variant of sgs_render, fixed device scratch)
against sgs_render through the C ABI: identical framebuffers and outputs, k-buffer updates
and REPROJECT included, on consecutive calls whose second input is smaller (the scratch
keeps the first call's contents), and host-side rejection of views beyond the scratch."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import small_camera, small_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from StochasticGSCudaBackend import _native
    return _native


def _dev(arrays):
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in arrays.items()}


def _outputs(H, W):
    import torch
    return [torch.full((H, W, c), float("nan"), device="cuda") for c in (3, 1, 1, 1)]


def _call(native, entry, t, cam, frame, mode, hist, v2p, kb):
    """One render through the C ABI with ``entry`` ("sgs_render" or "sgs_render_noalloc")."""
    import torch
    keep = []
    g = native.gaussians_struct(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], keep)
    s = native.settings_struct(cam["w2c"], cam["cam_position"], cam["width"], cam["height"], cam["focal_x"],
                               cam["focal_y"], cam["center_x"], cam["center_y"], cam["near_plane"], cam["far_plane"],
                               cam.get("proper_antialiasing", False), cam.get("supersampling_factor", 1),
                               cam.get("use_gps", False), cam.get("hybrid_scheduling", False),
                               cam.get("spatial_reuse", False), torch.device("cuda"), keep)
    H, W = cam["height"], cam["width"]
    outs = _outputs(H, W)
    out = native.sgs_history(*[o.data_ptr() for o in outs])
    h = native.sgs_history(*[x.data_ptr() for x in hist]) if hist is not None else None
    v2p_t = torch.from_numpy(np.asarray(v2p, np.float32)).cuda()
    kp, ko = (kb[0].data_ptr(), kb[1].data_ptr()) if kb is not None else (None, None)
    stream = torch.cuda.current_stream().cuda_stream
    args = (C.byref(g), C.byref(s), None, 0, kp, ko, C.byref(h) if h is not None else None, v2p_t.data_ptr(), mode,
            C.byref(out), frame, None)
    if entry == "sgs_render":
        n = native._lib.sgs_render_workspace_size(C.byref(s), g.n_gaussians, None)
        ws = native._workspace(n, torch.device("cuda"))
        rc = native._lib.sgs_render(*args, ws.data_ptr(), ws.numel(), stream)
    else:
        rc = native._lib.sgs_render_noalloc(*args, stream)
    assert rc == 0, native._lib.sgs_last_error().decode()
    torch.cuda.synchronize()
    return outs


def _equal(a, b):
    for x, y, name in zip(a, b, ("image", "depth", "alpha", "weight_mass")):
        x, y = x.cpu().numpy(), y.cpu().numpy()
        assert np.array_equal(x, y, equal_nan=False), f"{name}: max diff {np.nanmax(np.abs(x - y))}"


@pytest.mark.parametrize("flags", [dict(), dict(spatial_reuse=True), dict(use_gps=True, hybrid_scheduling=True),
                                   dict(hybrid_scheduling=True, supersampling_factor=2)])
def test_noalloc_equals_render(native, flags):
    import torch
    big = small_scene(4000, seed=61, log_scale_mean=-3.4)
    small = {k: v[:900].copy() for k, v in small_scene(4000, seed=62, log_scale_mean=-3.2).items()}
    eye = np.eye(4, dtype=np.float32)
    for case, (arrays, size) in enumerate(((big, (192, 144)), (small, (96, 64)))):
        t = _dev(arrays)
        cam = small_camera(*size, **flags)
        cam2 = small_camera(*size, angle=0.03, **flags)
        v2p = (np.asarray(cam["w2c"], np.float64) @ np.linalg.inv(np.asarray(cam2["w2c"], np.float64))).astype(
            np.float32)
        print(f"case: noalloc flags={flags} call={case} size={size} n={arrays['means'].shape[0]}", flush=True)
        kbs = None
        if flags.get("spatial_reuse"):
            ss = max(1, int(flags.get("supersampling_factor", 1)))
            P = size[0] * size[1] * ss * ss
            kbs = [(torch.full((4, P, 2), -1, dtype=torch.int32, device="cuda"),
                    torch.zeros((4, P, 4), device="cuda")) for _ in range(2)]
        hist = {"sgs_render": None, "sgs_render_noalloc": None}
        # RESET, STATIC, then REPROJECT to the moved camera
        for step, (c, mode, m) in enumerate(((cam, 0, eye), (cam, 1, eye), (cam2, 2, v2p))):
            res = {}
            for e, entry in enumerate(("sgs_render", "sgs_render_noalloc")):
                res[entry] = _call(native, entry, t, c, 100 + step, mode, hist[entry], m,
                                   kbs[e] if kbs else None)
            _equal(res["sgs_render"], res["sgs_render_noalloc"])
            hist = res
        if kbs:
            assert torch.equal(kbs[0][0], kbs[1][0]) and torch.equal(kbs[0][1], kbs[1][1])


def test_noalloc_rejects_views_beyond_scratch(native):
    import torch
    t = _dev(small_scene(100, seed=63))
    # 4096 x 4200 framebuffer words: more than 128 MiB
    cam = small_camera(2048, 2100, supersampling_factor=2)
    keep = []
    g = native.gaussians_struct(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"], keep)
    s = native.settings_struct(cam["w2c"], cam["cam_position"], 2048, 2100, cam["focal_x"], cam["focal_y"],
                               cam["center_x"], cam["center_y"], 0.2, 1000.0, False, 2, False, False, False,
                               torch.device("cuda"), keep)
    outs = _outputs(4, 4)
    out = native.sgs_history(*[o.data_ptr() for o in outs])
    rc = native._lib.sgs_render_noalloc(C.byref(g), C.byref(s), None, 0, None, None, None, None, 0, C.byref(out), 0,
                                        None, torch.cuda.current_stream().cuda_stream)
    assert rc == -2 and b"scratch" in native._lib.sgs_last_error()
