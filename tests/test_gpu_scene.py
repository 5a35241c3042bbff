"""GPU parity of scene loading (sgs_ply_decode, sgs_morton_order, sgs_gather_gaussians through
the C ABI) against the oracle's restatement of Model.py:22-79: arrays bit for bit, Morton codes
and order exactly. Also the model / renderer / camera-path layer above it: the renderer's
state machine against explicit render() calls, and the camera-path benchmark."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import small_scene
from oracle import scene_oracle
from plyfiles import gaussian_rows, random_arrays, write_ply
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


def _host(arrays):
    return {k: v.cpu().numpy() for k, v in arrays.items()}


def _assert_same(got, ref):
    assert set(got) == set(ref)
    for k in ref:
        assert got[k].shape == ref[k].shape, k
        np.testing.assert_array_equal(got[k].view(np.uint32), ref[k].view(np.uint32), err_msg=k)


@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian", "ascii"])
@pytest.mark.parametrize("bases", [1, 4, 16])
@pytest.mark.parametrize("morton", [True, False])
def test_load_ply_matches_oracle(gpu, tmp_path, fmt, bases, morton):
    from StochasticGSCudaBackend.ply import load_ply
    n = 300 if fmt == "ascii" else 2000
    arrays = random_arrays(n, bases, seed=bases + 17)
    path = tmp_path / "s.ply"
    write_ply(path, gaussian_rows(arrays), fmt)
    print_case(test="load_ply", fmt=fmt, bases=bases, morton=morton, n=n)
    got = _host(load_ply(path, morton=morton))
    _assert_same(got, scene_oracle.load_scene(path, morton=morton))


def test_load_ply_chunks_and_scalar_types(gpu, tmp_path):
    from StochasticGSCudaBackend.ply import load_ply
    arrays = random_arrays(1001, 9, seed=4)

    def extra(i, r):
        r = [(nm, "double", float(v) * 1.0000001) if nm.startswith(("scale_", "rot_")) else (nm, t, v)
             for nm, t, v in r]
        r = [(nm, "short", int(i % 700) - 350) if nm == "opacity" else (nm, t, v) for nm, t, v in r]
        return r[:6] + [("red", "uchar", i % 256)] + r[6:] + [("flag", "char", -3), ("id", "uint", 4000000000 + i)]

    path = tmp_path / "t.ply"
    write_ply(path, gaussian_rows(arrays, extra), "binary_big_endian", trailing_element=True)
    ref = scene_oracle.load_scene(path)
    for chunk in (1 << 30, 5 * 251, 251):   # whole file, 5 records, 1 record per piece (stride 251)
        print_case(test="load_ply_chunks", chunk=chunk)
        _assert_same(_host(load_ply(path, chunk_bytes=chunk)), ref)


def test_load_ply_empty_and_errors(gpu, tmp_path):
    from StochasticGSCudaBackend.ply import load_ply
    empty = tmp_path / "e.ply"
    hdr = ("ply\nformat binary_little_endian 1.0\nelement vertex 0\n" +
           "".join(f"property float {n}\n" for n in ("x", "y", "z", "scale_0", "scale_1", "scale_2", "rot_0",
                                                      "rot_1", "rot_2", "rot_3", "opacity", "f_dc_0", "f_dc_1",
                                                      "f_dc_2")) + "end_header\n")
    empty.write_bytes(hdr.encode())
    out = load_ply(empty)
    assert out["means"].shape == (0, 3) and out["sh"].shape == (0, 1, 3)
    trunc = tmp_path / "t.ply"
    write_ply(trunc, gaussian_rows(random_arrays(10, 4)))
    data = trunc.read_bytes()
    trunc.write_bytes(data[:-7])
    with pytest.raises(ValueError, match="announces"):
        load_ply(trunc)


def test_morton_order_matches_oracle(gpu):
    import torch
    rng = np.random.default_rng(9)
    m = rng.normal(size=(300_000, 3)).astype(np.float32) * np.float32(3.0)
    m[:1000] = m[1000:2000]                    # duplicates: ties keep index order
    m[5, 0], m[6, 1], m[7, 2] = np.nan, np.inf, -np.inf
    print_case(test="morton", n=len(m))
    order, codes = gpu.morton_order(torch.from_numpy(m).cuda(), return_codes=True)
    ref_codes = scene_oracle.morton_codes(m)
    ref_order = scene_oracle.morton_order(m)
    np.testing.assert_array_equal(order.cpu().numpy(), ref_order)
    np.testing.assert_array_equal(codes.cpu().numpy().view(np.uint64), ref_codes[ref_order])


def test_morton_degenerate(gpu):
    import torch
    for m in (np.zeros((5, 3), np.float32), np.full((3, 3), np.nan, np.float32), np.ones((1, 3), np.float32)):
        print_case(test="morton_degenerate", n=len(m))
        order, codes = gpu.morton_order(torch.from_numpy(m).cuda(), return_codes=True)
        np.testing.assert_array_equal(order.cpu().numpy(), np.arange(len(m)))
        assert not codes.cpu().numpy().any()


def test_gather_matches_indexing(gpu):
    import torch
    arrays = random_arrays(5000, 16, seed=2)
    perm = np.random.default_rng(3).permutation(5000)[:4000]
    print_case(test="gather", n=5000)
    got = gpu.gather_gaussians(to_device(arrays), torch.from_numpy(perm.astype(np.int32)).cuda())
    _assert_same(_host(got), {k: v[perm] for k, v in arrays.items()})


# ---------------------------------------------------------------------------------------
# model, renderer state machine, camera path
# ---------------------------------------------------------------------------------------
def _camera(width=64, height=48):
    from StochasticGSCudaBackend.renderer import PerspectiveCamera
    f = (height / 2.0) / np.tan(np.radians(30.0))
    return PerspectiveCamera(width, height, f, f)


def _view(angle, cam, frame=0):
    from StochasticGSCudaBackend.renderer import View
    from StochasticGSCudaBackend.synthetic import look_at, orbit_eye
    return View.from_w2c(cam, look_at(orbit_eye(angle, 3.0)), frame_idx=frame)


def test_model_build_renders_like_oracle_arrays(gpu, tmp_path):
    import torch
    from StochasticGSCudaBackend.model import StochasticGSModel
    from StochasticGSCudaBackend.renderer import StochasticGSRenderer
    arrays = small_scene(3000, seed=5, log_scale_mean=-3.5)
    path = tmp_path / "scene.ply"
    write_ply(path, gaussian_rows(arrays))
    model = StochasticGSModel(PLY_PATH=str(path)).build()
    ref = scene_oracle.load_scene(path)
    _assert_same(_host({"means": model.means, "scales": model.scales, "rotations": model.rotations,
                        "opacities": model.opacities, "sh": model.sh_coefficients}), ref)
    r = StochasticGSRenderer(model, SAMPLES_PER_PIXEL=3)
    r.frame_idx = 40
    view = _view(0.3, _camera())
    print_case(test="model_render")
    out = r.render_image(view)
    t = to_device(ref)
    cam = {"w2c": view.w2c.numpy(), "cam_position": view.position.numpy(), "width": 64, "height": 48,
           "focal_x": view.camera.focal_x, "focal_y": view.camera.focal_y, "center_x": 32.0, "center_y": 24.0,
           "near_plane": 0.2, "far_plane": 10000.0}
    hist = plain_spp(gpu, t, cam, 40, 3)
    assert_close(r.history, hist)
    alpha = hist[2].clamp(0, 1)
    np.testing.assert_array_equal(out["rgb"].cpu().numpy(), hist[0].clamp(0, 1).cpu().numpy())
    np.testing.assert_array_equal(out["alpha"].cpu().numpy(), alpha.cpu().numpy())
    torch.cuda.synchronize()


def test_renderer_state_machine(gpu):
    import torch
    from StochasticGSCudaBackend import TemporalMode
    from StochasticGSCudaBackend.model import StochasticGSModel
    from StochasticGSCudaBackend.renderer import StochasticGSRenderer
    arrays = small_scene(2000, seed=6, log_scale_mean=-3.5)
    t = to_device(arrays)
    model = StochasticGSModel.from_arrays(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"])
    cam = _camera()
    a, b = _view(0.0, cam), _view(0.05, cam)
    for temporal_reuse in (False, True):
        r = StochasticGSRenderer(model, SAMPLES_PER_PIXEL=2, TEMPORAL_REUSE=temporal_reuse)
        r.frame_idx = 0
        modes = []
        for v in (a, a, b):
            modes.append(r.next_temporal_mode(v, r.settings_for(v)) if r.settings is not None else TemporalMode.RESET)
            r.render_image(v)
        print_case(test="state_machine", temporal_reuse=temporal_reuse)
        assert modes[1] == TemporalMode.STATIC
        assert modes[2] == (TemporalMode.REPROJECT if temporal_reuse else TemporalMode.RESET)
        # the same sequence by explicit render() calls
        ca = {"w2c": a.w2c.numpy(), "cam_position": a.position.numpy(), "width": 64, "height": 48,
              "focal_x": cam.focal_x, "focal_y": cam.focal_y, "center_x": 32.0, "center_y": 24.0,
              "near_plane": 0.2, "far_plane": 10000.0}
        cb = dict(ca, w2c=b.w2c.numpy(), cam_position=b.position.numpy())
        h = plain_spp(gpu, t, ca, 0, 2)
        h = plain_spp(gpu, t, ca, 2, 2, mode=1, history=h)
        if temporal_reuse:
            v2p = (a.w2c @ b.c2w).numpy()
            h = plain_spp(gpu, t, cb, 4, 2, mode=2, history=h, v2p=v2p)
        else:
            h = plain_spp(gpu, t, cb, 4, 2)
        assert_close(r.history, h)


def test_renderer_composites_over_background(gpu):
    """Renderer.py:91-103 on a coloured background: rgb = clamp(image + (1 - clamp(alpha)) bg),
    depth = depth / max(alpha, 1e-8), alpha clamped, the history left as rendered; the
    benchmark form returns the composited rgb only."""
    import torch
    from StochasticGSCudaBackend.model import StochasticGSModel
    from StochasticGSCudaBackend.renderer import PerspectiveCamera, StochasticGSRenderer, View
    from StochasticGSCudaBackend.synthetic import look_at, orbit_eye
    arrays = small_scene(1500, seed=9, log_scale_mean=-3.4)
    t = to_device(arrays)
    model = StochasticGSModel.from_arrays(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"])
    bg = (0.2, 0.5, 0.9)
    f = 24.0 / np.tan(np.radians(30.0))
    cam = PerspectiveCamera(64, 48, f, f, background_color=bg)
    view = View.from_w2c(cam, look_at(orbit_eye(0.3, 3.0)), frame_idx=0)
    r = StochasticGSRenderer(model, SAMPLES_PER_PIXEL=3)
    r.frame_idx = 40
    print_case(test="composite")
    out = r.render_image(view)
    torch.cuda.synchronize()
    ca = {"w2c": view.w2c.numpy(), "cam_position": view.position.numpy(), "width": 64, "height": 48,
          "focal_x": f, "focal_y": f, "center_x": 32.0, "center_y": 24.0, "near_plane": 0.2, "far_plane": 10000.0}
    h = [x.cpu().numpy().astype(np.float64) for x in plain_spp(gpu, t, ca, 40, 3)]
    assert_close(r.history, plain_spp(gpu, t, ca, 40, 3))       # the history is not modified in place
    a = np.clip(h[2], 0.0, 1.0)
    rgb = np.clip(h[0] + (1.0 - a) * np.asarray(bg), 0.0, 1.0)
    depth = h[1] / np.maximum(a, 1e-8)
    assert 0.05 < a.mean() < 0.95 and (a == 0).any()           # covered and uncovered pixels
    np.testing.assert_allclose(out["rgb"].cpu().numpy(), rgb, atol=1e-6)
    np.testing.assert_allclose(out["alpha"].cpu().numpy(), a, atol=0)
    np.testing.assert_allclose(out["depth"].cpu().numpy(), depth, rtol=1e-6, atol=1e-6)
    bgpix = h[2][..., 0] == 0
    np.testing.assert_allclose(out["rgb"].cpu().numpy()[bgpix], np.broadcast_to(bg, (int(bgpix.sum()), 3)), atol=1e-7)
    assert np.all(out["depth"].cpu().numpy()[bgpix] == 0.0)
    chw = r.render_image(view, to_chw=True)
    assert chw["rgb"].shape == (3, 48, 64) and chw["depth"].shape == (1, 48, 64)
    bench = r.render_image(view, benchmark=True)
    assert set(bench) == {"rgb"}


def test_settings_change_resets(gpu):
    from StochasticGSCudaBackend import TemporalMode
    from StochasticGSCudaBackend.model import StochasticGSModel
    from StochasticGSCudaBackend.renderer import StochasticGSRenderer
    t = to_device(small_scene(500, seed=7))
    model = StochasticGSModel.from_arrays(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"])
    r = StochasticGSRenderer(model)
    v = _view(0.0, _camera())
    r.render_image(v)
    assert r.next_temporal_mode(v, r.settings_for(v)) == TemporalMode.STATIC
    r.render_image(v)
    v2 = _view(0.0, _camera(80, 48))          # another resolution: new k-buffer and history
    assert r.next_temporal_mode(v2, r.settings_for(v2)) == TemporalMode.RESET
    r.DISABLE_PROGRESSIVE_RENDERING = True
    assert r.next_temporal_mode(v2, r.settings_for(v2)) == TemporalMode.RESET
    assert r.next_temporal_mode(v2, r.settings_for(v2), benchmark=True) == TemporalMode.STATIC


def test_camera_path_benchmark(gpu, tmp_path):
    from StochasticGSCudaBackend.camera_path import (benchmark_camera_path, benchmark_camera_path_batched,
                                                     load_camera_path, save_camera_path)
    from StochasticGSCudaBackend.model import StochasticGSModel
    from StochasticGSCudaBackend.renderer import StochasticGSRenderer
    t = to_device(small_scene(3000, seed=8, log_scale_mean=-3.5))
    model = StochasticGSModel.from_arrays(t["means"], t["scales"], t["rotations"], t["opacities"], t["sh"])
    cam = _camera(96, 64)
    views = [_view(0.1 * k, cam, frame=k) for k in range(6)]
    path = tmp_path / "path.json"
    save_camera_path(path, views)
    loaded = load_camera_path(path, render_width=96, render_height=64)
    assert [v.frame_idx for v in loaded] == list(range(6))
    for v, w in zip(loaded, views):
        np.testing.assert_allclose(v.c2w_numpy, w.c2w_numpy, atol=1e-12)
        assert abs(v.camera.focal_x - cam.focal_x) < 1e-9
    r = StochasticGSRenderer(model, SAMPLES_PER_PIXEL=2)
    print_case(test="camera_path")
    res = benchmark_camera_path(r, loaded, num_iterations=2, output_directory=tmp_path, save_images=True)
    lines = (tmp_path / "per_frame_times.txt").read_text().splitlines()
    assert lines[0] == "average frame time in milliseconds over 2 passes" and len(lines) == 7
    assert lines[3].split("\t")[0] == "2" and float(lines[3].split("\t")[1]) > 0
    assert (tmp_path / "camera_path" / "00005.png").is_file()
    assert res["avg_fps"] > 0 and len(res["per_frame_ms"]) == 6
    # the batched path: each frame equals render_image right after a reset with the same frames
    b = benchmark_camera_path_batched(r, loaded, num_iterations=2, max_views=4)
    seq = StochasticGSRenderer(model, SAMPLES_PER_PIXEL=2)
    seq.frame_idx = 0
    outs = [o for g in b["renderers"] for o in g.outputs]
    for v, o in zip(loaded, outs):
        seq.render_image(v)
        assert_close(o, seq.history)
