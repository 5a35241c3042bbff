"""Pins the oracle's EWA projection (gs_utils.h:38-193) against an FP64 closed form."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import oracle_scene, small_camera, small_scene


def ewa_fp64(means, scales, quat, logit, cam, ss=1, aa=False):
    """FP64 EWA splatting (Zwicker et al. 2001) with the renderer's conventions."""
    W = np.asarray(cam["w2c"], np.float64).reshape(4, 4)
    fx, fy = cam["focal_x"] * ss, cam["focal_y"] * ss
    cx, cy = cam["center_x"] * ss, cam["center_y"] * ss
    width, height = cam["width"] * ss, cam["height"] * ss
    t = W[:3, :3] @ means + W[:3, 3]
    z = t[2]
    r, i, j, k = quat / np.linalg.norm(quat)
    R = np.array([
        [1 - 2 * (j * j + k * k), 2 * (i * j - r * k), 2 * (i * k + r * j)],
        [2 * (i * j + r * k), 1 - 2 * (i * i + k * k), 2 * (j * k - r * i)],
        [2 * (i * k - r * j), 2 * (j * k + r * i), 1 - 2 * (i * i + j * j)],
    ])
    S3 = R @ np.diag(np.exp(2.0 * scales)) @ R.T
    x, y = t[0] / z, t[1] / z
    xc = np.clip(x, (-0.15 * width - cx) / fx, (1.15 * width - cx) / fx)
    yc = np.clip(y, (-0.15 * height - cy) / fy, (1.15 * height - cy) / fy)
    J = np.array([[fx / z, 0.0, -fx * xc / z], [0.0, fy / z, -fy * yc / z]])
    M = J @ W[:3, :3]
    S2 = M @ S3 @ M.T
    det_raw = np.linalg.det(S2)
    S2 = S2 + (0.1 if aa else 0.3) * ss * ss * np.eye(2)
    det = np.linalg.det(S2)
    o = 1.0 / (1.0 + np.exp(-logit))
    if aa:
        o *= np.sqrt(max(det_raw / det, 0.0))
    conic = np.linalg.inv(S2)
    pt = np.log(255.0 * o)
    mean2d = np.array([x * fx + cx, y * fy + cy])
    ext = np.sqrt(2.0 * pt * np.array([S2[0, 0], S2[1, 1]]))
    return dict(depth=z, opacity=o, det=det, pt=pt, mean2d=mean2d, conic=np.array([conic[0, 0], conic[0, 1], conic[1, 1]]),
                cov=np.array([S2[0, 0], S2[0, 1], S2[1, 1]]), ext=ext)


@pytest.mark.parametrize("ss,aa", [(1, False), (2, False), (1, True)])
def test_projection_matches_fp64(oracle_mod, ss, aa):
    arrays = small_scene(400, seed=3, log_scale_mean=-3.5)
    sc = oracle_scene(arrays)
    cam = small_camera(96, 80, supersampling_factor=ss, proper_antialiasing=aa)
    checked = 0
    for g in range(sc.n):
        p = oracle_mod.project(sc, g, cam)
        ref = ewa_fp64(arrays["means"][g].astype(np.float64), arrays["scales"][g].astype(np.float64),
                       arrays["rotations"][g].astype(np.float64), float(arrays["opacities"][g]), cam, ss, aa)
        if not p.active:
            continue
        checked += 1
        rel = lambda a, b: np.max(np.abs(np.asarray(a, np.float64) - b) / (np.abs(b) + 1e-6))
        assert rel(p.depth, ref["depth"]) < 1e-6
        assert rel(p.opacity, ref["opacity"]) < 1e-5
        assert np.max(np.abs(np.array(p.mean2d) - ref["mean2d"])) < 2e-4 * ss
        assert rel(np.array(p.cov2d), ref["cov"]) < 2e-4
        assert rel(np.array(p.conic), ref["conic"]) < 2e-4
        assert rel(p.power_threshold, ref["pt"]) < 1e-4
        lo = ref["mean2d"] - ref["ext"]
        assert abs(p.bounds[0] - lo[0]) < 1e-3 * (1 + abs(lo[0])) and abs(p.bounds[2] - lo[1]) < 1e-3 * (1 + abs(lo[1]))
    assert checked > 50


def test_projection_culls(oracle_mod):
    base = small_scene(6, seed=1, log_scale_mean=-3.0)
    cam = small_camera(64, 48)
    eye = np.asarray(cam["cam_position"], np.float64)
    a = {k: v.copy() for k, v in base.items()}
    a["means"][:] = 0.0
    a["opacities"][:] = 3.0
    a["rotations"][:] = [1.0, 0.0, 0.0, 0.0]
    a["means"][0] = eye + 0.1 * (0.0 - eye) / np.linalg.norm(eye)   # closer than the near plane (0.2)
    a["opacities"][1] = -8.0                                            # opacity < 1/255
    a["rotations"][2] = 0.0                                             # zero quaternion
    a["means"][3] = [30.0, 0.0, 0.0]                                    # far outside the frustum
    sc = oracle_scene(a)
    act = [oracle_mod.project(sc, g, cam).active for g in range(6)]
    assert act == [0, 0, 0, 0, 1, 1]
    far_cam = dict(cam, far_plane=1.0)
    assert oracle_mod.project(sc, 4, far_cam).active == 0               # beyond the far plane
