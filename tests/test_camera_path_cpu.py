"""Camera paths in the reference's JSON layout (Trainer.py:40-62): a literal fixture,
tests/golden/camera_path_reference_format.json, written by hand in that layout (a list of
frames with a 3x3 camera-to-world ``rotation``, a ``position``, ``fx``/``fy`` at the recorded
``width``/``height`` and an ``id``), loaded by load_camera_path without a GPU."""
from __future__ import annotations

import json
import warnings
from pathlib import Path

import numpy as np

FIXTURE = Path(__file__).resolve().parent / "golden" / "camera_path_reference_format.json"


def test_fixture_loads_in_both_camera_modes():
    from StochasticGSCudaBackend.camera_path import load_camera_path
    from StochasticGSCudaBackend.renderer import PerspectiveCamera
    raw = json.loads(FIXTURE.read_text())
    # render_width x render_height cameras: focal lengths scaled from the recorded size
    views = load_camera_path(FIXTURE, render_width=1920, render_height=1080)
    assert [v.frame_idx for v in views] == [f["id"] for f in raw] == [10, 11, 12, 13]
    for v, f in zip(views, raw):
        c2w = np.eye(4)
        c2w[:3, :3] = f["rotation"]
        c2w[:3, 3] = f["position"]
        np.testing.assert_array_equal(v.c2w_numpy, c2w)
        assert v.camera.width == 1920 and v.camera.height == 1080
        assert v.camera.focal_x == f["fx"] * 1920 / 1280 and v.camera.focal_y == f["fy"] * 1080 / 720
        assert v.camera_index == f["id"]
        # the rotation is orthonormal and the camera looks at the origin (+z forward)
        R = c2w[:3, :3]
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-8)
        to_origin = -c2w[:3, 3] / np.linalg.norm(c2w[:3, 3])
        np.testing.assert_allclose(R[:, 2], to_origin, atol=1e-8)
        # world-to-view of the view is the inverse pose
        np.testing.assert_allclose(v.w2c.numpy().astype(np.float64) @ c2w, np.eye(4), atol=1e-6)
    # one shared (dataset) camera: every frame uses it, camera index 0
    cam = PerspectiveCamera(640, 360, 550.0, 550.0)
    shared = load_camera_path(FIXTURE, camera=cam)
    assert all(v.camera is cam and v.camera_index == 0 for v in shared)
    assert [v.frame_idx for v in shared] == [10, 11, 12, 13]


def test_missing_file_warns_and_gives_no_views(tmp_path):
    from StochasticGSCudaBackend.camera_path import load_camera_path
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        assert load_camera_path(tmp_path / "absent.json") == []
    assert len(w) == 1
