"""Seeded synthetic scenes and cameras for tests and benchmarks.

Gaussians are stored the way the renderer reads them (the reference's Model.py:22-79
layout): means (N,3), log scales (N,3), quaternions (N,4) as (r,i,j,k), opacity logits
(N,) and SH coefficients (N,16,3) for degree 3. Cameras follow the OpenCV convention
(+x right, +y down, +z forward) with a vertical field of view.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np


@dataclass(frozen=True)
class SceneSpec:
    n: int
    seed: int = 0
    means: str = "cube"            # "cube": U[-1,1]^3; "mixture": equal-weight Gaussian mixture
    n_components: int = 3
    component_sigma: float = 0.3
    log_scale_mean: float = -4.5
    log_scale_std: float = 0.5
    heavy_fraction: float = 0.0    # share of Gaussians with log scales from N(heavy_mean, heavy_std)
    heavy_mean: float = -1.5
    heavy_std: float = 0.3
    opacity_std: float = 1.5       # logits ~ N(0, opacity_std)
    sh_degree: int = 3


def make_scene(spec: SceneSpec) -> dict[str, np.ndarray]:
    """Gaussian arrays (float32, contiguous) for ``spec``."""
    rng = np.random.default_rng(spec.seed)
    n = spec.n
    f32 = np.float32
    if spec.means == "cube":
        means = rng.uniform(-1.0, 1.0, size=(n, 3)).astype(f32)
    elif spec.means == "mixture":
        centres = rng.uniform(-0.7, 0.7, size=(spec.n_components, 3)).astype(f32)
        comp = rng.integers(0, spec.n_components, size=n)
        means = centres[comp] + spec.component_sigma * rng.standard_normal((n, 3), dtype=f32)
    else:
        raise ValueError(f"unknown means distribution {spec.means!r}")
    scales = (spec.log_scale_mean + spec.log_scale_std * rng.standard_normal((n, 3), dtype=f32)).astype(f32)
    if spec.heavy_fraction > 0.0:
        heavy = rng.random(n) < spec.heavy_fraction
        k = int(heavy.sum())
        scales[heavy] = (spec.heavy_mean + spec.heavy_std * rng.standard_normal((k, 3), dtype=f32)).astype(f32)
    q = rng.standard_normal((n, 4), dtype=f32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    opacities = (spec.opacity_std * rng.standard_normal(n, dtype=f32)).astype(f32)
    bases = (spec.sh_degree + 1) ** 2
    sh = (0.05 * rng.standard_normal((n, bases, 3), dtype=f32)).astype(f32)
    sh[:, 0, :] = (0.5 + 0.2 * rng.standard_normal((n, 3), dtype=f32)).astype(f32)
    return {
        "means": np.ascontiguousarray(means, f32),
        "scales": np.ascontiguousarray(scales, f32),
        "rotations": np.ascontiguousarray(q, f32),
        "opacities": np.ascontiguousarray(opacities, f32),
        "sh": np.ascontiguousarray(sh, f32),
    }


def morton_order(means: np.ndarray, bits: int = 10) -> np.ndarray:
    """Permutation sorting Gaussians by the Morton code of their means quantised to
    2^bits cells per axis of the bounding box, the order the reference's model loader
    stores scenes in (Model.py:32-34), which keeps screen-space neighbours together."""
    lo = means.min(axis=0)
    span = np.maximum(means.max(axis=0) - lo, 1e-12)
    q = np.clip(((means - lo) / span * (1 << bits)).astype(np.int64), 0, (1 << bits) - 1)
    code = np.zeros(len(means), np.int64)
    for b in range(bits):
        for axis in range(3):
            code |= ((q[:, axis] >> b) & 1) << (3 * b + axis)
    return np.argsort(code, kind="stable")


def reorder(arrays: dict[str, np.ndarray], order: np.ndarray) -> dict[str, np.ndarray]:
    return {k: np.ascontiguousarray(v[order]) for k, v in arrays.items()}


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, -1.0, 0.0)) -> np.ndarray:
    """4x4 world-to-view matrix, +z towards ``target``, image +y along -``up``."""
    eye = np.asarray(eye, np.float64)
    z = np.asarray(target, np.float64) - eye
    z /= np.linalg.norm(z)
    y_dir = -np.asarray(up, np.float64)
    x = np.cross(y_dir, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z])
    w2c = np.eye(4)
    w2c[:3, :3] = R
    w2c[:3, 3] = -R @ eye
    return w2c


def camera(eye, width: int, height: int, fov_y_deg: float = 60.0, target=(0.0, 0.0, 0.0), near: float = 0.2,
           far: float = 10000.0, **flags) -> dict:
    """Pinhole camera dict with the fields of Settings (image pixels) plus any ``flags``."""
    f = (height / 2.0) / math.tan(math.radians(fov_y_deg) / 2.0)
    w2c = look_at(eye, target)
    cam = {
        "w2c": w2c.astype(np.float32),
        "cam_position": np.asarray(eye, np.float32),
        "width": int(width), "height": int(height),
        "focal_x": float(f), "focal_y": float(f),
        "center_x": width / 2.0, "center_y": height / 2.0,
        "near_plane": float(near), "far_plane": float(far),
        "proper_antialiasing": False, "supersampling_factor": 1,
        "use_gps": False, "hybrid_scheduling": False, "spatial_reuse": False,
    }
    cam.update(flags)
    return cam


def orbit_eye(angle_rad: float, radius: float = 3.0, height: float = 0.0) -> np.ndarray:
    return np.array([radius * math.sin(angle_rad), height, -radius * math.cos(angle_rad)])


# ------------------------------------------------------------------------------------------
# workloads of the benchmark configurations
# ------------------------------------------------------------------------------------------
def scene_c1() -> SceneSpec:
    """100k Gaussians, means U[-1,1]^3, log scales N(-4.5, 0.5)."""
    return SceneSpec(n=100_000, means="cube", log_scale_mean=-4.5, log_scale_std=0.5)


def scene_c2() -> SceneSpec:
    """1M Gaussians, 3-component mixture (sigma 0.3), log scales N(-5, 0.7)."""
    return SceneSpec(n=1_000_000, means="mixture", n_components=3, log_scale_mean=-5.0, log_scale_std=0.7)


def scene_c3() -> SceneSpec:
    """3M Gaussians, heavy tail: log scales N(-5, 1.0) with 1% from N(-1.5, 0.3)."""
    return SceneSpec(n=3_000_000, means="mixture", n_components=3, log_scale_mean=-5.0, log_scale_std=1.0,
                     heavy_fraction=0.01, heavy_mean=-1.5, heavy_std=0.3)


def scene_c4() -> SceneSpec:
    """6M Gaussians, 8-component mixture, log scales N(-5.5, 0.8), opacity logits N(0, 2)."""
    return SceneSpec(n=6_000_000, means="mixture", n_components=8, log_scale_mean=-5.5, log_scale_std=0.8,
                     opacity_std=2.0)


def camera_c1(**flags) -> dict:
    return camera(orbit_eye(0.0, 3.0), 800, 600, **flags)


def camera_c2(**flags) -> dict:
    return camera(orbit_eye(0.0, 3.0), 1920, 1080, **flags)


def camera_c3(**flags) -> dict:
    flags.setdefault("supersampling_factor", 2)
    return camera(orbit_eye(0.0, 8.0), 1920, 1080, **flags)


C4_RESOLUTIONS = [(256, 256), (640, 480), (1280, 720), (1920, 1080), (2048, 1152)]
C4_SPP = [1, 4, 16]


def views_c4(n_views: int = 64, seed: int = 1) -> list[dict]:
    """64 cameras on a ring of radius 3 with N(0, 0.1) jitter, uneven resolution and spp."""
    rng = np.random.default_rng(seed)
    views = []
    for v in range(n_views):
        eye = orbit_eye(2.0 * math.pi * v / n_views, 3.0) + rng.normal(0.0, 0.1, 3)
        w, h = C4_RESOLUTIONS[int(rng.integers(len(C4_RESOLUTIONS)))]
        cam = camera(eye, w, h)
        cam["spp"] = int(C4_SPP[int(rng.integers(len(C4_SPP)))])
        views.append(cam)
    return views


def scaled(spec: SceneSpec, n: int) -> SceneSpec:
    """The same distribution with a different Gaussian count (for small parity cases)."""
    return replace(spec, n=n)
