"""Camera paths from JSON and the frame-time benchmark over them (the reference's
Trainer.py:28-65 and :79-128).

A camera path file is a JSON list of frames, each with ``rotation`` (3x3, camera-to-world),
``position`` (3,), ``id``, and ``fx``, ``fy``, ``width``, ``height`` (the intrinsics the path
was recorded with). ``benchmark_camera_path`` renders every frame once to warm up, then
``num_iterations`` passes over the path, with a device synchronisation around every frame
when per-frame timings are on, and writes ``per_frame_times.txt`` in the reference's
format. ``benchmark_camera_path_batched`` renders the same frames through the batched
renderer, one call per group of frames.
"""
from __future__ import annotations

import json
import warnings
from pathlib import Path
from time import perf_counter

import numpy as np
import torch

from .renderer import PerspectiveCamera, StochasticGSRenderer, View


def load_camera_path(path, camera: PerspectiveCamera | None = None, render_width: int = 1920,
                     render_height: int = 1080, near_plane: float = 0.2, far_plane: float = 10000.0,
                     background_color=(0.0, 0.0, 0.0)) -> list[View]:
    """Views of a camera path file (Trainer.py:28-65). With ``camera`` every frame uses it
    (USE_DATASET_CAMERA, frame camera index 0); without, each frame gets a
    render_width x render_height camera with its focal lengths scaled from the recorded
    size. A missing file gives a warning and no views, like the reference."""
    path_file = Path(path)
    if not path_file.is_file():
        warnings.warn(f"no camera path file at {path_file}; returning no views")
        return []
    with open(path_file, "r") as f:
        camera_path = json.load(f)
    views: list[View] = []
    for frame in camera_path:
        c2w = np.eye(4, dtype=np.float64)
        c2w[:3, :3] = np.asarray(frame["rotation"], dtype=np.float64)
        c2w[:3, 3] = np.asarray(frame["position"], dtype=np.float64)
        if camera is not None:
            cam, index = camera, 0
        else:
            cam = PerspectiveCamera(width=render_width, height=render_height,
                                    focal_x=frame["fx"] * (render_width / frame["width"]),
                                    focal_y=frame["fy"] * (render_height / frame["height"]),
                                    near_plane=near_plane, far_plane=far_plane, background_color=background_color)
            index = frame["id"]
        views.append(View(camera=cam, c2w_numpy=c2w, frame_idx=frame["id"], camera_index=index,
                          global_frame_idx=frame["id"]))
    return views


def save_camera_path(path, views: list[View]) -> None:
    """Writes views in the camera path format (intrinsics of each view's camera)."""
    frames = [{"id": int(v.frame_idx), "rotation": v.c2w_numpy[:3, :3].tolist(),
               "position": v.c2w_numpy[:3, 3].tolist(), "fx": float(v.camera.focal_x), "fy": float(v.camera.focal_y),
               "width": int(v.camera.width), "height": int(v.camera.height)} for v in views]
    Path(path).write_text(json.dumps(frames, indent=1))


def _write_times(output_directory, views, per_frame_times, num_iterations) -> Path:
    frame_times_path = Path(output_directory) / "per_frame_times.txt"
    frame_times_path.parent.mkdir(parents=True, exist_ok=True)
    with open(frame_times_path, "w") as frame_times_file:
        frame_times_file.write(f"average frame time in milliseconds over {num_iterations} passes\n")
        for view, frame_time in zip(views, per_frame_times):
            frame_times_file.write(f"{view.frame_idx}\t{frame_time:.4f}\n")
    return frame_times_path


def save_image(path, rgb: torch.Tensor) -> None:
    """(H,W,3) or (3,H,W) colour in [0,1] as an 8-bit PNG."""
    from PIL import Image
    if rgb.shape[0] == 3 and rgb.dim() == 3 and rgb.shape[-1] != 3:
        rgb = rgb.permute(1, 2, 0)
    img = (rgb.clamp(0.0, 1.0) * 255.0 + 0.5).to(torch.uint8).cpu().numpy()
    Image.fromarray(img).save(path)


def benchmark_camera_path(renderer: StochasticGSRenderer, views: list[View], num_iterations: int = 10,
                          per_frame_timings: bool = True, output_directory=None, save_images: bool = False) -> dict:
    """FPS over the path (Trainer.py:79-128): one warm-up pass, then ``num_iterations``
    timed passes of ``render_image(view, benchmark=True)``."""
    for view in views:
        renderer.render_image(view, benchmark=True)
    torch.cuda.synchronize()
    per_frame_times = np.zeros(len(views), dtype=np.float64)
    start_time = perf_counter()
    for _ in range(num_iterations):
        for frame_index, view in enumerate(views):
            if per_frame_timings:
                torch.cuda.synchronize()
                frame_start_time = perf_counter()
                renderer.render_image(view, benchmark=True)
                torch.cuda.synchronize()
                per_frame_times[frame_index] += (perf_counter() - frame_start_time) * 1000
            else:
                renderer.render_image(view, benchmark=True)
    torch.cuda.synchronize()
    total_time = perf_counter() - start_time
    num_renders = max(len(views) * num_iterations, 1)
    result = {"frames": len(views), "passes": num_iterations, "avg_fps": num_renders / total_time,
              "avg_ms_per_image": total_time * 1000 / num_renders}
    if per_frame_timings:
        per_frame_times /= max(num_iterations, 1)
        result["per_frame_ms"] = per_frame_times.tolist()
        if output_directory is not None:
            result["per_frame_times_path"] = str(_write_times(output_directory, views, per_frame_times,
                                                              num_iterations))
    if save_images and output_directory is not None:
        out = Path(output_directory) / "camera_path"
        out.mkdir(parents=True, exist_ok=True)
        for view in views:
            save_image(out / f"{view.frame_idx:05d}.png", renderer.render_image(view, to_chw=True)["rgb"])
    return result


def batch_views_for_path(renderer: StochasticGSRenderer, views: list[View], first_frame: int = 0):
    """BatchViews rendering each path frame as ``render_image`` does right after a reset
    (RESET, SAMPLES_PER_PIXEL samples, consecutive RNG frames from ``first_frame``)."""
    from .batch import BatchView
    out = []
    frame = first_frame
    spp = max(renderer.SAMPLES_PER_PIXEL, 1)
    for view in views:
        out.append(BatchView(renderer.settings_for(view), spp=spp, frame_idx=frame))
        frame += spp
    return out


def benchmark_camera_path_batched(renderer: StochasticGSRenderer, views: list[View], num_iterations: int = 10,
                                  max_views: int = 64, tuning=None) -> dict:
    """Frames per second of the path through the batched renderer: the frames in groups of
    ``max_views`` views, each group one render call (results equal ``render_image`` of each
    frame after a reset, with the same RNG frames)."""
    from .batch import BatchRenderer
    m = renderer.model
    groups = []
    for g in range(0, len(views), max_views):
        bviews = batch_views_for_path(renderer, views[g:g + max_views], first_frame=g * max(renderer.SAMPLES_PER_PIXEL, 1))
        groups.append(BatchRenderer(m.means, m.scales, m.rotations, m.opacities, m.sh_coefficients, bviews, tuning,
                                    getattr(m, "cov3d_cache", None)))
    for r in groups:
        r.render()
    torch.cuda.synchronize()
    start_time = perf_counter()
    for _ in range(num_iterations):
        for r in groups:
            r.render()
    torch.cuda.synchronize()
    total_time = perf_counter() - start_time
    num_renders = max(len(views) * num_iterations, 1)
    return {"frames": len(views), "passes": num_iterations, "avg_fps": num_renders / total_time,
            "avg_ms_per_image": total_time * 1000 / num_renders, "renderers": groups}
