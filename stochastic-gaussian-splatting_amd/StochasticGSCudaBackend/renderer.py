"""Progressive per-view renderer around ``render`` (the reference's Renderer.py:16-103).

``StochasticGSRenderer.render_image(view)`` keeps the reference's state machine: a new
k-buffer and an empty history whenever a non-pose setting changes (or progressive rendering
is disabled), STATIC accumulation while the camera stands still, REPROJECT (temporal reuse)
or RESET when it moves; ``SAMPLES_PER_PIXEL`` render calls per image, the first in that mode
and the rest STATIC; the premultiplied history composited over the background.

When the settings allow it (stochastic transparency without GPS, hybrid scheduling, spatial
reuse or blue noise), the SAMPLES_PER_PIXEL samples of an image are rendered by one batched
call (sgs_render_batch: the view binned once, its samples rasterised together), which gives
the same history bit for bit as the per-sample loop.

``PerspectiveCamera`` and ``View`` carry the fields of the framework's camera and view
objects that the renderer reads (Renderer.py:51-71,80-89).
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass, field

import numpy as np
import torch

from .render import History, Settings, TemporalMode, blue_noise, create_k_buffer, render


@dataclass
class PerspectiveCamera:
    """Pinhole camera in pixels (OpenCV convention: +x right, +y down, +z forward)."""
    width: int
    height: int
    focal_x: float
    focal_y: float
    center_x: float | None = None   # default: width / 2
    center_y: float | None = None   # default: height / 2
    near_plane: float = 0.2
    far_plane: float = 10000.0
    background_color: tuple = (0.0, 0.0, 0.0)
    distortion: object = None

    def __post_init__(self):
        if self.center_x is None:
            self.center_x = self.width / 2.0
        if self.center_y is None:
            self.center_y = self.height / 2.0


@dataclass
class View:
    """A camera pose: ``c2w`` (4,4) camera-to-world with the camera looking along +z."""
    camera: PerspectiveCamera
    c2w_numpy: np.ndarray
    frame_idx: int = 0
    camera_index: int = 0
    global_frame_idx: int = 0
    _cache: dict = field(default_factory=dict, repr=False, compare=False)

    def __post_init__(self):
        self.c2w_numpy = np.asarray(self.c2w_numpy, dtype=np.float64).reshape(4, 4)

    @property
    def c2w(self) -> torch.Tensor:
        if "c2w" not in self._cache:
            self._cache["c2w"] = torch.as_tensor(self.c2w_numpy, dtype=torch.float32)
        return self._cache["c2w"]

    @property
    def w2c(self) -> torch.Tensor:
        if "w2c" not in self._cache:
            self._cache["w2c"] = torch.as_tensor(np.linalg.inv(self.c2w_numpy), dtype=torch.float32)
        return self._cache["w2c"]

    @property
    def position(self) -> torch.Tensor:
        return torch.as_tensor(self.c2w_numpy[:3, 3], dtype=torch.float32)

    @classmethod
    def from_w2c(cls, camera: PerspectiveCamera, w2c, **kw) -> "View":
        return cls(camera, np.linalg.inv(np.asarray(w2c, dtype=np.float64).reshape(4, 4)), **kw)


class StochasticGSRenderer:
    """Wrapper around the rendering backend (Renderer.py:26-103). The flags are the
    reference's configuration keys (Renderer.py:16-25)."""

    def __init__(self, model, *, PROPER_ANTIALIASING: bool = False, USE_GPS: bool = False,
                 HYBRID_SCHEDULING: bool = False, SUPERSAMPLING_FACTOR: int = 1, SAMPLES_PER_PIXEL: int = 1,
                 DISABLE_PROGRESSIVE_RENDERING: bool = False, SPATIAL_REUSE: bool = False,
                 TEMPORAL_REUSE: bool = False) -> None:
        if not torch.cuda.is_available():
            raise RuntimeError("renderer not implemented in CPU mode")   # Renderer.py:31-32
        self.model = model
        self.PROPER_ANTIALIASING = PROPER_ANTIALIASING
        self.USE_GPS = USE_GPS
        self.HYBRID_SCHEDULING = HYBRID_SCHEDULING
        self.SUPERSAMPLING_FACTOR = SUPERSAMPLING_FACTOR
        self.SAMPLES_PER_PIXEL = SAMPLES_PER_PIXEL
        self.DISABLE_PROGRESSIVE_RENDERING = DISABLE_PROGRESSIVE_RENDERING
        self.SPATIAL_REUSE = SPATIAL_REUSE
        self.TEMPORAL_REUSE = TEMPORAL_REUSE
        self.settings = None
        self.k_buffer = None
        self.history = None
        self.history_c2w = None
        self.history_w2c = None
        self.frame_idx = None   # None: the process-global frame counter; else the next RNG frame

    def _render_once(self, settings: Settings, view_to_previous_view: torch.Tensor,
                     temporal_mode: TemporalMode) -> None:
        """Renders a single sample and merges it into the history (Renderer.py:41-46)."""
        frame = None
        if self.frame_idx is not None:
            frame, self.frame_idx = self.frame_idx, self.frame_idx + 1
        m = self.model
        self.history = render(m.means, m.scales, m.rotations, m.opacities, m.sh_coefficients, settings, self.k_buffer,
                              self.history, view_to_previous_view, temporal_mode, frame_idx=frame,
                              cov3d_cache=getattr(m, "cov3d_cache", None))

    def _render_samples(self, settings: Settings, view_to_previous_view: torch.Tensor, temporal_mode: TemporalMode,
                        spp: int) -> None:
        """The SPP loop of Renderer.py:85-87 (first sample in temporal_mode, the rest STATIC),
        as one batched render of the view's spp samples (every feature, the k-buffer updated
        sample by sample; the same outputs bit for bit)."""
        if spp > 1:
            from .batch import BatchView, render_batch
            frame = None
            if self.frame_idx is not None:
                frame, self.frame_idx = self.frame_idx, self.frame_idx + spp
            reset = temporal_mode == TemporalMode.RESET
            bv = BatchView(settings, spp=spp, frame_idx=frame, temporal_mode=temporal_mode,
                           history=None if reset else self.history,
                           view_to_previous_view=view_to_previous_view if temporal_mode == TemporalMode.REPROJECT
                           else None, k_buffer=self.k_buffer if settings.spatial_reuse else None,
                           blue_noise=blue_noise())
            m = self.model
            self.history = render_batch(m.means, m.scales, m.rotations, m.opacities, m.sh_coefficients, [bv],
                                        cov3d_cache=getattr(m, "cov3d_cache", None))[0]
            return
        for _ in range(spp):
            self._render_once(settings, view_to_previous_view, temporal_mode)
            temporal_mode = TemporalMode.STATIC

    def settings_for(self, view: View) -> Settings:
        cam = view.camera
        return Settings(view.w2c, view.position, cam.width, cam.height, cam.focal_x, cam.focal_y, cam.center_x,
                        cam.center_y, cam.near_plane, cam.far_plane, self.PROPER_ANTIALIASING,
                        self.SUPERSAMPLING_FACTOR, self.USE_GPS, self.HYBRID_SCHEDULING, self.SPATIAL_REUSE)

    def next_temporal_mode(self, view: View, settings: Settings, benchmark: bool = False) -> TemporalMode:
        """The state transition of Renderer.py:73-84 (resets k-buffer and history when needed)."""
        reset = self.DISABLE_PROGRESSIVE_RENDERING and not benchmark
        if reset or self.settings is None or settings.as_tuple()[2:] != self.settings.as_tuple()[2:]:
            self.settings = settings
            self.k_buffer = create_k_buffer(settings)
            self.history = History(*(torch.empty(0) for _ in range(4)))
            return TemporalMode.RESET
        if (self.history_c2w == view.c2w_numpy).all():
            return TemporalMode.STATIC
        return TemporalMode.REPROJECT if self.TEMPORAL_REUSE else TemporalMode.RESET

    @torch.inference_mode()
    def render_image(self, view: View, to_chw: bool = False, benchmark: bool = False) -> dict[str, torch.Tensor]:
        """Renders an image for a given view (Renderer.py:48-103)."""
        if not isinstance(view.camera, PerspectiveCamera):
            raise TypeError("renderer only supports perspective cameras")
        if view.camera.distortion is not None:
            warnings.warn("found distortion parameters that will be ignored by the rasterizer")
        settings = self.settings_for(view)
        temporal_mode = self.next_temporal_mode(view, settings, benchmark)
        view_to_previous_view = self.history_w2c @ view.c2w if temporal_mode == TemporalMode.REPROJECT else view.c2w
        self._render_samples(settings, view_to_previous_view, temporal_mode, max(self.SAMPLES_PER_PIXEL, 1))
        self.history_c2w = view.c2w_numpy
        self.history_w2c = view.w2c

        # the history tensors are the next call's input (colour and depth carry a factor of
        # coverage): composite into new tensors, never in place
        image, depth, alpha, _ = self.history
        alpha = alpha.clamp(0.0, 1.0)
        bg = torch.as_tensor(view.camera.background_color, dtype=image.dtype, device=image.device)
        image = (image + (1.0 - alpha) * bg).clamp(0.0, 1.0)
        if benchmark:
            return {"rgb": image}
        depth = depth / alpha.clamp_min(1.0e-8)
        if to_chw:
            image = image.permute(2, 0, 1)
            depth = depth.permute(2, 0, 1)
            alpha = alpha.permute(2, 0, 1)
        return {"rgb": image, "depth": depth, "alpha": alpha}
