"""Python backend API of the renderer: the drop-in for the reference's
StochasticGSCudaBackend/StochasticGSCudaBackend/render.py:1-95.

Same names, fields, argument order and return types (``render``, ``Settings``,
``TemporalMode``, ``History``, ``blue_noise``, ``create_k_buffer``); the native module is
``_native`` (libsgs.so through ctypes) instead of the pybind11 ``_C``. Two keyword-only
additions to ``render``: ``frame_idx`` (the RNG frame, by default the process-global
counter the reference keeps in render.cu:56-57) and ``tuning`` (scheduling knobs).
"""
from __future__ import annotations

from enum import IntEnum
from functools import cache
from pathlib import Path
from typing import NamedTuple

import torch

from . import _native as _C
from ._native import Tuning


class Settings(NamedTuple):
    """Camera and feature flags of one render call (render.py:8-42).

    Conventions: ``w2c`` is the (4, 4) row-major matrix taking world points into the camera
    frame (+x right, +y down, +z forward, OpenCV style); ``cam_position`` is the camera centre
    in world coordinates (the viewing direction for SH shading starts there); sizes and
    intrinsics are in pixels of the output image, the principal point measured from the top
    left corner; geometry outside [near_plane, far_plane] of view depth is culled.
    ``proper_antialiasing`` uses the smaller screen-space dilation with opacity compensation
    instead of the plain dilation; ``supersampling_factor`` s renders s x s framebuffer cells
    per output pixel and box-filters them; ``use_gps`` splats Poisson-distributed points
    instead of per-fragment stochastic transparency; ``hybrid_scheduling`` moves Gaussians with
    large footprints to a tile-binned pass; ``spatial_reuse`` adds the four edge neighbours'
    samples, weighted by their winners' footprints, to every framebuffer cell."""
    w2c: torch.Tensor
    cam_position: torch.Tensor
    width: int
    height: int
    focal_x: float
    focal_y: float
    center_x: float
    center_y: float
    near_plane: float
    far_plane: float
    proper_antialiasing: bool
    supersampling_factor: int
    use_gps: bool
    hybrid_scheduling: bool
    spatial_reuse: bool

    def as_tuple(self) -> tuple:
        return (
            self.w2c,
            self.cam_position,
            self.width,
            self.height,
            self.focal_x,
            self.focal_y,
            self.center_x,
            self.center_y,
            self.near_plane,
            self.far_plane,
            self.proper_antialiasing,
            self.supersampling_factor,
            self.use_gps,
            self.hybrid_scheduling,
            self.spatial_reuse,
        )


class TemporalMode(IntEnum):
    """How a sample is merged with the previous result (resolve.h:11-15): RESET ignores it;
    STATIC (camera unchanged) keeps a running average weighted by sample mass; REPROJECT
    (camera moved) fetches it at each pixel's reprojected position and down-weights it where
    it disagrees with the new sample's neighbourhood."""
    RESET = 0
    STATIC = 1
    REPROJECT = 2


class History(NamedTuple):
    """The resolved image and what accumulating more samples into it needs: colour (H, W, 3)
    and depth (H, W, 1), both multiplied by coverage; coverage (H, W, 1); and the summed
    sample weight per pixel (H, W, 1)."""
    image: torch.Tensor
    depth: torch.Tensor
    alpha: torch.Tensor
    weight_mass: torch.Tensor


@cache
def blue_noise() -> torch.Tensor:
    """The (64, 128, 128) uint8 spatio-temporal blue-noise volume stored next to the native
    library as stbn_scalar_128x128x64.bin, or an empty tensor (white noise) if absent.
    The volume is not shipped (its licence is non-commercial)."""
    path = Path(_C.LIB_PATH).parent / "stbn_scalar_128x128x64.bin"
    if not path.is_file():
        return torch.empty(0, dtype=torch.uint8, device="cuda")
    return torch.frombuffer(bytearray(path.read_bytes()), dtype=torch.uint8).view(64, 128, 128).cuda()


def create_k_buffer(settings: Settings) -> tuple[torch.Tensor, torch.Tensor]:
    """Fresh spatial-reuse state for ``settings``: per framebuffer cell, k_buffer_size slots of
    (Gaussian index, depth bits) as int32 pairs, all -1 (no entry), and of four float32
    occlusion ratios, one per edge neighbour; both have zero slots without spatial reuse."""
    n_entries = (_C.k_buffer_size if settings.spatial_reuse else 0,
                 settings.width * settings.height * settings.supersampling_factor ** 2)
    return (
        torch.full((*n_entries, 2), -1, dtype=torch.int32, device="cuda"),
        torch.empty((*n_entries, 4), dtype=torch.float32, device="cuda"),
    )


def render(
    means: torch.Tensor,
    scales: torch.Tensor,
    rotations: torch.Tensor,
    opacities: torch.Tensor,
    sh_coefficients: torch.Tensor,
    settings: Settings,
    k_buffer: tuple[torch.Tensor, torch.Tensor],
    history: History,
    view_to_previous_view: torch.Tensor,
    temporal_mode: TemporalMode,
    *,
    frame_idx: int | None = None,
    tuning: Tuning | None = None,
    cov3d_cache: torch.Tensor | None = None,
) -> History:
    """Renders one stochastic sample and merges it into ``history`` (render.py:72-95).

    Stochastic transparency without hybrid scheduling accumulates by default with the tile
    schedule: each (bin, Gaussian) instance is listed per framebuffer bin and each bin's list
    is sorted into depth buckets so 16x16 tiles can stop early, where the reference scatters
    sorting-free (README.md:6). The framebuffer is the same bit for bit; ``tuning`` selects
    the scatter (``Tuning(schedule=2)``) or caps the tile schedule's per-call workspace
    (``max_workspace_mib``, default a quarter of the device's memory), above which the view
    scatters. ``cov3d_cache``: precompute_cov3d's output for these Gaussians, or None."""
    return History(*_C.render(
        means,
        scales,
        rotations,
        opacities,
        sh_coefficients,
        *settings.as_tuple(),
        blue_noise(),
        *k_buffer,
        *history,
        view_to_previous_view,
        int(temporal_mode),
        frame_idx=frame_idx,
        tuning=tuning,
        cov3d_cache=cov3d_cache,
    ))
