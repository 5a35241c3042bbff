"""A standard 3DGS model loaded from a PLY file (the reference's Model.py:16-79).

``StochasticGSModel(PLY_PATH=...).build()`` registers the buffers ``means``, ``scales``,
``rotations``, ``opacities`` and ``sh_coefficients`` (the names ``Renderer.py:43-46`` reads),
in Morton order of the means, on the current ROCm device. Decoding, ordering and
reordering run in libsgs.so (see ``ply.load_ply``). At load it also precomputes the
view-independent projection terms (3D covariance, sigmoid opacity; the precompute
README.md:139-142 suggests) into ``cov3d_cache``, which the renderer passes to every render;
call ``refresh_cache()`` after changing the parameters.
"""
from __future__ import annotations

import torch

from ._native import precompute_cov3d
from .ply import load_ply


class StochasticGSModel(torch.nn.Module):
    """Stores a standard 3DGS model (Model.py:12-20: configurable PLY_PATH and CAMERA_PATH)."""

    def __init__(self, name: str | None = None, PLY_PATH: str | None = None, CAMERA_PATH: str | None = None,
                 device: torch.device | str | None = None) -> None:
        super().__init__()
        self.name = name
        self.PLY_PATH = PLY_PATH
        self.CAMERA_PATH = CAMERA_PATH
        self.device = device

    def build(self) -> "StochasticGSModel":
        """Loads PLY_PATH (Model.py:22-79)."""
        if self.PLY_PATH is None:
            raise ValueError("PLY_PATH is not set")
        arrays = load_ply(self.PLY_PATH, self.device)
        self.register_buffer("means", arrays["means"])
        self.register_buffer("scales", arrays["scales"])
        self.register_buffer("rotations", arrays["rotations"])
        self.register_buffer("opacities", arrays["opacities"])
        self.register_buffer("sh_coefficients", arrays["sh"])
        self.refresh_cache()
        return self

    def refresh_cache(self) -> None:
        """Recomputes ``cov3d_cache`` from the current scales, rotations and opacities."""
        self.register_buffer("cov3d_cache", precompute_cov3d(self.means, self.scales, self.rotations, self.opacities))

    @classmethod
    def from_arrays(cls, means, scales, rotations, opacities, sh_coefficients, name: str | None = None):
        """A model around arrays already on the device (synthetic scenes, tests)."""
        m = cls(name)
        for k, v in (("means", means), ("scales", scales), ("rotations", rotations), ("opacities", opacities),
                     ("sh_coefficients", sh_coefficients)):
            m.register_buffer(k, v)
        m.refresh_cache()
        return m

    @property
    def n_primitives(self) -> int:
        return int(self.means.shape[0])
