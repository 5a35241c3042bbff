"""Batched multi-view rendering: many cameras of uneven resolution and spp in one
persistent launch (sgs_render_batch).

View v's result equals ``spp`` successive ``render()`` calls with frames
``frame_idx .. frame_idx + spp - 1``, the first with ``temporal_mode`` and the rest STATIC
(the SPP loop of the reference's Renderer.py:85-87), with identical framebuffer winners.
``BatchRenderer`` keeps the workspace, outputs and the encoded view table resident so a
render can be repeated or captured into a graph and replayed. ``StreamedBatchRenderer``
renders contiguous groups of the views as concurrent batches on their own streams, with the
same outputs.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import NamedTuple, Sequence

import numpy as np
import torch

from . import _native
from .render import History, Settings, TemporalMode


SCHEDULE_TILES = 1    # LDS framebuffer tiles, one workgroup per (view, 16x16 tile)
SCHEDULE_SCATTER = 2  # persistent scatter with atomicMin into device memory


class BatchTuning(NamedTuple):
    """sgs_batch_tuning; 0 keeps the library default. No field changes a result bit."""
    sequential_threshold: int = 0
    big_threshold: int = 0
    entry_fragments: int = 0
    queue_capacity: int = 0
    view_major: int = 0
    blocks_per_cu: int = 0
    schedule: int = 0
    max_tile_instances: int = 0
    tile_samples: int = 0

    def c(self) -> _native.sgs_batch_tuning:
        return _native.sgs_batch_tuning(*[int(v) for v in self])


@dataclass
class BatchView:
    settings: Settings
    spp: int = 1
    frame_idx: int | None = None  # None: the next spp frames of the process-global counter
    temporal_mode: TemporalMode = TemporalMode.RESET
    history: History | None = None
    view_to_previous_view: torch.Tensor | np.ndarray | None = None
    # tile schedule bins of this view (16, 32 or 64 pixels; 0 = 16 for framebuffers of at most
    # 2^19 pixels, else 32); no result bit depends on it
    bin_size: int = 0
    # the single-view features (settings.use_gps / hybrid_scheduling / spatial_reuse come from
    # settings): the view's k-buffer (create_k_buffer) when spatial_reuse, updated in place like
    # render()'s; a (64,128,128) uint8 blue-noise volume or None; the GPS point cap (0: the
    # reference's)
    k_buffer: tuple | None = None
    blue_noise: torch.Tensor | None = None
    max_gps_points: int = 0


def _host(t, shape) -> np.ndarray:
    if isinstance(t, torch.Tensor):
        t = t.detach().to("cpu")
    return np.ascontiguousarray(np.asarray(t, np.float32).reshape(shape))


class BatchRenderer:
    """Renders a fixed list of views repeatedly; outputs are ``self.outputs`` (one History
    per view, overwritten by every render)."""

    def __init__(self, means, scales, rotations, opacities, sh_coefficients, views: Sequence[BatchView],
                 tuning: BatchTuning | None = None, cov3d_cache: torch.Tensor | None = None):
        self.device = means.device
        self._keep: list = []
        self.gaussians = _native.gaussians_struct(means, scales, rotations, opacities, sh_coefficients, self._keep,
                                                  cov3d_cache)
        self.n_views = len(views)
        self.views = list(views)
        self.tuning = tuning.c() if tuning is not None else None
        self._tp = C.byref(self.tuning) if self.tuning is not None else None
        self.c_views = (_native.sgs_view * max(1, self.n_views))()
        self.outputs: list[History] = []
        opts = dict(dtype=torch.float32, device=self.device)
        for i, v in enumerate(self.views):
            s = v.settings
            H, W = int(s.height), int(s.width)
            out = History(torch.empty((H, W, 3), **opts), torch.empty((H, W, 1), **opts),
                          torch.empty((H, W, 1), **opts), torch.empty((H, W, 1), **opts))
            self.outputs.append(out)
            cv = self.c_views[i]
            cv.w2c[:] = _host(s.w2c, 16).tolist()
            cv.cam_position[:] = _host(s.cam_position, 3).tolist()
            cv.width, cv.height = W, H
            cv.focal_x, cv.focal_y = float(s.focal_x), float(s.focal_y)
            cv.center_x, cv.center_y = float(s.center_x), float(s.center_y)
            cv.near_plane, cv.far_plane = float(s.near_plane), float(s.far_plane)
            cv.proper_antialiasing = int(bool(s.proper_antialiasing))
            cv.supersampling_factor = int(s.supersampling_factor)
            cv.spp = int(v.spp)
            cv.temporal_mode = int(v.temporal_mode)
            cv.bin_size = int(v.bin_size)
            cv.use_gps = int(bool(s.use_gps))
            cv.hybrid_scheduling = int(bool(s.hybrid_scheduling))
            cv.spatial_reuse = int(bool(s.spatial_reuse))
            cv.max_gps_points = int(v.max_gps_points)
            if v.blue_noise is not None and v.blue_noise.numel() > 0:
                bn = v.blue_noise.to(device=self.device, dtype=torch.uint8).contiguous()
                if bn.numel() != 64 * 128 * 128:
                    raise ValueError(f"view {i}: blue_noise must be a (64,128,128) uint8 volume")
                self._keep.append(bn)
                cv.blue_noise, cv.blue_noise_numel = bn.data_ptr(), bn.numel()
            if s.spatial_reuse:
                if v.k_buffer is None:
                    raise ValueError(f"view {i}: spatial reuse needs the view's k-buffer (create_k_buffer)")
                kp, ko = v.k_buffer
                P = W * H * max(1, int(s.supersampling_factor)) ** 2
                if (kp.shape != (_native.k_buffer_size, P, 2) or ko.shape != (_native.k_buffer_size, P, 4)
                        or kp.dtype != torch.int32 or ko.dtype != torch.float32 or not kp.is_contiguous()
                        or not ko.is_contiguous() or kp.device != self.device or ko.device != self.device):
                    raise ValueError(f"view {i}: k-buffer does not match the settings (use create_k_buffer)")
                cv.k_buffer_primitives, cv.k_buffer_occlusion = kp.data_ptr(), ko.data_ptr()
            v2p = np.eye(4, dtype=np.float32) if v.view_to_previous_view is None else _host(v.view_to_previous_view, 16)
            cv.view_to_previous_view[:] = np.asarray(v2p, np.float32).reshape(16).tolist()
            if int(v.temporal_mode) != 0:
                if v.history is None:
                    raise ValueError(f"view {i}: STATIC and REPROJECT need a history")
                hist = [_native._f32(h, n, self.device) for h, n in zip(v.history, History._fields)]
                self._keep.extend(hist)
                cv.history = _native.sgs_history(*[h.data_ptr() for h in hist])
            cv.out = _native.sgs_history(*[o.data_ptr() for o in out])
        self.set_frames([v.frame_idx for v in self.views])
        nbytes = _native._lib.sgs_render_batch_workspace_size(self.c_views, self.n_views,
                                                               self.gaussians.n_gaussians, self._tp)
        self.workspace = _native._workspace(nbytes, self.device)
        tb = _native._lib.sgs_batch_view_table_bytes(self.n_views)
        self._table_host = torch.empty(tb, dtype=torch.uint8, pin_memory=True)
        self.table = torch.empty(tb, dtype=torch.uint8, device=self.device)
        self._copied = None
        self.upload_table()

    def set_frames(self, frames: Sequence[int | None]) -> None:
        """Sets each view's first RNG frame (None: take spp frames from the global counter).
        Call upload_table() afterwards to make a captured graph see them."""
        for i, f in enumerate(frames):
            if f is None:
                f = _native.next_frame_idx()
                for _ in range(self.c_views[i].spp - 1):
                    _native.next_frame_idx()
            self.c_views[i].frame_idx = int(f) & 0xFFFFFFFF

    def upload_table(self) -> None:
        """Encodes the views into the pinned staging buffer and copies it to the device table
        on the current stream (after the previous copy has read the staging buffer)."""
        if self._copied is not None:
            self._copied.synchronize()
        rc = _native._lib.sgs_batch_encode_views(self.c_views, self.n_views, self._table_host.data_ptr())
        _native._check(rc, "sgs_batch_encode_views")
        with torch.cuda.device(self.device):
            self.table.copy_(self._table_host, non_blocking=True)
            self._copied = torch.cuda.Event()
            self._copied.record()

    def render(self, stage_events: Sequence["torch.cuda.Event"] | None = None) -> list[History]:
        """stage_events: 5 recorded-once events marked before the clears, before the
        accumulation, before its main kernel (tile rasteriser or persistent scatter), after
        the accumulation and after the resolve (for timing)."""
        ev = None
        if stage_events is not None:
            if len(stage_events) != 5:
                raise ValueError("stage_events takes 5 events")
            ev = (C.c_void_p * 5)(*[e.cuda_event for e in stage_events])
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            rc = _native._lib.sgs_render_batch(C.byref(self.gaussians), self.c_views, self.n_views,
                                               self.table.data_ptr(), self._tp, self.workspace.data_ptr(),
                                               self.workspace.numel(), ev, stream)
        _native._check(rc, "sgs_render_batch")
        return self.outputs

    def tile_stats(self) -> tuple[int, int, bool]:
        """(instances binned by the last render's tile schedule, (view, Gaussian) pairs that
        touched a tile, whether the instances overflowed max_tile_instances); waits for the
        current stream."""
        n, vis, o = C.c_uint64(0), C.c_uint64(0), C.c_uint32(0)
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            rc = _native._lib.sgs_batch_tile_stats(self.workspace.data_ptr(), stream, C.byref(n), C.byref(vis),
                                                   C.byref(o))
        _native._check(rc, "sgs_batch_tile_stats")
        return int(n.value), int(vis.value), bool(o.value)

    def capture(self) -> "torch.cuda.CUDAGraph":
        """Captures render() into a graph (HIP graph on ROCm); replay() re-renders with the
        view table's current contents."""
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self.render()  # warm-up on the capture stream
        torch.cuda.current_stream(self.device).wait_stream(side)
        with torch.cuda.graph(graph):
            self.render()
        return graph


def split_views(views: Sequence[BatchView], groups: int) -> list[list[int]]:
    """min(groups, len(views)) non-empty contiguous groups of view indices, each boundary placed
    where the framebuffer pixel-samples before it are closest to its share of the total."""
    prefix = [0]
    for v in views:
        ss = max(1, int(v.settings.supersampling_factor))
        prefix.append(prefix[-1] + int(v.settings.width) * ss * int(v.settings.height) * ss * int(v.spp))
    n = len(views)
    g = max(1, min(int(groups), n))
    bounds = [0]
    for k in range(1, g):
        target = prefix[n] * k / g
        bounds.append(min(range(bounds[-1] + 1, n - (g - k) + 1), key=lambda b: abs(prefix[b] - target)))
    bounds.append(n)
    return [list(range(bounds[k], bounds[k + 1])) for k in range(g)]


class StreamedBatchRenderer:
    """The views as ``streams`` contiguous groups (split_views), each rendered by its own
    BatchRenderer (own workspace) on its own HIP stream of default priority, forked from and
    joined to the caller's current stream: one group's binning passes (memory bound) run beside
    another group's rasteriser (VALU bound). A view's result depends only on the view, so the
    outputs equal one BatchRenderer over all views bit for bit, and frames taken from the
    process-wide counter are drawn in view order as there. ``tuning`` is one BatchTuning for
    every group or a list with one per group (e.g. each group's instance capacity)."""

    def __init__(self, means, scales, rotations, opacities, sh_coefficients, views: Sequence[BatchView],
                 tuning: BatchTuning | Sequence[BatchTuning] | None = None, streams: int = 2,
                 cov3d_cache: torch.Tensor | None = None):
        import torch
        if streams < 1:
            raise ValueError("streams must be at least 1")
        self.device = means.device
        self.views = list(views)
        self.groups = split_views(self.views, streams)
        tunings = [tuning] * len(self.groups) if tuning is None or isinstance(tuning, BatchTuning) else list(tuning)
        if len(tunings) != len(self.groups):
            raise ValueError(f"{len(tunings)} tunings for {len(self.groups)} view groups")
        self.parts = [BatchRenderer(means, scales, rotations, opacities, sh_coefficients,
                                    [self.views[i] for i in g], t, cov3d_cache)
                      for g, t in zip(self.groups, tunings)]
        self.streams = [torch.cuda.Stream(device=self.device) for _ in self.parts]
        self.outputs: list[History] = [None] * len(self.views)  # type: ignore[list-item]
        for g, r in zip(self.groups, self.parts):
            for j, i in enumerate(g):
                self.outputs[i] = r.outputs[j]

    @property
    def workspace_bytes(self) -> int:
        return sum(r.workspace.numel() for r in self.parts)

    def set_frames(self, frames: Sequence[int | None]) -> None:
        """As BatchRenderer.set_frames, per view of the whole list."""
        for g, r in zip(self.groups, self.parts):
            r.set_frames([frames[i] for i in g])

    def upload_table(self) -> None:
        for r in self.parts:
            r.upload_table()

    def render(self) -> list[History]:
        import torch
        cur = torch.cuda.current_stream(self.device)
        for s in self.streams:
            s.wait_stream(cur)
        for r, s in zip(self.parts, self.streams):
            with torch.cuda.stream(s):
                r.render()
        for s in self.streams:
            cur.wait_stream(s)
        return self.outputs

    def tile_stats(self) -> tuple[int, int, bool]:
        """Summed over the groups (BatchRenderer.tile_stats)."""
        stats = [r.tile_stats() for r in self.parts]
        return sum(x[0] for x in stats), sum(x[1] for x in stats), any(x[2] for x in stats)

    def part_tile_stats(self) -> list[tuple[int, int, bool]]:
        return [r.tile_stats() for r in self.parts]

    def capture(self) -> "torch.cuda.CUDAGraph":
        """Captures render() (every group's stream, forked from and joined to the capturing
        stream) into one graph."""
        import torch
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self.render()  # warm-up on the capture stream
        torch.cuda.current_stream(self.device).wait_stream(side)
        with torch.cuda.graph(graph):
            self.render()
        return graph


def render_batch(means, scales, rotations, opacities, sh_coefficients, views: Sequence[BatchView],
                 tuning: BatchTuning | None = None, cov3d_cache: torch.Tensor | None = None) -> list[History]:
    """Renders every view (all of its samples) in one persistent launch; fresh outputs."""
    r = BatchRenderer(means, scales, rotations, opacities, sh_coefficients, views, tuning, cov3d_cache)
    return r.render()
