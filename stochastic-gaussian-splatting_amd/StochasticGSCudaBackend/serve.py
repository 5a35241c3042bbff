"""Serving loop: continuous batching of per-stream render requests over one resident scene.

A *stream* is a client viewport (fixed resolution) whose frames are resolved with the
temporal REPROJECT mode against that stream's own history (RESET for its first frame), and
with spatial reuse over the stream's own k-buffer when its StreamSpec asks for it (the
spatio-temporal resolve), as the reference's Renderer.py:73-87 does for one camera path. A *request* is one
frame of one stream: a camera pose and an spp count.

The server keeps the Gaussians and a ring of ``chain + 1`` output buffers per stream
resident in device memory (frame t reads the buffer frame t-1 wrote and writes the next
one), and renders the requests that have arrived in batches through sgs_render_batch. Up to
``inflight`` batches are on the GPU at once, each on its own stream with its own
preallocated workspace arena; a request is admitted into the batch being formed while

  * the batch has fewer than ``max_views`` views,
  * the batch's workspace (framebuffers of every sample, tile lists, resolve buffers, as
    sgs_render_batch_workspace_size reports) still fits the arena, and
  * its stream has no frame in another batch on the GPU and fewer than ``chain`` frames in
    this one (consecutive frames of a stream share a batch: the library resolves frame t+1
    after frame t, whose outputs are its history).

Requests that arrive while batches are on the GPU are admitted as soon as a slot frees. A
batching window ``max_wait`` (seconds) holds a free slot until the oldest pending request has
waited that long or enough requests are pending to fill a batch: batches come out larger and
the latency distribution tighter (wait + batch time instead of a queue's exponential tail).
Each request's random numbers are fixed by (stream, frame): frame t of stream s uses RNG
frames (s << 24) + (sum of the spp of frames 0..t-1 of s) onwards, so a result is the same
whichever batch it lands in, and equal to rendering the stream alone.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Sequence

import numpy as np
import torch

from . import _native
from .batch import BatchTuning
from .render import History, TemporalMode

STREAM_SHIFT = 24  # RNG frames per stream: 2^24


@dataclass(frozen=True)
class StreamSpec:
    """A client viewport: image size and pinhole intrinsics (image pixels)."""
    width: int
    height: int
    focal_x: float
    focal_y: float
    center_x: float
    center_y: float
    near_plane: float = 0.2
    far_plane: float = 10000.0
    proper_antialiasing: bool = False
    supersampling_factor: int = 1
    spatial_reuse: bool = False   # spatio-temporal resolve: the stream keeps a k-buffer on the device
    use_gps: bool = False


@dataclass
class Request:
    stream: int
    w2c: np.ndarray           # (4,4) float32 world-to-view
    cam_position: np.ndarray  # (3,) float32
    spp: int
    arrival: float = 0.0      # seconds on the server clock
    # set by the server
    frame: int = -1           # index of this frame in its stream
    frame_idx: int = 0        # first RNG frame
    mode: int = 0
    v2p: np.ndarray | None = None  # view_to_previous_view (REPROJECT)
    admitted: float = math.nan
    done: float = math.nan
    output: History | None = None   # copy of the outputs (keep_outputs)
    out_buffers: History | None = None  # the stream's ring slot this frame writes

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Stream:
    spec: StreamSpec
    ring: list               # chain + 1 History sets
    k_buffer: tuple | None   # spatial reuse: the stream's k-buffer (create_k_buffer layout)
    head: int = 0            # ring slot of the latest admitted frame's output
    frames: int = 0          # frames submitted
    finished: int = 0        # frames finished (they finish in order)
    spp_used: int = 0        # RNG frames consumed by submitted frames
    prev_w2c: np.ndarray | None = None  # pose of the last submitted frame
    busy: bool = False       # frames of it are in a batch on the GPU


@dataclass
class _Slot:
    stream: torch.cuda.Stream
    workspace: torch.Tensor
    table_host: torch.Tensor
    table: torch.Tensor
    views: C.Array
    requests: list = field(default_factory=list)
    done_event: torch.cuda.Event | None = None


def admit(pending: deque, can_take: Callable[[Request, int], bool], fits: Callable[[Request, int, int], bool],
          max_views: int) -> tuple[list, deque]:
    """Admission policy of one batch: takes pending requests in arrival order while the
    batch has room (max_views) and memory (``fits(req, position, k)``). ``can_take(req, k)``
    decides whether a request can join when k frames of its stream are already in the batch;
    once a stream's request is refused, its later requests wait too (a stream's frames run in
    order). Returns (admitted, still pending in arrival order)."""
    chosen: list[Request] = []
    kept: deque[Request] = deque()
    in_batch: dict[int, int] = {}
    blocked: set[int] = set()
    while pending and len(chosen) < max_views:
        req = pending.popleft()
        k = in_batch.get(req.stream, 0)
        if req.stream in blocked or not can_take(req, k):
            kept.append(req)
            blocked.add(req.stream)
            continue
        if not fits(req, len(chosen), k):
            if not chosen:
                raise RuntimeError(f"a frame of stream {req.stream} alone exceeds the workspace arena")
            kept.append(req)
            break
        chosen.append(req)
        in_batch[req.stream] = k + 1
    kept.extend(pending)
    return chosen, kept


def _history(spec: StreamSpec, device) -> History:
    opts = dict(dtype=torch.float32, device=device)
    H, W = spec.height, spec.width
    return History(torch.zeros((H, W, 3), **opts), torch.zeros((H, W, 1), **opts), torch.zeros((H, W, 1), **opts),
                   torch.zeros((H, W, 1), **opts))


def _k_buffer(spec: StreamSpec, device) -> tuple | None:
    if not spec.spatial_reuse:
        return None
    P = spec.width * spec.height * max(1, int(spec.supersampling_factor)) ** 2
    return (torch.full((_native.k_buffer_size, P, 2), -1, dtype=torch.int32, device=device),
            torch.zeros((_native.k_buffer_size, P, 4), dtype=torch.float32, device=device))


def _hist_struct(h: History) -> _native.sgs_history:
    return _native.sgs_history(*[t.data_ptr() for t in h])


class Server:
    """Continuous-batching renderer for ``streams`` over one resident scene."""

    def __init__(self, means, scales, rotations, opacities, sh_coefficients, streams: Sequence[StreamSpec],
                 max_views: int = 32, inflight: int = 2, chain: int = 3, max_wait: float = 0.0,
                 arena_bytes: int | None = None,
                 tuning: BatchTuning | None = None, keep_outputs: bool = False,
                 clock: Callable[[], float] = time.perf_counter, cov3d_cache: torch.Tensor | None = None):
        if max_views < 1 or inflight < 1 or chain < 1:
            raise ValueError("max_views, inflight and chain must be >= 1")
        self.chain = chain
        self.max_wait = float(max_wait)
        self.device = means.device
        self._keep: list = []
        self.gaussians = _native.gaussians_struct(means, scales, rotations, opacities, sh_coefficients, self._keep,
                                                  cov3d_cache)
        self.n_gaussians = int(self.gaussians.n_gaussians)
        self.streams = [_Stream(s, [_history(s, self.device) for _ in range(chain + 1)], _k_buffer(s, self.device))
                        for s in streams]
        if len(self.streams) > (1 << (32 - STREAM_SHIFT)):
            raise ValueError(f"at most {1 << (32 - STREAM_SHIFT)} streams")
        self.max_views = max_views
        self.keep_outputs = keep_outputs
        self.clock = clock
        if arena_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            arena_bytes = int(0.5 * free) // inflight
        self.arena_bytes = int(arena_bytes) & ~255
        # The tile schedule's instance list takes up to half of the arena (14 bytes per
        # instance) rather than the library's per-view estimate: an overflowing batch falls
        # back to the much slower scatter schedule, which would show as a latency spike.
        tuning = BatchTuning() if tuning is None else tuning
        if tuning.max_tile_instances == 0:
            tuning = tuning._replace(max_tile_instances=max(1 << 16, min(1 << 31, self.arena_bytes // 32)))
        self.tuning = tuning.c()
        self._tp = C.byref(self.tuning)
        self.overflowed_batches = 0
        table_bytes = _native._lib.sgs_batch_view_table_bytes(max_views)
        self.slots: list[_Slot] = []
        for _ in range(inflight):
            self.slots.append(_Slot(stream=torch.cuda.Stream(device=self.device),
                                    workspace=_native._workspace(self.arena_bytes, self.device),
                                    table_host=torch.empty(table_bytes, dtype=torch.uint8, pin_memory=True),
                                    table=torch.empty(table_bytes, dtype=torch.uint8, device=self.device),
                                    views=(_native.sgs_view * max_views)()))
        self.pending: deque[Request] = deque()
        self.finished: list[Request] = []
        self.batches = 0
        self.batch_sizes: list[int] = []

    # ---------------------------------------------------------------- requests
    def submit(self, req: Request) -> Request:
        """Queues a request; fixes its frame number, RNG frames and temporal mode."""
        st = self.streams[req.stream]
        if req.spp < 1:
            raise ValueError("spp must be >= 1")
        if st.spp_used + req.spp > (1 << STREAM_SHIFT):
            raise ValueError(f"stream {req.stream} ran out of RNG frames")
        req.frame = st.frames
        req.frame_idx = (req.stream << STREAM_SHIFT) + st.spp_used
        req.mode = int(TemporalMode.RESET if st.frames == 0 else TemporalMode.REPROJECT)
        # view_to_previous_view (Renderer.py:84): previous w2c @ inverse(current w2c)
        if st.prev_w2c is not None:
            req.v2p = (st.prev_w2c.astype(np.float64) @ np.linalg.inv(req.w2c.astype(np.float64))).astype(np.float32)
        st.prev_w2c = np.asarray(req.w2c, np.float32)
        st.frames += 1
        st.spp_used += req.spp
        self.pending.append(req)
        return req

    # ---------------------------------------------------------------- batching
    def _fill_view(self, cv: _native.sgs_view, req: Request, k: int) -> None:
        """View of ``req``, the k-th frame of its stream in the batch being formed."""
        st = self.streams[req.stream]
        s = st.spec
        cv.w2c[:] = np.asarray(req.w2c, np.float32).reshape(16).tolist()
        cv.cam_position[:] = np.asarray(req.cam_position, np.float32).reshape(3).tolist()
        cv.width, cv.height = s.width, s.height
        cv.focal_x, cv.focal_y, cv.center_x, cv.center_y = s.focal_x, s.focal_y, s.center_x, s.center_y
        cv.near_plane, cv.far_plane = s.near_plane, s.far_plane
        cv.proper_antialiasing = int(s.proper_antialiasing)
        cv.supersampling_factor = int(s.supersampling_factor)
        cv.spp = int(req.spp)
        cv.frame_idx = int(req.frame_idx) & 0xFFFFFFFF
        cv.temporal_mode = req.mode
        v2p = np.eye(4, dtype=np.float32) if req.v2p is None else req.v2p
        cv.view_to_previous_view[:] = v2p.reshape(16).tolist()
        n = len(st.ring)
        hist, out = st.ring[(st.head + k) % n], st.ring[(st.head + k + 1) % n]
        cv.history = _hist_struct(hist) if req.mode != int(TemporalMode.RESET) else _native.sgs_history()
        cv.out = _hist_struct(out)
        cv.use_gps = int(s.use_gps)
        cv.spatial_reuse = int(s.spatial_reuse)
        cv.k_buffer_primitives, cv.k_buffer_occlusion = ((st.k_buffer[0].data_ptr(), st.k_buffer[1].data_ptr())
                                                         if st.k_buffer is not None else (None, None))
        req.out_buffers = out

    def _free_slot(self) -> _Slot | None:
        for slot in self.slots:
            if slot.done_event is None:
                return slot
        return None

    def _launch(self, slot: _Slot) -> bool:
        """Admits pending requests into ``slot`` and launches the batch; False if none fits."""

        def can_take(req: Request, k: int) -> bool:
            st = self.streams[req.stream]
            return not st.busy and req.frame == st.finished + k and k < self.chain

        def fits(req: Request, position: int, k: int) -> bool:
            self._fill_view(slot.views[position], req, k)
            return _native._lib.sgs_render_batch_workspace_size(slot.views, position + 1, self.n_gaussians,
                                                                self._tp) <= self.arena_bytes

        chosen, self.pending = admit(self.pending, can_take, fits, self.max_views)
        if not chosen:
            return False
        now = self.clock()
        n = len(chosen)
        rc = _native._lib.sgs_batch_encode_views(slot.views, n, slot.table_host.data_ptr())
        _native._check(rc, "sgs_batch_encode_views")
        with torch.cuda.device(self.device), torch.cuda.stream(slot.stream):
            slot.table.copy_(slot.table_host, non_blocking=True)
            rc = _native._lib.sgs_render_batch(C.byref(self.gaussians), slot.views, n, slot.table.data_ptr(),
                                               self._tp, slot.workspace.data_ptr(), slot.workspace.numel(), None,
                                               slot.stream.cuda_stream)
            _native._check(rc, "sgs_render_batch")
            slot.done_event = torch.cuda.Event()
            slot.done_event.record(slot.stream)
        for req in chosen:
            req.admitted = now
            st = self.streams[req.stream]
            st.busy = True
            st.head = (st.head + 1) % len(st.ring)
        slot.requests = chosen
        self.batches += 1
        self.batch_sizes.append(n)
        return True

    def _complete(self, slot: _Slot) -> None:
        now = self.clock()
        n_inst, visible, overflowed = C.c_uint64(), C.c_uint64(), C.c_uint32()
        rc = _native._lib.sgs_batch_tile_stats(slot.workspace.data_ptr(), slot.stream.cuda_stream, C.byref(n_inst),
                                               C.byref(visible), C.byref(overflowed))
        _native._check(rc, "sgs_batch_tile_stats")
        self.overflowed_batches += int(overflowed.value != 0)
        for req in slot.requests:
            st = self.streams[req.stream]
            st.busy = False
            st.finished += 1
            req.done = now
            if self.keep_outputs:
                # copied before a later batch of the stream can overwrite this ring slot
                with torch.cuda.device(self.device), torch.cuda.stream(slot.stream):
                    req.output = History(*[t.clone() for t in req.out_buffers])
                slot.stream.synchronize()
            self.finished.append(req)
        slot.requests = []
        slot.done_event = None

    def poll(self) -> int:
        """Completes every batch the GPU has finished; returns how many requests finished."""
        n = 0
        for slot in self.slots:
            if slot.done_event is not None and slot.done_event.query():
                n += len(slot.requests)
                self._complete(slot)
        return n

    def pump(self) -> None:
        """One scheduling round: complete finished batches, then launch into free slots once
        the batching window of the oldest pending request has passed (or a batch is full)."""
        self.poll()
        while self.pending:
            slot = self._free_slot()
            if slot is None:
                break
            if (len(self.pending) < self.max_views and self.max_wait > 0.0
                    and self.clock() - self.pending[0].arrival < self.max_wait):
                break
            if not self._launch(slot):
                break

    def drain(self) -> None:
        """Runs until every submitted request has finished (the batching window still applies)."""
        while self.pending or any(s.done_event is not None for s in self.slots):
            self.pump()
        torch.cuda.synchronize(self.device)

    # ---------------------------------------------------------------- real-time feed
    def run(self, requests: Sequence[Request], t0: float | None = None) -> list[Request]:
        """Feeds ``requests`` (``arrival`` in seconds after ``t0``, rewritten to the absolute
        clock) in real time and returns them when all have finished; latency = completion seen
        by the host - arrival."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = self.clock() if t0 is None else t0
        for r in reqs:
            r.arrival += t0
        i = 0
        while i < len(reqs) or self.pending or any(s.done_event is not None for s in self.slots):
            now = self.clock()
            while i < len(reqs) and reqs[i].arrival <= now:
                self.submit(reqs[i])
                i += 1
            self.pump()
        torch.cuda.synchronize(self.device)
        return reqs


def latency_stats(reqs: Sequence[Request]) -> dict:
    lat = np.array([r.latency for r in reqs]) * 1e3
    return {"requests": len(reqs), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "mean_ms": round(float(lat.mean()), 3),
            "max_ms": round(float(lat.max()), 3)}
