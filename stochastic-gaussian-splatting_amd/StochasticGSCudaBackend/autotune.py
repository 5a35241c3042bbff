"""Per-scene autotuning of the scheduling knobs.

None of the knobs searched here changes a result bit (the framebuffer winners are the
closest accepted fragment per (sub)pixel whatever the schedule), so tuning is a pure timing
search: a coordinate descent over the candidate values of each knob, timing the stage the
knobs act on with HIP events on the current stream, a few repetitions per trial, keeping the
median. The max_gps_points cap does change results and is never tuned.

* ``tune_render`` tunes the single-view path for one camera: the schedule (tiles or the
  reference's scatter) for stochastic transparency without hybrid scheduling, tile size,
  binning pixel and point thresholds, max_binned_gaussians and the sequential threshold
  (Tuning), timing the accumulation stage (sgs_accumulate).
* ``tune_batch`` tunes a batched render: schedule and samples per LDS pass of the tile
  schedule (BatchTuning), the views' bin sizes (16-pixel bins for framebuffers up to a
  pixel-count threshold, 32 above; ``TuneResult.views`` holds the views with the chosen bin
  sizes) and the number of view groups rendered on concurrent streams
  (StreamedBatchRenderer; ``TuneResult.streams``, with one BatchTuning per group in
  ``TuneResult.part_tunings``), timing whole renders; each batch's instance capacity is set
  from its measured instance count with a margin, which shrinks the workspace.
"""
from __future__ import annotations

import statistics
from dataclasses import dataclass, field
from typing import Callable, Sequence

import torch

from . import _native
from .render import Settings, Tuning


@dataclass
class TuneResult:
    best: object
    ms: float
    trials: list = field(default_factory=list)  # (config, ms) in evaluation order
    views: list | None = None  # tune_batch: the views with the chosen bin sizes
    streams: int = 1  # tune_batch: view groups on concurrent streams (1: one BatchRenderer)
    part_tunings: list | None = None  # tune_batch with streams > 1: one BatchTuning per group
    # tune_batch's tile instance capacities are measured on the tuned poses; a render from
    # other poses that exceeds them falls back to the (bit-identical, much slower) scatter path


def coordinate_search(evaluate: Callable[[dict], float], space: dict[str, Sequence], start: dict,
                      passes: int = 2, min_gain: float = 0.0) -> tuple[dict, float, list]:
    """Minimises ``evaluate`` over the product of ``space`` by coordinate descent from
    ``start``: each pass tries every value of each knob with the others fixed and keeps the
    best; stops after ``passes`` passes or a pass without improvement. Each configuration
    is evaluated at most once. A candidate replaces the current best only if it is faster by
    more than the fraction ``min_gain`` (timing noise then does not move the result away from
    the defaults)."""
    seen: dict[tuple, float] = {}
    trials: list = []

    def cost(cfg: dict) -> float:
        key = tuple(sorted(cfg.items()))
        if key not in seen:
            seen[key] = float(evaluate(dict(cfg)))
            trials.append((dict(cfg), seen[key]))
        return seen[key]

    best = dict(start)
    best_ms = cost(best)
    for _ in range(passes):
        improved = False
        for knob, values in space.items():
            for v in values:
                cand = dict(best)
                cand[knob] = v
                ms = cost(cand)
                if ms < best_ms * (1.0 - min_gain):
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    return best, best_ms, trials


def _time_ms(fn: Callable[[], None], repeats: int) -> float:
    stream = torch.cuda.current_stream()
    fn()  # warm-up (also allocates)
    times = []
    for _ in range(repeats):
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        fn()
        end.record(stream)
        end.synchronize()
        times.append(start.elapsed_time(end))
    return statistics.median(times)


RENDER_SPACE = {
    "tile_size": [16, 32, 64],
    "binning_pixel_threshold": [128, 256, 512, 1024, 2048, 4096],
    "binning_point_threshold": [128, 256, 512, 1024, 2048],
    "max_binned_gaussians": [1 << 14, 1 << 16, 1 << 18, 1 << 20],
    "sequential_threshold": [2, 4, 8, 16],
    "schedule": [1, 2],
    "bin_size": [16, 32],
}


def tune_render(means, scales, rotations, opacities, sh_coefficients, settings: Settings, *, frame_idx: int = 0,
                repeats: int = 3, space: dict | None = None, passes: int = 2) -> TuneResult:
    """Fastest Tuning for the accumulation of ``settings`` (its own GPS / hybrid choice).
    Knobs that do not act on that schedule are left at their defaults."""
    if space is None:
        # schedule and bin_size act only on stochastic transparency without hybrid scheduling
        space = {k: v for k, v in RENDER_SPACE.items() if k not in ("schedule", "bin_size")}
        if not settings.hybrid_scheduling:
            space = {"sequential_threshold": space["sequential_threshold"]}
            if not settings.use_gps:
                space["schedule"] = list(RENDER_SPACE["schedule"])
                space["bin_size"] = list(RENDER_SPACE["bin_size"])
        elif settings.use_gps:
            space.pop("tile_size")
            space.pop("binning_pixel_threshold")
        else:
            space.pop("binning_point_threshold")
    n = int(means.shape[0])
    space = {k: [v for v in vals if k != "max_binned_gaussians" or v <= max(n, 1)] or [max(n, 1)]
             for k, vals in space.items()}
    w2c = settings.w2c.to(device=means.device, dtype=torch.float32)
    cam = settings.cam_position.to(device=means.device, dtype=torch.float32)
    fb = torch.empty((settings.height * max(1, settings.supersampling_factor),
                      settings.width * max(1, settings.supersampling_factor)), dtype=torch.int64, device=means.device)
    args = (means, scales, rotations, opacities, sh_coefficients, w2c, cam, *settings.as_tuple()[2:])

    def evaluate(cfg: dict) -> float:
        tuning = Tuning(**cfg)
        return _time_ms(lambda: _native.accumulate(*args, frame_idx=frame_idx, tuning=tuning, framebuffer=fb), repeats)

    start = {k: 0 for k in space}  # 0 = library default
    best, ms, trials = coordinate_search(evaluate, {k: [0] + list(v) for k, v in space.items()}, start, passes)
    return TuneResult(Tuning(**best), ms, [(Tuning(**c), t) for c, t in trials])


# small_bin_max_pixels: views whose framebuffer has at most this many pixels get 16-pixel
# bins, the others 32 (-1: each view's own bin_size, by default the library's choice);
# streams: view groups rendered by separate batches on concurrent streams
BATCH_SPACE = {"schedule": [1, 2], "tile_samples": [0, 8, 16], "small_bin_max_pixels": [-1, 0, 1 << 18, 1 << 21],
               "streams": [1, 2, 4]}


def views_with_bins(views, small_bin_max_pixels: int) -> list:
    """The views with 16-pixel bins for framebuffers of at most ``small_bin_max_pixels``
    pixels and 32-pixel bins above (-1: unchanged)."""
    import dataclasses
    if small_bin_max_pixels < 0:
        return list(views)
    out = []
    for v in views:
        ss = max(1, int(v.settings.supersampling_factor))
        px = int(v.settings.width) * ss * int(v.settings.height) * ss
        out.append(dataclasses.replace(v, bin_size=16 if px <= small_bin_max_pixels else 32))
    return out


def tune_batch(means, scales, rotations, opacities, sh_coefficients, views, *, repeats: int = 2,
               space: dict | None = None, capacity_margin: float = 1.25) -> TuneResult:
    """Fastest BatchTuning for rendering ``views`` (BatchView list) in one call; the tile
    instance capacity of the result (of each group's, with streams > 1) is the measured
    instance count x ``capacity_margin``.

    That capacity belongs to the tuned camera poses: other poses over the same scene can need
    more tile instances, and a batch that exceeds the capacity still renders the same bits but
    through the scatter fallback, an order of magnitude slower. Re-tune when the poses change
    materially, check ``tile_stats()[2]`` (overflow) after a render, or clear the capacity
    (``max_tile_instances=0``: the library's per-pixel default, usually larger) to trade
    memory for headroom."""
    from .batch import BatchRenderer, BatchTuning, StreamedBatchRenderer
    space = dict(BATCH_SPACE if space is None else space)

    def renderer(cfg: dict, tuning=None):
        cfg = dict(cfg)
        vs = views_with_bins(views, cfg.pop("small_bin_max_pixels", -1))
        streams = int(cfg.pop("streams", 1))
        tuning = BatchTuning(**cfg) if tuning is None else tuning
        if streams > 1:
            return StreamedBatchRenderer(means, scales, rotations, opacities, sh_coefficients, vs, tuning, streams)
        return BatchRenderer(means, scales, rotations, opacities, sh_coefficients, vs, tuning)

    def evaluate(cfg: dict) -> float:
        r = renderer(cfg)
        ms = _time_ms(r.render, repeats)
        del r
        return ms

    # a knob leaves its first (default) value only for a gain above 1 %: repeated C4 trials
    # differ by about 0.5 %
    best, ms, trials = coordinate_search(evaluate, space, {k: v[0] for k, v in space.items()}, passes=1,
                                         min_gain=0.01)
    best_views = views_with_bins(views, best.get("small_bin_max_pixels", -1))
    streams = int(best.get("streams", 1))
    tuning = BatchTuning(**{k: v for k, v in best.items() if k not in ("small_bin_max_pixels", "streams")})
    part_tunings = None
    if tuning.schedule in (0, 1):
        probe = renderer(best, tuning)
        probe.render()
        stats = probe.part_tile_stats() if streams > 1 else [probe.tile_stats()]
        del probe
        caps = [tuning._replace(max_tile_instances=min(int(n * capacity_margin) + 1, (1 << 31)))
                if not o and n > 0 else tuning for n, _, o in stats]
        if streams > 1:
            part_tunings = caps
        else:
            tuning = caps[0]
    return TuneResult(tuning, ms, [(dict(c), t) for c, t in trials], best_views, streams, part_tunings)
