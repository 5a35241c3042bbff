#!/usr/bin/env python3
"""Benchmark: pixel-samples/s of the stochastic splatting renderer on the 6M-Gaussian,
64-view workload (uneven resolutions from 256x256 to 2048x1152, spp in {1, 4, 16}).

One step renders every view with all of its samples (RESET then STATIC accumulation, like
the reference's SPP loop in Renderer.py:85-87), inputs resident in HBM before timing.
Prints one JSON line. Extra fields: the plain per-view path beside the batched one, the
roofline of the dominant kernel, and the CPU oracle timed on this host.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--views 64] [--gaussians 6000000]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
for _p in (str(REPO), str(REPO / "stochastic-gaussian-splatting_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--gaussians", type=int, default=6_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["plain", "batched", "all"], default="all")
    return ap.parse_args()


def view_settings(cam: dict, device):
    import torch
    from StochasticGSCudaBackend import Settings
    return Settings(torch.from_numpy(cam["w2c"]).to(device), torch.from_numpy(cam["cam_position"]).to(device),
                    cam["width"], cam["height"], cam["focal_x"], cam["focal_y"], cam["center_x"], cam["center_y"],
                    cam["near_plane"], cam["far_plane"], False, 1, False, False, False)


def samples_per_step(views) -> int:
    return sum(v["width"] * v["height"] * v["spp"] for v in views)


def run_plain(scene, views, settings, k_buffers, frame_base: int):
    """Per view: spp sequential render() calls of the drop-in API (RESET, then STATIC)."""
    import torch
    from StochasticGSCudaBackend import History, TemporalMode, render
    eye = torch.eye(4, device=scene["means"].device)
    empty = History(*(torch.empty(0) for _ in range(4)))
    out = []
    frame = frame_base
    for v, (cam, s) in enumerate(zip(views, settings)):
        hist = empty
        mode = TemporalMode.RESET
        for _ in range(cam["spp"]):
            hist = render(scene["means"], scene["scales"], scene["rotations"], scene["opacities"], scene["sh"], s,
                          k_buffers[v], hist, eye, mode, frame_idx=frame)
            frame += 1
            mode = TemporalMode.STATIC
        out.append(hist)
    return out


def time_scatter_launches(scene, views, reps: int = 1):
    """Average duration of the accumulation launch (framebuffer clear + ST scatter kernel),
    measured with HIP events on the stream the library launches on, and its counters."""
    import torch
    from StochasticGSCudaBackend import _native
    dev = scene["means"].device
    stream = torch.cuda.current_stream(dev)
    fbs, args = [], []
    for cam in views:
        w2c = torch.from_numpy(cam["w2c"]).to(dev)
        cp = torch.from_numpy(cam["cam_position"]).to(dev)
        a = (scene["means"], scene["scales"], scene["rotations"], scene["opacities"], scene["sh"], w2c, cp,
             cam["width"], cam["height"], cam["focal_x"], cam["focal_y"], cam["center_x"], cam["center_y"],
             cam["near_plane"], cam["far_plane"], False, 1, False, False, False)
        args.append(a)
        fbs.append(torch.empty((cam["height"], cam["width"]), dtype=torch.int64, device=dev))
    # algorithmic work per launch from the counting variant (outside the timed region)
    counters = torch.zeros(4, dtype=torch.int64, device=dev)
    work = []
    for a, fb in zip(args, fbs):
        _native.accumulate(*a, frame_idx=0, framebuffer=fb, counters=counters)
        work.append(counters.cpu().tolist())
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(len(views) * reps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(len(views) * reps)]
    k = 0
    for r in range(reps):
        for a, fb in zip(args, fbs):
            starts[k].record(stream)
            _native.accumulate(*a, frame_idx=1 + r, framebuffer=fb)
            ends[k].record(stream)
            k += 1
    torch.cuda.synchronize()
    ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return ms, work


def roofline_scatter(scene_n: int, views, ms, work):
    """Algorithmic bytes of one accumulation launch: 12 B of means for every Gaussian (depth
    cull), 32 B more (scales, rotation, opacity) for every Gaussian inside [near, far], 8 B
    clear of each framebuffer word, 8 B per atomicMin issued."""
    n_views = len(views)
    reps = len(ms) // n_views
    bytes_total, t_total, atomics = 0.0, 0.0, 0
    for v, cam in enumerate(views):
        tested, accepted, projected, _ = work[v]
        P = cam["width"] * cam["height"]
        b = 12.0 * scene_n + 32.0 * scene_n + 8.0 * P + 8.0 * accepted
        for r in range(reps):
            bytes_total += b
            t_total += ms[r * n_views + v] * 1e-3
        atomics += accepted * reps
    achieved = bytes_total / t_total / 1e9
    return {
        "kernel": "sgs::st_scatter_kernel (+ framebuffer clear)",
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "avg_launch_ms": round(t_total * 1e3 / len(ms), 4),
        "atomics_per_s": round(atomics / t_total, 1),
        "fragments_tested_per_launch": int(sum(w[0] for w in work) / n_views),
    }


def cpu_baseline(scene_np, views, budget_s: float):
    """The oracle (scalar C, OpenMP over Gaussians) rendering views of the same workload,
    one sample each, until the time budget is spent."""
    from oracle import oracle
    threads = oracle.n_threads_available()
    sc = oracle.OracleScene(scene_np["means"], scene_np["scales"], scene_np["rotations"], scene_np["opacities"],
                            scene_np["sh"])
    order = sorted(range(len(views)), key=lambda v: views[v]["width"] * views[v]["height"])
    done, samples, t0 = [], 0, time.perf_counter()
    for v in order:
        cam = views[v]
        oracle.render(sc, cam, 0, n_threads=threads)
        done.append(v)
        samples += cam["width"] * cam["height"]
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    res = sorted({f"{views[v]['width']}x{views[v]['height']}" for v in done})
    return {"value": round(samples / dt / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"{len(done)} views of the workload ({', '.join(res)}), 1 spp each, full scene, "
                      f"accumulate+sample+resolve, {dt:.1f} s"}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (sharding is described in DESIGN.md)"}))
        return 2
    import numpy as np
    import torch
    from StochasticGSCudaBackend import create_k_buffer
    from StochasticGSCudaBackend.synthetic import make_scene, scaled, scene_c4, views_c4

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t0 = time.perf_counter()
    spec = scaled(scene_c4(), args.gaussians)
    scene_np = make_scene(spec)
    scene = {k: torch.from_numpy(v).to(dev) for k, v in scene_np.items()}
    views = views_c4(args.views)
    settings = [view_settings(v, dev) for v in views]
    k_buffers = [create_k_buffer(s) for s in settings]
    torch.cuda.synchronize()
    log(f"scene {spec.n} Gaussians, {len(views)} views, {samples_per_step(views) / 1e6:.1f} M pixel-samples/step "
        f"(setup {time.perf_counter() - t0:.1f} s)")

    S = samples_per_step(views)
    result = {}
    # ---- plain: per-view sequential render() calls -------------------------------------
    for w in range(args.warmup):
        run_plain(scene, views, settings, k_buffers, frame_base=1_000_000 * (w + 1))
        torch.cuda.synchronize()
        log(f"plain warmup {w + 1}/{args.warmup} done")
    step_ms = []
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        run_plain(scene, views, settings, k_buffers, frame_base=10_000 * k)
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - ts) * 1e3)
        log(f"plain step {k + 1}/{args.steps}: {step_ms[-1]:.1f} ms")
    plain_s = time.perf_counter() - t_all
    plain_value = S * args.steps / plain_s / 1e6
    result["plain"] = {"value": round(plain_value, 2), "ms_per_step": round(plain_s * 1e3 / args.steps, 2),
                       "p50_step_ms": round(float(np.percentile(step_ms, 50)), 2),
                       "p99_step_ms": round(float(np.percentile(step_ms, 99)), 2)}

    # ---- roofline of the dominant kernel ----------------------------------------------------
    ms, work = time_scatter_launches(scene, views)
    roof = roofline_scatter(spec.n, views, ms, work)
    log(f"scatter launches: avg {roof['avg_launch_ms']} ms, {roof['achieved']} GB/s algorithmic")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene_np, views, args.cpu_seconds)
        log(f"cpu oracle: {cpu['value']} Msamples/s on {cpu['cores']} threads")

    value = plain_value
    ms_per_step = plain_s * 1e3 / args.steps
    line = {
        "metric": "pixel-samples/s (Msamples/s) batched vs plain, 6M Gaussians; p50/p99 frame ms",
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{spec.n} Gaussians (8-component mixture, SH3), {len(views)} views on a ring, "
                               f"resolutions 256x256..2048x1152, spp in {{1,4,16}}, ST scatter, ss=1",
                   "path": "plain per-view render() calls", "pixel_samples_per_step": S},
        "plain": result["plain"],
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
