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
    ap.add_argument("--serve-seconds", type=float, default=6.0,
                    help="length of the serving run at 80%% of saturated load (0: skip serving)")
    ap.add_argument("--serve-streams", type=int, default=32)
    ap.add_argument("--serve-inflight", type=int, default=1, help="batches on the GPU at once")
    ap.add_argument("--serve-window-ms", type=float, default=35.0,
                    help="batching window: a free slot waits until the oldest request is this old")
    ap.add_argument("--only", choices=["plain", "batched", "all"], default="all",
                    help="run only one path (profiling); the JSON line then reports that path")
    ap.add_argument("--schedule", choices=["tiles", "scatter"], default="tiles",
                    help="batched accumulation: LDS tiles (default) or persistent scatter")
    ap.add_argument("--tile-samples", type=int, choices=[8, 16], default=8,
                    help="samples per LDS pass of the tile rasteriser (with --no-autotune)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --schedule/--tile-samples instead of tuning them on the workload first")
    ap.add_argument("--view-major", action="store_true",
                    help="batched work items view by view instead of Gaussian chunk by chunk")
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy 1 GiB once first, to calibrate HBM byte counters against a known size")
    ap.add_argument("--no-morton", action="store_true",
                    help="keep generation order instead of the Morton order scenes are stored in")
    return ap.parse_args()


def view_settings(cam: dict, device):
    import torch
    from StochasticGSCudaBackend import Settings
    return Settings(torch.from_numpy(cam["w2c"]).to(device), torch.from_numpy(cam["cam_position"]).to(device),
                    cam["width"], cam["height"], cam["focal_x"], cam["focal_y"], cam["center_x"], cam["center_y"],
                    cam["near_plane"], cam["far_plane"], False, 1, False, False, False)


def host_settings(cam: dict):
    import torch
    from StochasticGSCudaBackend import Settings
    return Settings(torch.from_numpy(cam["w2c"]), torch.from_numpy(cam["cam_position"]), cam["width"], cam["height"],
                    cam["focal_x"], cam["focal_y"], cam["center_x"], cam["center_y"], cam["near_plane"],
                    cam["far_plane"], False, 1, False, False, False)


def frames_for_step(views, frame_base: int) -> list[int]:
    """First RNG frame of every view, in the order run_plain assigns them."""
    out, f = [], frame_base
    for cam in views:
        out.append(f)
        f += cam["spp"]
    return out


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
    scatter = _native.Tuning(schedule=2)  # the reference's schedule (the default is the tile one)
    for r in range(reps):
        for a, fb in zip(args, fbs):
            starts[k].record(stream)
            _native.accumulate(*a, frame_idx=1 + r, framebuffer=fb, tuning=scatter)
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


def roofline_render_samples(scene_np, views, scene, step_ms, work, atomics_per_s):
    """The per-render-sample roofline of a whole step (every view, every sample): bytes
    B = 12 N + 32 N_z + 16 P + 204 U + 24 Q (1 + [not RESET]) per (view, sample), N_z the
    Gaussians inside [near, far], P framebuffer words, U distinct winners, Q output pixels, and
    A = accepted fragments, one 64-bit atomicMin each in a scatter; t_roof = max(B / 8 TB/s,
    A / R_atomic) with R_atomic the scatter's measured atomic rate. U comes from one
    tile-schedule framebuffer per view and A from the scatter's counters of one sample per view,
    both taken for each of the view's samples. Outside the timed region."""
    import numpy as np
    import torch
    from StochasticGSCudaBackend import _native
    means = scene_np["means"].astype(np.float64)
    n = means.shape[0]
    dev = scene["means"].device
    bytes_total, atomics = 0.0, 0.0
    for v, cam in enumerate(views):
        w2c = np.asarray(cam["w2c"], np.float64)
        depth = means @ w2c[2, :3] + w2c[2, 3]
        n_z = int(np.count_nonzero((depth >= cam["near_plane"]) & (depth <= cam["far_plane"])))
        fb = _native.accumulate(scene["means"], scene["scales"], scene["rotations"], scene["opacities"], scene["sh"],
                                torch.from_numpy(np.asarray(cam["w2c"], np.float32)).to(dev),
                                torch.from_numpy(np.asarray(cam["cam_position"], np.float32)).to(dev),
                                cam["width"], cam["height"], cam["focal_x"], cam["focal_y"], cam["center_x"],
                                cam["center_y"], cam["near_plane"], cam["far_plane"], False, 1, False, False, False,
                                frame_idx=7)
        won = fb[(fb >> 32) != 0xFFFFFFFF] & 0xFFFFFFFF
        u = int(torch.unique(won).numel())
        P = Q = cam["width"] * cam["height"]
        spp = int(cam["spp"])
        per = 12.0 * n + 32.0 * n_z + 16.0 * P + 204.0 * u
        bytes_total += spp * per + 24.0 * Q * (1 + 2 * (spp - 1))
        atomics += spp * work[v][1]
    t = step_ms * 1e-3
    t_hbm = bytes_total / (HBM_PEAK_GBS * 1e9)
    t_atomic = atomics / atomics_per_s
    return {
        "definition": "per (view, sample): B = 12N + 32N_z + 16P + 204U + 24Q(1 + [mode != RESET]) bytes, "
                      "A = accepted fragments (atomicMin each in a scatter); t_roof = max(B / 8 TB/s, A / R_atomic)",
        "bytes_per_step": round(bytes_total),
        "atomics_per_step": round(atomics),
        "r_atomic_per_s": round(atomics_per_s, 1),
        "t_hbm_ms": round(t_hbm * 1e3, 3),
        "t_atomic_ms": round(t_atomic * 1e3, 3),
        "step_ms": round(step_ms, 3),
        "achieved_gbs": round(bytes_total / t / 1e9, 1),
        "hbm_frac": round(t_hbm / t, 4),
        "atomic_frac": round(t_atomic / t, 4),
        "frac": round(max(t_hbm, t_atomic) / t, 4),
    }


# HBM bytes per launch of the tile rasteriser on the default workload (6M Gaussians, 64 views,
# 8 samples per LDS pass), from rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (calibrated on a
# 1 GiB device copy); the commands and raw counters are in profiles/tile_raster_c4_6M_64views.md.
PROFILED_RASTER_TRAFFIC = {"bytes": 6.58e9, "profile": "profiles/tile_raster_c4_6M_64views.md"}


def roofline_tiles(views, raster_ms: list, n_instances: int, visible_pairs: int, default_workload: bool):
    """Algorithmic bytes of the tile-rasteriser stage (its one-layer launch for the views of one
    sample and its 8-layer launch for the others, timed together): 44 B of Gaussian parameters
    per visible (view, Gaussian) pair (means 12 B + the cached covariance and opacity 32 B, read
    at least once), 8 B per (view, tile, Gaussian) list entry (index + depth bucket) and 8 B
    per framebuffer word written (every sample of every pixel, once)."""
    words = sum(cam["width"] * cam["height"] * cam["spp"] for cam in views)
    b = 44.0 * visible_pairs + 8.0 * n_instances + 8.0 * words
    t = sum(raster_ms) / len(raster_ms) * 1e-3
    achieved = b / t / 1e9
    return {
        "kernel": "sgs::batch_tile_raster_kernel<1> + <8> (one stage)",
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": PROFILED_RASTER_TRAFFIC["bytes"] if default_workload else None,
        "traffic_source": PROFILED_RASTER_TRAFFIC["profile"] if default_workload else None,
        "issue_bound": "VALU (1.0 wave-instruction per CU-cycle, SQ counters in the same profile)",
        "avg_launch_ms": round(t * 1e3, 3),
        "algorithmic_bytes_per_launch": int(b),
        "instances_per_launch": int(n_instances),
        "visible_pairs_per_launch": int(visible_pairs),
        "framebuffer_words_per_launch": int(words),
    }


def roofline_batch(scene_n: int, views, acc_ms: list, work):
    """Algorithmic bytes of one persistent accumulation launch: per view 44 B of parameters
    for every Gaussian (12 B means for the depth cull + 32 B scales/rotation/opacity) and 8 B
    per atomicMin issued (sample s of view v: the per-sample count of the counting pass
    times spp)."""
    atomics = sum(work[v][1] * cam["spp"] for v, cam in enumerate(views))
    tested = sum(work[v][0] for v in range(len(views)))
    b = 44.0 * scene_n * len(views) + 8.0 * atomics
    t = sum(acc_ms) / len(acc_ms) * 1e-3
    achieved = b / t / 1e9
    return {
        "kernel": "sgs::batch_accumulate_kernel",
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "avg_launch_ms": round(t * 1e3, 3),
        "algorithmic_bytes_per_launch": int(b),
        "atomics_per_s": round(atomics / t, 1),
        "fragment_alphas_per_s": round(tested / t, 1),
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


# ---------------------------------------------------------------------------------------------
# serving: 32 streams on smooth trajectories, Poisson arrivals, continuous batching
# ---------------------------------------------------------------------------------------------
SERVE_SIZES = [(1280, 720), (1920, 1080), (2560, 1440)]
SERVE_SPP = [1, 2, 8]


def serving_streams(n_streams: int, seed: int = 3):
    import numpy as np
    from StochasticGSCudaBackend.serve import StreamSpec
    from StochasticGSCudaBackend.synthetic import camera, orbit_eye
    rng = np.random.default_rng(seed)
    specs, paths = [], []
    for _ in range(n_streams):
        w, h = SERVE_SIZES[int(rng.integers(len(SERVE_SIZES)))]
        c = camera(orbit_eye(0.0, 3.0), w, h)
        specs.append(StreamSpec(w, h, c["focal_x"], c["focal_y"], c["center_x"], c["center_y"], c["near_plane"],
                                c["far_plane"]))
        # orbit phase, angular speed (rad/s), radius, height amplitude
        paths.append((rng.uniform(0, 2 * math.pi), rng.uniform(0.2, 0.6) * rng.choice([-1, 1]),
                      rng.uniform(2.6, 3.4), rng.uniform(-0.3, 0.3)))
    return specs, paths


def serving_requests(specs, paths, times, seed: int):
    """One request per arrival time: a uniformly chosen stream, its pose at that time, spp
    from {1, 2, 8}."""
    import numpy as np
    from StochasticGSCudaBackend.serve import Request
    from StochasticGSCudaBackend.synthetic import camera, orbit_eye
    rng = np.random.default_rng(seed)
    reqs = []
    for t in times:
        s = int(rng.integers(len(specs)))
        phase, omega, radius, height = paths[s]
        eye = orbit_eye(phase + omega * t, radius, height * math.sin(0.5 * t))
        c = camera(eye, specs[s].width, specs[s].height)
        reqs.append(Request(stream=s, w2c=c["w2c"], cam_position=c["cam_position"],
                            spp=int(SERVE_SPP[int(rng.integers(len(SERVE_SPP)))]), arrival=float(t)))
    return reqs


def bench_serving(args, scene, dev):
    """Saturated throughput (every request queued at t=0), then Poisson arrivals at 80 % of it:
    p50/p99 latency (arrival to completion seen by the host). The plain per-request path is
    timed on the same requests (spp single-view render() calls each) and its latency at the
    same arrivals follows from a FIFO single-server queue over those service times."""
    import numpy as np
    import torch
    from StochasticGSCudaBackend import _native
    from StochasticGSCudaBackend.serve import Server, latency_stats
    specs, paths = serving_streams(args.serve_streams)
    server = Server(scene["means"], scene["scales"], scene["rotations"], scene["opacities"], scene["sh"], specs,
                    max_views=args.serve_streams, inflight=args.serve_inflight,
                    max_wait=args.serve_window_ms * 1e-3)
    # warm-up and saturation: 4 frames per stream queued at once
    warm = serving_requests(specs, paths, [0.0] * args.serve_streams, seed=10)
    server.run(warm)
    sat = serving_requests(specs, paths, [0.0] * (4 * args.serve_streams), seed=11)
    t0 = time.perf_counter()
    server.run(sat)
    sat_s = time.perf_counter() - t0
    sat_rate = len(sat) / sat_s
    log(f"serving saturated: {sat_rate:.1f} req/s ({server.batches} batches so far)")
    # 80 % load, Poisson arrivals
    rate = 0.8 * sat_rate
    rng = np.random.default_rng(12)
    n = max(1, int(rate * args.serve_seconds))
    times = np.cumsum(rng.exponential(1.0 / rate, n))
    load = serving_requests(specs, paths, times, seed=13)
    b0, ov0 = server.batches, server.overflowed_batches
    t0 = time.perf_counter()
    done = server.run(load)
    run_s = time.perf_counter() - t0
    overflowed = server.overflowed_batches - ov0
    stats = latency_stats(done)
    samples = sum(specs[r.stream].width * specs[r.stream].height * r.spp for r in done)
    batches = server.batches - b0
    log(f"serving at {rate:.1f} req/s: p50 {stats['p50_ms']} ms, p99 {stats['p99_ms']} ms, "
        f"{len(done) / batches:.1f} requests per batch")
    # plain per-request path: service time of each request rendered alone
    sample_reqs = load[:min(len(load), 96)]
    hist_cache = {}
    service = []
    kb = (torch.empty((0, 0, 2), dtype=torch.int32, device=dev), torch.empty((0, 0, 4), device=dev))
    eye4 = torch.eye(4, device=dev)
    noise = torch.empty(0, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    for r in sample_reqs:
        sp = specs[r.stream]
        if (sp.width, sp.height) not in hist_cache:
            hist_cache[(sp.width, sp.height)] = tuple(torch.zeros((sp.height, sp.width, c), device=dev)
                                                      for c in (3, 1, 1, 1))
        hist = hist_cache[(sp.width, sp.height)]
        w2c = torch.from_numpy(np.asarray(r.w2c, np.float32)).to(dev)
        cp = torch.from_numpy(np.asarray(r.cam_position, np.float32)).to(dev)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        h = hist
        for k in range(r.spp):
            h = _native.render(scene["means"], scene["scales"], scene["rotations"], scene["opacities"], scene["sh"],
                               w2c, cp, sp.width, sp.height, sp.focal_x, sp.focal_y, sp.center_x, sp.center_y,
                               sp.near_plane, sp.far_plane, False, 1, False, False, False, noise, kb[0], kb[1], *h,
                               eye4, 2 if k == 0 else 1, frame_idx=r.frame_idx + k)
        torch.cuda.synchronize()
        service.append(time.perf_counter() - ts)
    mean_service = float(np.mean(service))
    svc = np.resize(np.array(service), len(load))
    arr = np.array([r.arrival for r in load]) - load[0].arrival
    done_t, lat = 0.0, []
    for a, sv in zip(arr, svc):
        done_t = max(done_t, a) + sv
        lat.append(done_t - a)
    lat = np.array(lat) * 1e3
    del server
    torch.cuda.empty_cache()
    return {"streams": args.serve_streams, "inflight": args.serve_inflight,
            "window_ms": args.serve_window_ms, "saturated_req_per_s": round(sat_rate, 1),
            "offered_req_per_s": round(rate, 1), "completed_req_per_s": round(len(done) / run_s, 1),
            "msamples_per_s": round(samples / run_s / 1e6, 1), "requests": len(done),
            "requests_per_batch": round(len(done) / max(batches, 1), 2),
            "p50_ms": stats["p50_ms"], "p99_ms": stats["p99_ms"],
            "p99_over_p50": round(stats["p99_ms"] / max(stats["p50_ms"], 1e-9), 3),
            "overflowed_batches": overflowed,
            "plain": {"req_per_s_capacity": round(1.0 / mean_service, 1),
                      "p50_ms_at_offered_load": round(float(np.percentile(lat, 50)), 3),
                      "p99_ms_at_offered_load": round(float(np.percentile(lat, 99)), 3),
                      "service_ms_mean": round(mean_service * 1e3, 3)}}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this renderer runs on one GPU; --gpus must be 1 (views and requests would shard "
                                   "across GPUs with the scene replicated and no exchange; not built)"}))
        return 2
    import numpy as np
    import torch
    from StochasticGSCudaBackend import create_k_buffer
    from StochasticGSCudaBackend.synthetic import make_scene, morton_order, reorder, scaled, scene_c4, views_c4

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t0 = time.perf_counter()
    spec = scaled(scene_c4(), args.gaussians)
    scene_np = make_scene(spec)
    if not args.no_morton:
        # scenes are stored Morton-ordered, as the reference's model loader does (Model.py:32-34)
        scene_np = reorder(scene_np, morton_order(scene_np["means"]))
    scene = {k: torch.from_numpy(v).to(dev) for k, v in scene_np.items()}
    views = views_c4(args.views)
    settings = [view_settings(v, dev) for v in views]
    k_buffers = [create_k_buffer(s) for s in settings]
    torch.cuda.synchronize()
    log(f"scene {spec.n} Gaussians, {len(views)} views, {samples_per_step(views) / 1e6:.1f} M pixel-samples/step "
        f"(setup {time.perf_counter() - t0:.1f} s)")

    if args.copy_probe:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    S = samples_per_step(views)
    result = {}
    plain_out0 = None
    if args.only in ("plain", "all"):
        result["plain"], plain_out0 = bench_plain(args, scene, views, settings, k_buffers, S)
    if args.only in ("batched", "all"):
        result["batched"], acc_ms = bench_batched(args, scene, views, S, plain_out0, dev)
    serving = None
    if args.only == "all" and args.serve_seconds > 0:
        serving = bench_serving(args, scene, dev)
        if "plain" in result:
            result["batched"]["speedup_vs_plain"] = round(result["batched"]["value"] / result["plain"]["value"], 3)

    # ---- roofline of the dominant kernel (and of the reference's scatter schedule) -------------
    ms, work = time_scatter_launches(scene, views)
    roof_plain = roofline_scatter(spec.n, views, ms, work)
    log(f"scatter launches: avg {roof_plain['avg_launch_ms']} ms, {roof_plain['achieved']} GB/s algorithmic")
    roof = roof_plain
    if "batched" in result:
        if args.schedule == "tiles":
            default_workload = (spec.n == 6_000_000 and len(views) == 64 and args.tile_samples == 8)
            roof = roofline_tiles(views, acc_ms, result["batched"]["tile_instances"], result["batched"]["visible_pairs"],
                                  default_workload)
            log(f"tile raster: {roof['avg_launch_ms']} ms, {roof['achieved']} GB/s algorithmic")
        else:
            roof = roofline_batch(spec.n, views, acc_ms, work)
            log(f"batched accumulate: {roof['avg_launch_ms']} ms, {roof['achieved']} GB/s algorithmic")
    roof_samples = None
    if "batched" in result:
        roof_samples = roofline_render_samples(scene_np, views, scene, result["batched"]["ms_per_step"], work,
                                               roof_plain["atomics_per_s"])
        log(f"render-sample roofline: HBM {roof_samples['hbm_frac']}, atomics {roof_samples['atomic_frac']} "
            f"of the {roof_samples['step_ms']} ms step")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene_np, views, args.cpu_seconds)
        log(f"cpu oracle: {cpu['value']} Msamples/s on {cpu['cores']} threads")

    head = result.get("batched", result.get("plain"))
    line = {
        "metric": "pixel-samples/s (Msamples/s) batched vs plain, 6M Gaussians; p50/p99 frame ms",
        "value": head["value"],
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{spec.n} Gaussians (8-component mixture, SH3, Morton order), {len(views)} views "
                               f"on a ring, resolutions 256x256..2048x1152, spp in {{1,4,16}}, ST scatter, ss=1",
                   "path": (f"batched render of all views and samples, {args.schedule} schedule" if "batched" in result
                            else "plain per-view render() calls"),
                   "pixel_samples_per_step": S},
        "batched": result.get("batched"),
        "plain": result.get("plain"),
        "p50_frame_ms": result["plain"]["p50_frame_ms"] if "plain" in result else None,
        "p99_frame_ms": result["plain"]["p99_frame_ms"] if "plain" in result else None,
        "roofline": roof,
        "roofline_scatter": roof_plain,
        "roofline_render_samples": roof_samples,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_plain(args, scene, views, settings, k_buffers, S):
    """Per-view sequential render() calls of the drop-in API: throughput and frame latency."""
    import numpy as np
    import torch
    for w in range(args.warmup):
        run_plain(scene, views, settings, k_buffers, frame_base=1_000_000 * (w + 1))
        torch.cuda.synchronize()
        log(f"plain warmup {w + 1}/{args.warmup} done")
    out0 = None
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        out = run_plain(scene, views, settings, k_buffers, frame_base=10_000 * k)
        torch.cuda.synchronize()
        if k == 0:
            out0 = out
        log(f"plain step {k + 1}/{args.steps}: {(time.perf_counter() - ts) * 1e3:.1f} ms")
    plain_s = time.perf_counter() - t_all
    # per-frame latency: one view with all its samples, synchronised (outside the timed loop)
    frame_ms = []
    for v in range(len(views)):
        ts = time.perf_counter()
        run_plain(scene, views[v:v + 1], settings[v:v + 1], k_buffers[v:v + 1], frame_base=500_000 + 100 * v)
        torch.cuda.synchronize()
        frame_ms.append((time.perf_counter() - ts) * 1e3)
    return {"value": round(S * args.steps / plain_s / 1e6, 2), "ms_per_step": round(plain_s * 1e3 / args.steps, 2),
            "p50_frame_ms": round(float(np.percentile(frame_ms, 50)), 3),
            "p99_frame_ms": round(float(np.percentile(frame_ms, 99)), 3)}, out0


def bench_batched(args, scene, views, S, plain_out0, dev):
    """Every view and sample in one batched render (or, when the autotuner picks it, in view
    groups rendered by concurrent batches on their own streams); stage timing with HIP events
    on the library's stream of a single batch; graph replay; bit comparison with the plain
    outputs of the same frames."""
    import numpy as np
    import torch
    from StochasticGSCudaBackend.batch import (SCHEDULE_SCATTER, SCHEDULE_TILES, BatchRenderer, BatchTuning, BatchView,
                                               StreamedBatchRenderer)
    bviews = [BatchView(host_settings(cam), spp=cam["spp"], frame_idx=0) for cam in views]
    tuning = BatchTuning(schedule=SCHEDULE_TILES if args.schedule == "tiles" else SCHEDULE_SCATTER,
                         tile_samples=args.tile_samples, view_major=int(args.view_major))
    tuned = None
    streams, part_tunings = 1, None
    if not args.no_autotune:
        from StochasticGSCudaBackend.autotune import tune_batch
        ts = time.perf_counter()
        res = tune_batch(scene["means"], scene["scales"], scene["rotations"], scene["opacities"], scene["sh"], bviews,
                         repeats=2)
        tuning = res.best._replace(view_major=int(args.view_major))
        bviews = res.views
        streams, part_tunings = res.streams, res.part_tunings
        args.schedule = "tiles" if tuning.schedule in (0, SCHEDULE_TILES) else "scatter"
        args.tile_samples = tuning.tile_samples or 8
        tuned = {"best": dict(tuning._asdict()), "streams": streams,
                 "part_max_tile_instances": [int(t.max_tile_instances) for t in part_tunings] if part_tunings else None,
                 "view_bin_sizes": sorted({v.bin_size for v in bviews}),
                 "seconds": round(time.perf_counter() - ts, 1),
                 "trials": [[dict(c), round(ms, 3)] for c, ms in res.trials]}
        log(f"autotuned batch: {tuned['best']} from {tuned['trials']}")
    br = BatchRenderer(scene["means"], scene["scales"], scene["rotations"], scene["opacities"], scene["sh"], bviews,
                       tuning)
    log(f"batched workspace {br.workspace.numel() / 2**30:.2f} GiB")

    def diff_vs_plain(r):
        if plain_out0 is None:
            return None
        r.set_frames(frames_for_step(views, 0))
        r.upload_table()
        r.render()
        torch.cuda.synchronize()
        d = 0.0
        for a, b in zip(r.outputs, plain_out0):
            for x, y in zip(a, b):
                d = max(d, float((x - y).abs().max()))
        return d

    max_diff = diff_vs_plain(br)
    log(f"batched vs plain (same frames): max abs difference {max_diff}")
    stream = torch.cuda.current_stream(dev)
    for w in range(args.warmup):
        br.set_frames(frames_for_step(views, 2_000_000 + 10_000 * w))
        br.upload_table()
        br.render()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
    for e4 in ev:
        for e in e4:
            e.record(stream)
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        br.set_frames(frames_for_step(views, 10_000 * k))
        br.upload_table()
        br.render(stage_events=ev[k])
        torch.cuda.synchronize()
        log(f"batched step {k + 1}/{args.steps}: {(time.perf_counter() - ts) * 1e3:.1f} ms")
    batch_s = time.perf_counter() - t_all
    main_ms = [e5[2].elapsed_time(e5[3]) for e5 in ev]
    stage = {"clear_ms": round(float(np.mean([e5[0].elapsed_time(e5[1]) for e5 in ev])), 3),
             "binning_ms": round(float(np.mean([e5[1].elapsed_time(e5[2]) for e5 in ev])), 3),
             "accumulate_kernel_ms": round(float(np.mean(main_ms)), 3),
             "accumulate_total_ms": round(float(np.mean([e5[1].elapsed_time(e5[3]) for e5 in ev])), 3),
             "resolve_ms": round(float(np.mean([e5[3].elapsed_time(e5[4]) for e5 in ev])), 3)}
    n_inst, n_visible, overflowed = br.tile_stats()
    log(f"stages {stage}; tile instances {n_inst} (overflow {overflowed})")
    single = {"value": round(S * args.steps / batch_s / 1e6, 2), "ms_per_step": round(batch_s * 1e3 / args.steps, 2)}
    head = br
    if streams > 1:
        # the headline steps: the view groups as concurrent batches on their own streams
        del br
        torch.cuda.empty_cache()
        head = StreamedBatchRenderer(scene["means"], scene["scales"], scene["rotations"], scene["opacities"],
                                     scene["sh"], bviews, part_tunings or tuning, streams)
        log(f"{streams} view groups {[len(g) for g in head.groups]}, workspace {head.workspace_bytes / 2**30:.2f} GiB")
        max_diff = diff_vs_plain(head)
        log(f"streamed batches vs plain (same frames): max abs difference {max_diff}")
        for w in range(args.warmup):
            head.set_frames(frames_for_step(views, 2_000_000 + 10_000 * w))
            head.upload_table()
            head.render()
        torch.cuda.synchronize()
        t_all = time.perf_counter()
        for k in range(args.steps):
            ts = time.perf_counter()
            head.set_frames(frames_for_step(views, 10_000 * k))
            head.upload_table()
            head.render()
            torch.cuda.synchronize()
            log(f"streamed batched step {k + 1}/{args.steps}: {(time.perf_counter() - ts) * 1e3:.1f} ms")
        batch_s = time.perf_counter() - t_all
    graph = head.capture()
    torch.cuda.synchronize()
    g_ms = []
    for k in range(args.steps):
        ts = time.perf_counter()
        head.set_frames(frames_for_step(views, 3_000_000 + 10_000 * k))
        head.upload_table()
        graph.replay()
        torch.cuda.synchronize()
        g_ms.append((time.perf_counter() - ts) * 1e3)
    return {"value": round(S * args.steps / batch_s / 1e6, 2), "ms_per_step": round(batch_s * 1e3 / args.steps, 2),
            "streams": streams, "single_batch": single,
            "graph_ms_per_step": round(float(np.mean(g_ms)), 2), "stages": stage,
            "tile_instances": n_inst, "visible_pairs": n_visible, "tile_overflow": overflowed, "autotune": tuned,
            "max_abs_diff_vs_plain": max_diff}, main_ms


if __name__ == "__main__":
    sys.exit(main())
