#!/usr/bin/env python3
"""Benchmark: pixel-samples/s of the stochastic splatting renderer on the 6M-Gaussian,
64-view workload (uneven resolutions from 256x256 to 2048x1152, spp in {1, 4, 16}).

One step renders every view with all of its samples (RESET then STATIC accumulation, like
the reference's SPP loop in Renderer.py:85-87), inputs resident in HBM before timing.
Prints one JSON line. Beside the headline (batched renderer) it reports the plain per-view
path, the serving loop (one and two batches in flight), the other workloads (1M Gaussians
at 1080p with 16 samples: scatter, hybrid and tile schedules; 3M heavy-tailed Gaussians
with point splatting at 2x supersampling), the device's measured copy and random-atomicMin
ceilings, the roofline of the dominant kernel and of the whole step with HBM bytes and
VALU instruction counts read from the committed rocprofv3 summary
(profiles/c4_batched_kernels.json), and the CPU oracle timed on this host.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--views 64] [--gaussians 6000000]

Profiling (the commands that make profiles/c4_batched_kernels.*):

    python bench.py --profile                       # only batched renders of the workload
    python bench.py --summarize-profile TRACE FETCH WRITE SQ   # rocprofv3 output dirs -> profiles/
    python bench.py --summarize-trace DIR OUT.md    # per-(kernel, grid) table of any kernel trace
"""
from __future__ import annotations

import argparse
import csv
import hashlib
import json
import math
import os
import statistics
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
for _p in (str(REPO), str(REPO / "stochastic-gaussian-splatting_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
N_CU = 256                 # MI355X compute units (8 XCDs x 32)
N_XCD = 8
PEAK_CLOCK_GHZ = 2.4       # MI355X peak engine clock
# VALU issue: each of a CU's four 16-lane SIMDs takes a wave64 VALU instruction every four
# cycles (a quad-cycle): 256 CUs x 4 SIMDs x 2.4 GHz / 4 = 614.4 G SIMD quad-cycles/s. gfx950
# issues a second VALU instruction in some quad-cycles: SQ_ACTIVE_INST_VALU counts the
# quad-cycles of every wave's VALU instructions (both of a pair) and SQ_ACTIVE_INST_VALU2 the
# quad-cycles with two, so ACTIVE - ACTIVE2 quad-cycles had at least one VALU instruction (the
# busy fraction, at most one) while instructions per quad-cycle can exceed one
VALU_PEAK_G_PER_S = N_CU * PEAK_CLOCK_GHZ
PROFILE_JSON = REPO / "profiles" / "c4_batched_kernels.json"
PROFILE_MD = REPO / "profiles" / "c4_batched_kernels.md"
COPY_CALIBRATION_BYTES = 1 << 30
SOURCE_DIRS = [REPO / "stochastic-gaussian-splatting_amd" / "csrc", REPO / "include"]

# kernels of one batched render with the tile schedule, in launch order
RASTER_KERNELS = ["sgs::batch_tile_raster_kernel<1u, 16u>", "sgs::batch_tile_raster_kernel<8u, 16u>",
                  "sgs::batch_tile_raster_kernel<16u, 8u>"]
BATCH_KERNELS = ["sgs::chunk_centroid_kernel", "sgs::view_key_range_kernel", "sgs::cov3d_kernel",
                 "sgs::batch_bin_kernel<false>", "sgs::exclusive_scan_kernel<sgs::CountInput>",
                 "sgs::batch_bin_kernel<true>", "sgs::tile_sort_kernel", *RASTER_KERNELS,
                 "sgs::batch_resolve_kernel<16u>"]


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--gaussians", type=int, default=6_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="budget of the CPU oracle on the headline workload (OpenMP, all cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the 1M and 3M Gaussian workloads")
    ap.add_argument("--serve-seconds", type=float, default=6.0,
                    help="length of each serving run at 80%% of saturated load (0: skip serving)")
    ap.add_argument("--serve-streams", type=int, default=32)
    ap.add_argument("--serve-inflight", type=str, default="1,2",
                    help="batches on the GPU at once, one serving run per value (the first is the headline)")
    ap.add_argument("--serve-window-ms", type=float, default=35.0,
                    help="batching window: a free slot waits until the oldest request is this old")
    ap.add_argument("--serve-temporal-only", action="store_true",
                    help="streams resolve with REPROJECT alone instead of the spatio-temporal resolve (spatial "
                         "reuse + REPROJECT)")
    ap.add_argument("--only", choices=["plain", "batched", "all"], default="all",
                    help="run only one path; the JSON line then reports that path")
    ap.add_argument("--schedule", choices=["tiles", "scatter"], default="tiles",
                    help="batched accumulation: LDS tiles (default) or persistent scatter")
    ap.add_argument("--tile-samples", type=int, choices=[0, 8, 16], default=0,
                    help="samples per LDS pass of the tile rasteriser (with --no-autotune; 0: 8 for views of up "
                         "to 8 samples, 16 per half tile above)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --schedule/--tile-samples instead of tuning them on the workload first")
    ap.add_argument("--view-major", action="store_true",
                    help="batched work items view by view instead of Gaussian chunk by chunk")
    ap.add_argument("--no-morton", action="store_true",
                    help="keep generation order instead of the Morton order scenes are stored in")
    ap.add_argument("--profile", action="store_true",
                    help="profiling run: a 1 GiB calibration copy, then --profile-renders single-batch renders of "
                         "the workload with the default tuning and nothing else")
    ap.add_argument("--profile-renders", type=int, default=3)
    ap.add_argument("--summarize-profile", nargs=4, metavar=("TRACE", "FETCH", "WRITE", "SQ"),
                    help="rocprofv3 output directories of the --profile runs (kernel trace; --pmc FETCH_SIZE; "
                         "--pmc WRITE_SIZE; --pmc SQ counters) -> profiles/c4_batched_kernels.{json,md}")
    ap.add_argument("--summarize-trace", nargs=2, metavar=("DIR", "OUT"),
                    help="rocprofv3 kernel-trace directory -> markdown table per (kernel, grid)")
    ap.add_argument("--trace-title", help="title of the --summarize-trace table")
    ap.add_argument("--trace-command", help="the traced command, quoted above the --summarize-trace table")
    return ap.parse_args()


# ---------------------------------------------------------------------------------------------
# rocprofv3 summaries
# ---------------------------------------------------------------------------------------------
def sources_sha256() -> str:
    """Digest of the HIP sources and headers: a profile summary is current for the build it
    was taken from."""
    h = hashlib.sha256()
    for d in SOURCE_DIRS:
        for p in sorted(d.glob("*")):
            if p.suffix in (".hip", ".h"):
                h.update(p.name.encode())
                h.update(p.read_bytes())
    return h.hexdigest()


def short_kernel_name(name: str) -> str:
    """'void ns::kernel<8u>(args...)' -> 'ns::kernel<8u>'."""
    n = name.strip()
    if n.startswith("void "):
        n = n[5:]
    prefix = ""
    if n.startswith("(anonymous namespace)::"):
        prefix, n = "(anonymous namespace)::", n[len("(anonymous namespace)::"):]
    depth, cut = 0, len(n)
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    return prefix + n[:cut]


def _csv_rows(directory: Path, suffix: str) -> list[dict]:
    rows = []
    for p in sorted(Path(directory).rglob(f"*{suffix}")):
        with open(p, newline="") as f:
            rows.extend(csv.DictReader(f))
    if not rows:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    return rows


def trace_groups(trace_dir: Path) -> dict:
    """(kernel, grid size in lanes) -> durations in ns, from run_kernel_trace.csv."""
    groups: dict = {}
    for r in _csv_rows(trace_dir, "kernel_trace.csv"):
        grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        key = (short_kernel_name(r["Kernel_Name"]), grid)
        groups.setdefault(key, []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return groups


def counter_groups(pmc_dir: Path) -> dict:
    """(kernel, grid) -> counter name -> per-dispatch values, from run_counter_collection.csv."""
    groups: dict = {}
    for r in _csv_rows(pmc_dir, "counter_collection.csv"):
        key = (short_kernel_name(r["Kernel_Name"]), int(r["Grid_Size"]))
        groups.setdefault(key, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    return groups


def summarize_trace(trace_dir: str, out_md: str, title: str | None = None, command: str | None = None) -> None:
    """Markdown table of a kernel trace: every (kernel, grid) group with its dispatch count
    and average / median duration, longest total first (under a title and the command that
    made the trace, when given)."""
    groups = trace_groups(Path(trace_dir))
    rows = sorted(groups.items(), key=lambda kv: -sum(kv[1]))
    lines = []
    if title:
        lines += [f"# {title}", ""]
    if command:
        lines += ["Command (repository root):", "", f"    {command}", "",
                  f"    python3 bench.py --summarize-trace <trace dir> {out_md}", ""]
    lines += ["| kernel | grid (lanes) | dispatches | average ms | median ms | total ms |", "|---|---|---|---|---|---|"]
    for (name, grid), d in rows[:40]:
        lines.append(f"| `{name}` | {grid:,} | {len(d)} | {statistics.mean(d) / 1e6:.4f} | "
                     f"{statistics.median(d) / 1e6:.4f} | {sum(d) / 1e6:.2f} |")
    Path(out_md).write_text("\n".join(lines) + "\n")
    print("\n".join(lines))


def summarize_profile(trace_dir: str, fetch_dir: str, write_dir: str, sq_dir: str) -> None:
    """profiles/c4_batched_kernels.json (read by the benchmark) and .md from the rocprofv3
    runs of `bench.py --profile`: per kernel of the batched render its dispatches per
    render, average duration, and the median per-dispatch counters; HBM bytes are
    FETCH_SIZE and WRITE_SIZE scaled by what they read for the calibration copy of 1 GiB."""
    trace = trace_groups(Path(trace_dir))
    pmc = {}
    for d in (fetch_dir, write_dir, sq_dir):
        for key, counters in counter_groups(Path(d)).items():
            pmc.setdefault(key, {}).update(counters)
    copy_key = max((k for k in pmc if k[0].endswith("copy_kernel")), key=lambda k: k[1])
    fetch_kb = statistics.median(pmc[copy_key]["FETCH_SIZE"])
    write_kb = statistics.median(pmc[copy_key]["WRITE_SIZE"])
    cal = {"copy_bytes": COPY_CALIBRATION_BYTES, "copy_fetch_kb": fetch_kb, "copy_write_kb": write_kb,
           "fetch_scale": COPY_CALIBRATION_BYTES / (fetch_kb * 1024.0),
           "write_scale": COPY_CALIBRATION_BYTES / (write_kb * 1024.0)}
    renders = None
    kernels = {}
    # every renderer kernel of the profiled run (the --profile run launches nothing else of
    # the library), in launch order where known
    names = sorted({k[0] for k in trace if k[0].startswith("sgs::")},
                   key=lambda n: (BATCH_KERNELS.index(n) if n in BATCH_KERNELS else len(BATCH_KERNELS), n))
    for name in names:
        keys = [k for k in trace if k[0] == name]
        key = max(keys, key=lambda k: (len(trace[k]), k[1]))   # the renders' grid
        durations = trace[key]
        c = pmc.get(key, {})
        med = {n: statistics.median(v) for n, v in c.items()}
        entry = {"grid_lanes": key[1], "dispatches": len(durations), "avg_ms": statistics.mean(durations) / 1e6,
                 "median_ms": statistics.median(durations) / 1e6, "counters": med}
        if "FETCH_SIZE" in med and "WRITE_SIZE" in med:
            entry["fetch_bytes"] = med["FETCH_SIZE"] * 1024.0 * cal["fetch_scale"]
            entry["write_bytes"] = med["WRITE_SIZE"] * 1024.0 * cal["write_scale"]
            entry["hbm_bytes"] = entry["fetch_bytes"] + entry["write_bytes"]
        if "GRBM_GUI_ACTIVE" in med and "SQ_INSTS_VALU" in med:
            cycles = med["GRBM_GUI_ACTIVE"] / N_XCD
            entry["valu_per_cu_cycle"] = med["SQ_INSTS_VALU"] / (N_CU * cycles)
            entry["clock_ghz"] = cycles / (entry["median_ms"] * 1e6)
            if "SQ_ACTIVE_INST_VALU" in med and "SQ_ACTIVE_INST_VALU2" in med:
                # SIMD quad-cycles with a VALU issue over the N_CU x 4 SIMDs x cycles / 4 available
                entry["valu_busy_quad_cycles"] = med["SQ_ACTIVE_INST_VALU"] - med["SQ_ACTIVE_INST_VALU2"]
                entry["valu_busy"] = entry["valu_busy_quad_cycles"] / (N_CU * cycles)
        kernels[name] = entry
        if name == "sgs::cov3d_kernel":
            renders = len(durations)
    out = {"workload": "6M Gaussians (8-component mixture, SH3, Morton order), 64 views on a ring, 256x256.."
                       "2048x1152, spp in {1,4,16}; one batch, tile schedule, default tuning",
           "sources_sha256": sources_sha256(), "renders": renders, "calibration": cal, "kernels": kernels,
           # the commands that wrote the four directories given (their paths as given)
           "commands": {
               "trace": f"rocprofv3 --kernel-trace --stats --output-format csv -d {trace_dir} -o run -- "
                        "python3 bench.py --profile",
               "fetch": f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d {fetch_dir} -o run -- "
                        "python3 bench.py --profile",
               "write": f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d {write_dir} -o run -- "
                        "python3 bench.py --profile",
               "sq": "rocprofv3 --pmc SQ_INSTS_VALU SQ_WAVES SQ_BUSY_CYCLES SQ_INSTS_LDS SQ_INSTS_SALU "
                     "SQ_LDS_BANK_CONFLICT SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_VALU2 GRBM_GUI_ACTIVE "
                     f"--output-format csv -d {sq_dir} -o run -- python3 bench.py --profile",
               "summary": f"python3 bench.py --summarize-profile {trace_dir} {fetch_dir} {write_dir} {sq_dir}"}}
    PROFILE_JSON.parent.mkdir(exist_ok=True)
    PROFILE_JSON.write_text(json.dumps(out, indent=1) + "\n")
    lines = [
        "# Kernel profile: one batched render of 6M Gaussians x 64 views (tile schedule)", "",
        f"Workload: {out['workload']}. MI355X (gfx950). Source digest `{out['sources_sha256'][:16]}`. "
        f"{renders} renders profiled.", "", "Commands (repository root):", "",
        *[f"    {v}" for v in out["commands"].values()], "",
        f"Calibration: the 1 GiB copy reads FETCH_SIZE {fetch_kb:,.0f} KB and WRITE_SIZE {write_kb:,.0f} KB: "
        f"FETCH_SIZE x{cal['fetch_scale']:.3f}, WRITE_SIZE x{cal['write_scale']:.3f}.", "",
        "| kernel | per render | avg ms | fetched GB | written GB | measured GB/s | of 8 TB/s | VALU wave-instr |"
        " VALU instr per SIMD quad-cycle | VALU busy | clock GHz |",
        "|---|---|---|---|---|---|---|---|---|---|---|"]
    total_ms = total_b = 0.0
    for name, e in kernels.items():
        per = e["dispatches"] / max(renders or 1, 1)
        b = e.get("hbm_bytes")
        gbs = b / (e["avg_ms"] * 1e-3) / 1e9 if b is not None else None
        total_ms += e["avg_ms"] * per
        total_b += (b or 0.0) * per
        lines.append(
            f"| `{name}` | {per:g} | {e['avg_ms']:.3f} | "
            f"{(e.get('fetch_bytes') or 0) / 1e9:.3f} | {(e.get('write_bytes') or 0) / 1e9:.3f} | "
            f"{gbs:.0f} | {gbs / HBM_PEAK_GBS:.3f} | {e['counters'].get('SQ_INSTS_VALU', float('nan')):.3e} | "
            f"{e.get('valu_per_cu_cycle', float('nan')):.3f} | {e.get('valu_busy', float('nan')):.3f} | "
            f"{e.get('clock_ghz', float('nan')):.2f} |"
            if gbs is not None else f"| `{name}` | {per:g} | {e['avg_ms']:.3f} | | | | | | | | |")
    lines += ["", f"One render: {total_ms:.2f} ms of kernels, {total_b / 1e9:.2f} GB of HBM traffic "
                  f"({total_b / (total_ms * 1e-3) / 1e9:.0f} GB/s over the kernel time)."]
    PROFILE_MD.write_text("\n".join(lines) + "\n")
    print("\n".join(lines))


def load_profile() -> dict | None:
    if not PROFILE_JSON.is_file():
        return None
    p = json.loads(PROFILE_JSON.read_text())
    p["stale"] = p.get("sources_sha256") != sources_sha256()
    return p


# ---------------------------------------------------------------------------------------------
# workload helpers
# ---------------------------------------------------------------------------------------------
def view_settings(cam: dict, device, **flags):
    import torch
    from StochasticGSCudaBackend import Settings
    return Settings(torch.from_numpy(cam["w2c"]).to(device), torch.from_numpy(cam["cam_position"]).to(device),
                    cam["width"], cam["height"], cam["focal_x"], cam["focal_y"], cam["center_x"], cam["center_y"],
                    cam["near_plane"], cam["far_plane"], bool(cam.get("proper_antialiasing", False)),
                    int(cam.get("supersampling_factor", 1)), bool(flags.get("use_gps", cam.get("use_gps", False))),
                    bool(flags.get("hybrid_scheduling", cam.get("hybrid_scheduling", False))), False)


def host_settings(cam: dict):
    return view_settings(cam, "cpu")


def frames_for_step(views, frame_base: int) -> list[int]:
    """First RNG frame of every view, in the order run_plain assigns them."""
    out, f = [], frame_base
    for cam in views:
        out.append(f)
        f += cam["spp"]
    return out


def samples_per_step(views) -> int:
    return sum(v["width"] * v["height"] * max(1, int(v.get("supersampling_factor", 1))) ** 2 * v["spp"]
               for v in views)


def device_scene(arrays: dict, dev):
    import torch
    return {k: torch.from_numpy(v).to(dev) for k, v in arrays.items()}


def make_workload_scene(spec, morton: bool = True) -> dict:
    from StochasticGSCudaBackend.synthetic import make_scene, morton_order, reorder
    arrays = make_scene(spec)
    # scenes are stored Morton-ordered, as the reference's model loader does (Model.py:32-34)
    return reorder(arrays, morton_order(arrays["means"])) if morton else arrays


def run_plain(scene, views, settings, k_buffers, frame_base: int, tuning=None):
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
                          k_buffers[v], hist, eye, mode, frame_idx=frame, tuning=tuning)
            frame += 1
            mode = TemporalMode.STATIC
        out.append(hist)
    return out


# ---------------------------------------------------------------------------------------------
# rooflines
# ---------------------------------------------------------------------------------------------
def time_scatter_launches(scene, views, reps: int = 1):
    """Average duration of the accumulation launch (framebuffer clear + ST scatter kernel, the
    reference's schedule), measured with HIP events on the stream the library launches on, and
    its counters (from a counting launch outside the timed region)."""
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
    counters = torch.zeros(4, dtype=torch.int64, device=dev)
    work = []
    for a, fb in zip(args, fbs):
        _native.accumulate(*a, frame_idx=0, framebuffer=fb, counters=counters)
        work.append(counters.cpu().tolist())
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(len(views) * reps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(len(views) * reps)]
    k = 0
    scatter = _native.Tuning(schedule=2)
    for r in range(reps):
        for a, fb in zip(args, fbs):
            starts[k].record(stream)
            _native.accumulate(*a, frame_idx=1 + r, framebuffer=fb, tuning=scatter)
            ends[k].record(stream)
            k += 1
    torch.cuda.synchronize()
    ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return ms, work


def roofline_scatter(scene_n: int, views, ms, work, atomic_rate: float, copy_gbs: float):
    """The plain scatter kernel (one view, one sample per launch): algorithmic bytes 12 B of
    means per Gaussian (depth cull), 32 B more (scales, rotation, opacity) per Gaussian in
    range, 8 B clear of each framebuffer word and 8 B per atomicMin issued; atomics against the
    measured random-atomicMin rate of the device."""
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
    rate = atomics / t_total
    return {"kernel": "sgs::st_scatter_kernel (+ framebuffer clear)", "bound": "atomic",
            "achieved_atomics_per_s": round(rate, 1), "peak_atomics_per_s": round(atomic_rate, 1),
            "frac": round(rate / atomic_rate, 4),
            "hbm": {"algorithmic_gbs": round(achieved, 2), "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "frac_of_copy_ceiling": round(achieved / copy_gbs, 4)},
            "avg_launch_ms": round(t_total * 1e3 / len(ms), 4),
            "fragments_tested_per_launch": int(sum(w[0] for w in work) / n_views)}


def distinct_winners(scene, views):
    """Distinct winning Gaussians of one sample of every view (tile schedule, frame 7)."""
    import numpy as np
    import torch
    from StochasticGSCudaBackend import _native
    dev = scene["means"].device
    out = []
    for cam in views:
        fb = _native.accumulate(scene["means"], scene["scales"], scene["rotations"], scene["opacities"], scene["sh"],
                                torch.from_numpy(np.asarray(cam["w2c"], np.float32)).to(dev),
                                torch.from_numpy(np.asarray(cam["cam_position"], np.float32)).to(dev),
                                cam["width"], cam["height"], cam["focal_x"], cam["focal_y"], cam["center_x"],
                                cam["center_y"], cam["near_plane"], cam["far_plane"], False, 1, False, False, False,
                                frame_idx=7)
        won = fb[(fb >> 32) != 0xFFFFFFFF] & 0xFFFFFFFF
        out.append(int(torch.unique(won).numel()))
    return out


def step_bytes(scene_np, views, winners):
    """Algorithmic HBM bytes of one step, two ways.

    batched: what the batched renderer must move: per view the Gaussian parameters once
    (12 B means for every Gaussian, 32 B covariance and opacity for those inside [near, far]),
    per sample and framebuffer word 8 B written and 8 B read back by the resolve, per view the
    winners' means and SH (204 B per distinct winner of a sample, a lower bound) and 24 B per
    output pixel written (+24 B read for a history; none here: every view starts with RESET).
    per_sample: the reference's formulation, where every (view, sample) is its own render:
    B = 12 N + 32 N_z + 16 P + 204 U + 24 Q (1 + [mode != RESET]) per (view, sample)."""
    import numpy as np
    means = scene_np["means"].astype(np.float64)
    n = means.shape[0]
    batched = per_sample = 0.0
    for cam, u in zip(views, winners):
        w2c = np.asarray(cam["w2c"], np.float64)
        depth = means @ w2c[2, :3] + w2c[2, 3]
        n_z = int(np.count_nonzero((depth >= cam["near_plane"]) & (depth <= cam["far_plane"])))
        P = Q = cam["width"] * cam["height"]
        spp = int(cam["spp"])
        batched += 12.0 * n + 32.0 * n_z + 16.0 * P * spp + 204.0 * u + 24.0 * Q
        per_sample += spp * (12.0 * n + 32.0 * n_z + 16.0 * P + 204.0 * u) + 24.0 * Q * (1 + 2 * (spp - 1))
    return batched, per_sample


def kernel_table(profile, views, n_instances, visible_pairs, scene_n):
    """Per kernel of the batched render: profiled duration and counters (profile summary),
    measured HBM bytes and fraction of 8 TB/s, VALU issue fraction, and where a per-unit
    algorithmic byte count is defined, that and its fraction (over the profiled duration)."""
    if profile is None:
        return None
    words = samples_per_step(views)
    pairs_all = scene_n * len(views)
    # pair records are 4 B (8-bit tile corners) while every view's framebuffer has at most
    # 255 16-pixel tiles a side, else 8 B
    fb_side = max(max(v["width"], v["height"]) * max(1, int(v.get("supersampling_factor", 1)))
                  for v in views)
    record = 4.0 if (fb_side + 15) // 16 <= 255 else 8.0
    algo = {
        # 32 B read (scales, rotation) + opacity, 32 B written per Gaussian
        "sgs::cov3d_kernel": 64.0 * scene_n,
        # 44 B of parameters per Gaussian, a pair record written per (view, Gaussian)
        "sgs::batch_bin_kernel<false>": 44.0 * scene_n + record * pairs_all,
        # the pair record read back, 12 B of means per Gaussian, a 2 B depth bucket and a 4 B
        # Gaussian index written per instance
        "sgs::batch_bin_kernel<true>": record * pairs_all + 12.0 * scene_n + 6.0 * n_instances,
        # every bucket and index read once, a 4 B sorted index written per instance
        "sgs::tile_sort_kernel": 10.0 * n_instances,
    }
    raster_algo = 44.0 * visible_pairs + 4.0 * n_instances + 8.0 * words
    rows = []
    for name, e in profile["kernels"].items():
        per = e["dispatches"] / max(profile.get("renders") or 1, 1)
        row = {"kernel": name, "per_render": per, "profile_avg_ms": round(e["avg_ms"], 4)}
        if "hbm_bytes" in e:
            row["measured_bytes"] = round(e["hbm_bytes"])
            gbs = e["hbm_bytes"] / (e["avg_ms"] * 1e-3) / 1e9
            row["measured_gbs"] = round(gbs, 1)
            row["measured_frac"] = round(gbs / HBM_PEAK_GBS, 4)
        if "SQ_INSTS_VALU" in e["counters"]:
            row["valu_wave_instructions"] = e["counters"]["SQ_INSTS_VALU"]
        if "valu_busy" in e:
            row["valu_busy_frac"] = round(e["valu_busy"], 4)
        if name in algo:
            row["algorithmic_bytes"] = round(algo[name])
            row["algorithmic_frac"] = round(algo[name] / (e["avg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        rows.append(row)
    stage = {"kernels": " + ".join(RASTER_KERNELS), "algorithmic_bytes": round(raster_algo)}
    return rows, stage


def roofline_raster(profile, views, raster_ms: list, n_instances: int, visible_pairs: int, copy_gbs: float):
    """The dominant kernel: the tile rasteriser (its one-layer launch for views of one sample,
    its 8-layer launch over whole tiles for views of up to 8 samples and its 16-layer launch
    over half tiles for views of more, timed together with HIP events). It is bound by
    VALU issue: the SIMD quad-cycles with a VALU issue per launch (SQ_ACTIVE_INST_VALU less
    SQ_ACTIVE_INST_VALU2 of the profile summary) over the live launch time, against 256 CUs x 4
    SIMDs x 2.4 GHz / 4, with the wave64 VALU instructions (SQ_INSTS_VALU) and their rate beside
    it (above the busy rate where two instructions issue in one quad-cycle). Beside those the HBM
    view: algorithmic bytes (44 B per visible (view, Gaussian) pair, a 4 B instance word per
    list entry, 8 B per framebuffer word written) and the measured bytes (profile)."""
    words = samples_per_step(views)
    algo = 44.0 * visible_pairs + 4.0 * n_instances + 8.0 * words
    t = sum(raster_ms) / len(raster_ms) * 1e-3
    out = {"kernel": "sgs::batch_tile_raster_kernel<1, 16> + <8, 16> + <16, 8> (one stage)",
           "avg_launch_ms": round(t * 1e3, 3)}
    ks = [profile["kernels"].get(k) for k in RASTER_KERNELS] if profile else [None]
    if all(k is not None and "valu_busy_quad_cycles" in k for k in ks):
        busy = sum(k["valu_busy_quad_cycles"] for k in ks)
        valu = sum(k["counters"]["SQ_INSTS_VALU"] for k in ks)
        traffic = sum(k.get("hbm_bytes", 0.0) for k in ks)
        achieved = busy / t / 1e9
        out.update({"bound": "valu", "achieved": round(achieved, 1), "peak": VALU_PEAK_G_PER_S,
                    "unit": "G VALU-busy SIMD quad-cycles/s", "frac": round(achieved / VALU_PEAK_G_PER_S, 4),
                    "traffic": round(traffic),
                    "valu_busy_quad_cycles_per_launch": busy,
                    "valu_wave_instructions_per_launch": valu,
                    "valu_instructions_g_per_s": round(valu / t / 1e9, 1),
                    "profile_avg_ms": round(sum(k["avg_ms"] for k in ks), 3),
                    "profile": str(PROFILE_JSON.relative_to(REPO)), "profile_stale": profile["stale"]})
    else:
        out.update({"bound": "hbm", "achieved": round(algo / t / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(algo / t / 1e9 / HBM_PEAK_GBS, 4), "traffic": None})
    gbs = algo / t / 1e9
    out["hbm"] = {"algorithmic_bytes": round(algo), "algorithmic_gbs": round(gbs, 1),
                  "frac": round(gbs / HBM_PEAK_GBS, 4), "frac_of_copy_ceiling": round(gbs / copy_gbs, 4)}
    if out.get("traffic"):
        out["hbm"]["measured_gbs"] = round(out["traffic"] / t / 1e9, 1)
        out["hbm"]["measured_frac"] = round(out["traffic"] / t / 1e9 / HBM_PEAK_GBS, 4)
    out.update({"instances_per_launch": int(n_instances), "visible_pairs_per_launch": int(visible_pairs),
                "framebuffer_words_per_launch": int(words)})
    return out


def roofline_batch_scatter(scene_n: int, views, acc_ms: list, work, atomic_rate: float):
    """The scatter schedule's persistent launch: atomics (per-sample count x spp) against the
    measured random-atomicMin rate; HBM: 44 B per (view, Gaussian) + 8 B per atomic."""
    atomics = sum(work[v][1] * cam["spp"] for v, cam in enumerate(views))
    b = 44.0 * scene_n * len(views) + 8.0 * atomics
    t = sum(acc_ms) / len(acc_ms) * 1e-3
    rate = atomics / t
    return {"kernel": "sgs::batch_accumulate_kernel", "bound": "atomic", "achieved": round(rate / 1e9, 3),
            "peak": round(atomic_rate / 1e9, 3), "unit": "G atomicMin/s", "frac": round(rate / atomic_rate, 4),
            "traffic": None, "avg_launch_ms": round(t * 1e3, 3),
            "hbm": {"algorithmic_gbs": round(b / t / 1e9, 1), "frac": round(b / t / 1e9 / HBM_PEAK_GBS, 4)}}


# ---------------------------------------------------------------------------------------------
# CPU oracle baselines
# ---------------------------------------------------------------------------------------------
def cpu_baselines(scene_np, views, budget_s: float, others: dict):
    """The oracle (scalar C, -O2 without fast math) on this host, in one run:
    * headline: views of the 6M-Gaussian workload, smallest first, one sample each, full scene,
      OpenMP over Gaussians on every core this process may use, until ``budget_s`` is spent;
    * single thread: one 256x256 view of that workload, the 100k-Gaussian 800x600 frame in
      full (4 samples), and the 256x256 centre crops of the 1M (1080p, one sample) and 3M
      (point splatting at 2x supersampling, one sample) workloads, these four run side by side
      on four cores, one thread each;
    * the same three other workloads with OpenMP on every core, one after another.
    Crops are reported as crop pixel-samples per second (the crops are the scenes' densest
    region)."""
    from concurrent.futures import ThreadPoolExecutor

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
    head = {"value": round(samples / dt / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"{len(done)} views of the 6M-Gaussian workload ({', '.join(res)}), 1 spp each, full scene, "
                      f"accumulate+sample+resolve, OpenMP on {threads} threads, {dt:.1f} s"}

    scenes = {k: oracle.OracleScene(a["means"], a["scales"], a["rotations"], a["opacities"], a["sh"])
              for k, (a, _) in others.items()}

    def crop_of(cam):
        W, H = oracle.fb_dims(cam)
        x0, y0 = (W - 256) // 2, (H - 256) // 2
        return (x0, y0, x0 + 256, y0 + 256)

    def job(kind: str, n_threads: int):
        ts = time.perf_counter()
        if kind == "c4_view":
            cam = views[order[0]]
            oracle.render(sc, cam, 0, n_threads=n_threads)
            n = cam["width"] * cam["height"]
        elif kind == "c1":
            cam = others["c1"][1]
            hist = None
            for k in range(4):
                hist = oracle.render(scenes["c1"], cam, k, history=hist, temporal_mode=min(k, 1),
                                     n_threads=n_threads)[:4]
            n = 4 * cam["width"] * cam["height"]
        else:
            cam = others[kind][1]
            oracle.accumulate(scenes[kind], cam, 0, crop=crop_of(cam), n_threads=n_threads)
            n = 256 * 256
        s = time.perf_counter() - ts
        return {"msamples_per_s": round(n / s / 1e6, 5), "seconds": round(s, 2), "pixel_samples": n}

    kinds = ["c4_view"] + [k for k in ("c1", "c2", "c3") if k in others]
    with ThreadPoolExecutor(max_workers=len(kinds)) as ex:
        single = dict(zip(kinds, ex.map(lambda k: job(k, 1), kinds)))
    omp = {k: job(k, threads) for k in kinds if k != "c4_view"}
    head["single_thread"] = {"value": single["c4_view"]["msamples_per_s"], "cores": 1,
                             "sample": f"one {views[order[0]]['width']}x{views[order[0]]['height']} view of the "
                                       f"6M-Gaussian workload, 1 spp, {single['c4_view']['seconds']} s"}
    head["configs"] = {
        "note": "single-thread runs side by side on separate cores; OpenMP runs one after another on "
                f"{threads} threads",
        **{k: {"single_thread": single[k], "openmp": omp[k],
               "sample": {"c1": "100k Gaussians, 800x600, 4 samples (RESET + 3 STATIC), full frame, "
                                "accumulate+sample+resolve",
                          "c2": "1M Gaussians, 1920x1080 stochastic transparency, 256x256 centre crop, one sample, "
                                "accumulation",
                          "c3": "3M heavy-tailed Gaussians, point splatting (max 4096 points) at 2x2 supersampling, "
                                "256x256 centre crop of the 3840x2160 framebuffer, one sample, accumulation"}[k]}
           for k in kinds if k != "c4_view"}}
    return head


# ---------------------------------------------------------------------------------------------
# the other workloads on the GPU (the reference's benchmark method, Trainer.py:86-112)
# ---------------------------------------------------------------------------------------------
def time_frames(render_frame, n_iter: int = 10) -> dict:
    """A warm-up frame, then n_iter frames each bracketed by device synchronisations (as the
    reference's per-frame timing does)."""
    import numpy as np
    import torch
    render_frame(0)
    torch.cuda.synchronize()
    ms = []
    for k in range(n_iter):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        render_frame(1 + k)
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - ts) * 1e3)
    return {"ms_per_frame": round(float(np.mean(ms)), 3), "p50_ms": round(float(np.percentile(ms, 50)), 3),
            "p99_ms": round(float(np.percentile(ms, 99)), 3)}


def bench_configs(others: dict, dev) -> dict:
    """1M Gaussians at 1920x1080 with 16 samples per frame (stochastic transparency: the
    reference's scatter schedule, hybrid scheduling, the default tile schedule, and the batched
    renderer with the frame's 16 samples in one call) and 3M heavy-tailed Gaussians with point
    splatting (max 4096 points) at 2x2 supersampling, 8 samples per frame (scatter, hybrid).
    Msamples/s = framebuffer (sub)pixels x samples per frame / frame time."""
    import torch
    from StochasticGSCudaBackend import History, Tuning, create_k_buffer, render
    from StochasticGSCudaBackend.batch import BatchView, render_batch
    out = {}
    eye = torch.eye(4, device=dev)
    empty = History(*(torch.empty(0) for _ in range(4)))
    for key, spp, modes in (("c2", 16, [("scatter", {}, Tuning(schedule=2)), ("hybrid", {"hybrid_scheduling": True}, None),
                                        ("tiles", {}, None), ("batched", {}, "batch")]),
                            ("c3", 8, [("scatter", {"use_gps": True}, Tuning(max_gps_points=4096)),
                                       ("hybrid", {"use_gps": True, "hybrid_scheduling": True},
                                        Tuning(max_gps_points=4096))])):
        if key not in others:
            continue
        arrays, cam = others[key]
        scene = device_scene(arrays, dev)
        ss = max(1, int(cam.get("supersampling_factor", 1)))
        per_frame = cam["width"] * cam["height"] * ss * ss * spp
        res = {"pixel_samples_per_frame": per_frame, "samples_per_frame": spp}
        for name, flags, tuning in modes:
            s = view_settings(cam, dev, **flags)
            kb = create_k_buffer(s)

            if tuning == "batch":
                def frame(k, s=s):
                    render_batch(scene["means"], scene["scales"], scene["rotations"], scene["opacities"], scene["sh"],
                                 [BatchView(s, spp=spp, frame_idx=1000 * k)])
            else:
                def frame(k, s=s, kb=kb, tuning=tuning):
                    hist = empty
                    for j in range(spp):
                        hist = render(scene["means"], scene["scales"], scene["rotations"], scene["opacities"],
                                      scene["sh"], s, kb, hist, eye, 0 if j == 0 else 1, frame_idx=1000 * k + j,
                                      tuning=tuning)
            t = time_frames(frame)
            t["msamples_per_s"] = round(per_frame / (t["ms_per_frame"] * 1e-3) / 1e6, 1)
            res[name] = t
            log(f"{key} {name}: {t['ms_per_frame']} ms per frame, {t['msamples_per_s']} Msamples/s")
        out[key] = res
        del scene
        torch.cuda.empty_cache()
    return out


# ---------------------------------------------------------------------------------------------
# serving: 32 streams on smooth trajectories, Poisson arrivals, continuous batching
# ---------------------------------------------------------------------------------------------
SERVE_SIZES = [(1280, 720), (1920, 1080), (2560, 1440)]
SERVE_SPP = [1, 2, 8]


def serving_streams(n_streams: int, seed: int = 3, spatial_reuse: bool = True):
    import numpy as np
    from StochasticGSCudaBackend.serve import StreamSpec
    from StochasticGSCudaBackend.synthetic import camera, orbit_eye
    rng = np.random.default_rng(seed)
    specs, paths = [], []
    for _ in range(n_streams):
        w, h = SERVE_SIZES[int(rng.integers(len(SERVE_SIZES)))]
        c = camera(orbit_eye(0.0, 3.0), w, h)
        specs.append(StreamSpec(w, h, c["focal_x"], c["focal_y"], c["center_x"], c["center_y"], c["near_plane"],
                                c["far_plane"], spatial_reuse=spatial_reuse))
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


def bench_serving(args, scene, dev, inflight: int, with_plain: bool):
    """Saturated throughput (every request queued at t=0), then Poisson arrivals at 80 % of it:
    p50/p99 latency (arrival to completion seen by the host). The plain per-request path is
    timed on the same requests (spp single-view render() calls each) and its latency at the
    same arrivals follows from a FIFO single-server queue over those service times."""
    import numpy as np
    import torch
    from StochasticGSCudaBackend import _native
    from StochasticGSCudaBackend.serve import Server, latency_stats
    reuse = not args.serve_temporal_only
    specs, paths = serving_streams(args.serve_streams, spatial_reuse=reuse)
    server = Server(scene["means"], scene["scales"], scene["rotations"], scene["opacities"], scene["sh"], specs,
                    max_views=args.serve_streams, inflight=inflight, max_wait=args.serve_window_ms * 1e-3)
    warm = serving_requests(specs, paths, [0.0] * args.serve_streams, seed=10)
    server.run(warm)
    sat = serving_requests(specs, paths, [0.0] * (4 * args.serve_streams), seed=11)
    t0 = time.perf_counter()
    server.run(sat)
    sat_s = time.perf_counter() - t0
    sat_rate = len(sat) / sat_s
    log(f"serving (inflight {inflight}) saturated: {sat_rate:.1f} req/s")
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
    log(f"serving (inflight {inflight}) at {rate:.1f} req/s: p50 {stats['p50_ms']} ms, p99 {stats['p99_ms']} ms, "
        f"{len(done) / batches:.1f} requests per batch")
    out = {"streams": args.serve_streams, "inflight": inflight, "window_ms": args.serve_window_ms,
           "resolve": "spatio-temporal (spatial reuse + REPROJECT)" if reuse else "temporal (REPROJECT)",
           "saturated_req_per_s": round(sat_rate, 1), "offered_req_per_s": round(rate, 1),
           "completed_req_per_s": round(len(done) / run_s, 1), "msamples_per_s": round(samples / run_s / 1e6, 1),
           "requests": len(done), "requests_per_batch": round(len(done) / max(batches, 1), 2),
           "p50_ms": stats["p50_ms"], "p99_ms": stats["p99_ms"],
           "p99_over_p50": round(stats["p99_ms"] / max(stats["p50_ms"], 1e-9), 3), "overflowed_batches": overflowed}
    if with_plain:
        sample_reqs = load[:min(len(load), 96)]
        hist_cache, kb_cache = {}, {}
        service = []
        eye4 = torch.eye(4, device=dev)
        noise = torch.empty(0, dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        for r in sample_reqs:
            sp = specs[r.stream]
            if (sp.width, sp.height) not in hist_cache:
                hist_cache[(sp.width, sp.height)] = tuple(torch.zeros((sp.height, sp.width, c), device=dev)
                                                          for c in (3, 1, 1, 1))
            hist = hist_cache[(sp.width, sp.height)]
            if (sp.width, sp.height) not in kb_cache:
                P, K = sp.width * sp.height, _native.k_buffer_size
                kb_cache[(sp.width, sp.height)] = (
                    (torch.full((K, P, 2), -1, dtype=torch.int32, device=dev), torch.zeros((K, P, 4), device=dev))
                    if reuse else (torch.empty((0, 0, 2), dtype=torch.int32, device=dev),
                                   torch.empty((0, 0, 4), device=dev)))
            kb = kb_cache[(sp.width, sp.height)]
            w2c = torch.from_numpy(np.asarray(r.w2c, np.float32)).to(dev)
            cp = torch.from_numpy(np.asarray(r.cam_position, np.float32)).to(dev)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            h = hist
            for k in range(r.spp):
                h = _native.render(scene["means"], scene["scales"], scene["rotations"], scene["opacities"],
                                   scene["sh"], w2c, cp, sp.width, sp.height, sp.focal_x, sp.focal_y, sp.center_x,
                                   sp.center_y, sp.near_plane, sp.far_plane, False, 1, False, False, reuse, noise,
                                   kb[0], kb[1], *h, eye4, 2 if k == 0 else 1, frame_idx=r.frame_idx + k)
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
        out["plain"] = {"req_per_s_capacity": round(1.0 / mean_service, 1),
                        "p50_ms_at_offered_load": round(float(np.percentile(lat, 50)), 3),
                        "p99_ms_at_offered_load": round(float(np.percentile(lat, 99)), 3),
                        "service_ms_mean": round(mean_service * 1e3, 3)}
    del server
    torch.cuda.empty_cache()
    return out


# ---------------------------------------------------------------------------------------------
# the headline workload
# ---------------------------------------------------------------------------------------------
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
        args.tile_samples = tuning.tile_samples
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
    for e5 in ev:
        for e in e5:
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
             "resolve_ms": round(float(np.mean([e5[3].elapsed_time(e5[4]) for e5 in ev])), 3),
             "device_ms": round(float(np.mean([e5[0].elapsed_time(e5[4]) for e5 in ev])), 3)}
    n_inst, n_visible, overflowed = br.tile_stats()
    log(f"stages {stage}; tile instances {n_inst} (overflow {overflowed})")
    single = {"value": round(S * args.steps / batch_s / 1e6, 2), "ms_per_step": round(batch_s * 1e3 / args.steps, 2),
              "tile_overflow": bool(overflowed)}
    head = br
    head_overflow = bool(overflowed)
    if streams > 1:
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
        head_overflow = head.tile_stats()[2]
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
            "tile_instances": n_inst, "visible_pairs": n_visible, "tile_overflow": bool(head_overflow),
            "autotune": tuned, "max_abs_diff_vs_plain": max_diff}, main_ms


def profile_mode(args, scene, views) -> int:
    """Only what the profile summary needs: a 1 GiB calibration copy, then single-batch renders
    of the workload with the default tuning (the configuration the autotuner selects for each
    batch), so every dispatch of the batch kernels in the trace belongs to these renders."""
    import torch
    from StochasticGSCudaBackend import microbench
    from StochasticGSCudaBackend.batch import BatchRenderer, BatchView
    src = torch.ones(COPY_CALIBRATION_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    microbench._check(microbench.library().sgs_mb_copy(dst.data_ptr(), src.data_ptr(), COPY_CALIBRATION_BYTES,
                                                       torch.cuda.current_stream().cuda_stream), "copy")
    torch.cuda.synchronize()
    del src, dst
    bviews = [BatchView(host_settings(cam), spp=cam["spp"], frame_idx=0) for cam in views]
    br = BatchRenderer(scene["means"], scene["scales"], scene["rotations"], scene["opacities"], scene["sh"], bviews)
    for k in range(args.profile_renders):
        br.set_frames(frames_for_step(views, 10_000 * k))
        br.upload_table()
        br.render()
        torch.cuda.synchronize()
        log(f"profile render {k + 1}/{args.profile_renders}")
    n_inst, n_visible, overflowed = br.tile_stats()
    print(json.dumps({"profile_renders": args.profile_renders, "tile_instances": n_inst, "visible_pairs": n_visible,
                      "overflow": bool(overflowed)}), flush=True)
    return 0


def main() -> int:
    args = parse()
    if args.summarize_profile:
        summarize_profile(*args.summarize_profile)
        return 0
    if args.summarize_trace:
        summarize_trace(*args.summarize_trace, title=args.trace_title, command=args.trace_command)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this renderer runs on one GPU; --gpus must be 1 (views and requests would shard "
                                   "across GPUs with the scene replicated and no exchange; not built)"}))
        return 2
    import torch
    from StochasticGSCudaBackend import create_k_buffer
    from StochasticGSCudaBackend.synthetic import (camera_c1, camera_c2, camera_c3, scaled, scene_c1, scene_c2, scene_c3,
                                                   scene_c4, views_c4)

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t0 = time.perf_counter()
    spec = scaled(scene_c4(), args.gaussians)
    scene_np = make_workload_scene(spec, morton=not args.no_morton)
    scene = device_scene(scene_np, dev)
    views = views_c4(args.views)
    if args.profile:
        return profile_mode(args, scene, views)
    settings = [view_settings(v, dev) for v in views]
    k_buffers = [create_k_buffer(s) for s in settings]
    torch.cuda.synchronize()
    S = samples_per_step(views)
    log(f"scene {spec.n} Gaussians, {len(views)} views, {S / 1e6:.1f} M pixel-samples/step "
        f"(setup {time.perf_counter() - t0:.1f} s)")

    from StochasticGSCudaBackend import microbench
    ceilings = {"copy": microbench.copy_ceiling(), "atomic_min_u64": microbench.atomic_min_rate()}
    copy_gbs = ceilings["copy"]["gbs"]
    atomic_rate = ceilings["atomic_min_u64"]["ops_per_s"]
    log(f"ceilings: copy {copy_gbs} GB/s, random u64 atomicMin {atomic_rate / 1e9:.2f} G/s")

    result = {}
    plain_out0 = None
    if args.only in ("plain", "all"):
        result["plain"], plain_out0 = bench_plain(args, scene, views, settings, k_buffers, S)
    acc_ms = None
    if args.only in ("batched", "all"):
        result["batched"], acc_ms = bench_batched(args, scene, views, S, plain_out0, dev)
        if "plain" in result:
            result["batched"]["speedup_vs_plain"] = round(result["batched"]["value"] / result["plain"]["value"], 3)
    serving = None
    if args.only == "all" and args.serve_seconds > 0:
        serving = [bench_serving(args, scene, dev, int(n), with_plain=(i == 0))
                   for i, n in enumerate(args.serve_inflight.split(","))]

    # ---- rooflines --------------------------------------------------------------------------
    profile = load_profile()
    ms, work = time_scatter_launches(scene, views)
    roof_scatter = roofline_scatter(spec.n, views, ms, work, atomic_rate, copy_gbs)
    log(f"scatter launches: avg {roof_scatter['avg_launch_ms']} ms, {roof_scatter['frac']} of the atomic rate")
    roof = roof_scatter
    kernels = None
    step = None
    if "batched" in result:
        b = result["batched"]
        if args.schedule == "tiles":
            roof = roofline_raster(profile, views, acc_ms, b["tile_instances"], b["visible_pairs"], copy_gbs)
            log(f"tile raster: {roof['avg_launch_ms']} ms, {roof['frac']} of {roof['unit']} peak")
            kt = kernel_table(profile, views, b["tile_instances"], b["visible_pairs"], spec.n)
            if kt is not None:
                kernels = {"rows": kt[0], "raster_stage": kt[1], "profile": str(PROFILE_JSON.relative_to(REPO)),
                           "profile_stale": profile["stale"]}
        else:
            roof = roofline_batch_scatter(spec.n, views, acc_ms, work, atomic_rate)
        winners = distinct_winners(scene, views)
        batched_b, per_sample_b = step_bytes(scene_np, views, winners)
        t_single = b["stages"]["device_ms"] * 1e-3
        step = {"definition": "batched-design algorithmic bytes of one step (parameters once per view, 16 B per "
                              "framebuffer word per sample, 204 B per distinct winner, 24 B per output pixel) over "
                              "the single batch's device time; measured bytes: the profile summary's HBM traffic of "
                              "every kernel of one render",
                "algorithmic_bytes": round(batched_b), "device_ms": round(t_single * 1e3, 3),
                "algorithmic_gbs": round(batched_b / t_single / 1e9, 1),
                "frac": round(batched_b / t_single / 1e9 / HBM_PEAK_GBS, 4),
                "per_sample_formulation": {"bytes": round(per_sample_b),
                                           "note": "B = 12N + 32N_z + 16P + 204U + 24Q(1 + [mode != RESET]) per "
                                                   "(view, sample): what rendering each sample as its own call "
                                                   "must move; the batched design reads each view's Gaussians once"},
                "global_atomics": 0 if args.schedule == "tiles" else None,
                "atomic_note": "the tile schedule resolves fragments with LDS ds_min_u64 and issues no global "
                               "atomics" if args.schedule == "tiles" else None}
        if profile is not None and args.schedule == "tiles":
            per_render = sum(e.get("hbm_bytes", 0.0) * e["dispatches"] / max(profile.get("renders") or 1, 1)
                             for e in profile["kernels"].values())
            step.update({"measured_bytes": round(per_render),
                         "measured_gbs": round(per_render / t_single / 1e9, 1),
                         "measured_frac": round(per_render / t_single / 1e9 / HBM_PEAK_GBS, 4),
                         "profile_stale": profile["stale"]})

    configs = None
    if args.only == "all" and not args.no_configs:
        others = {"c1": (make_workload_scene(scene_c1()), camera_c1()),
                  "c2": (make_workload_scene(scene_c2()), camera_c2()),
                  "c3": (make_workload_scene(scene_c3()), camera_c3(use_gps=True, max_gps_points=4096))}
        configs = bench_configs(others, dev)
    else:
        others = {}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baselines(scene_np, views, args.cpu_seconds, others)
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
        "roofline_step": step,
        "kernels": kernels,
        "roofline_scatter": roof_scatter,
        "ceilings": ceilings,
        "serving": serving[0] if serving else None,
        "serving_inflight": serving,
        "configs": configs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
