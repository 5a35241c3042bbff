"""bench.py's command-line contract on the CPU: a request for more than one GPU ends with
exit status 2 and one JSON line that says why, before any GPU work."""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

REPO = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=REPO)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def _write_csv(path: Path, header: list[str], rows: list[list]) -> None:
    import csv
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(header)
        w.writerows(rows)


def test_profile_summary_from_rocprof_csv(tmp_path, monkeypatch):
    """The summarizer groups rocprofv3 dispatches by (kernel, grid), scales the byte counters
    by the 1 GiB calibration copy and derives VALU issue per CU-cycle."""
    sys.path.insert(0, str(REPO))
    import bench
    monkeypatch.setattr(bench, "PROFILE_JSON", tmp_path / "p.json")
    monkeypatch.setattr(bench, "PROFILE_MD", tmp_path / "p.md")
    copy = "(anonymous namespace)::copy_kernel(HIP_vector_type<unsigned int, 4u> const*, unsigned long)"
    raster = "void sgs::batch_tile_raster_kernel<8u>(sgs::SceneArgs, unsigned int)"
    cov = "sgs::cov3d_kernel(sgs::SceneArgs, HIP_vector_type<float, 4u>*)"
    th = ["Kind", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"]
    trace_rows = [["KERNEL_DISPATCH", copy, 0, 250_000, 524288, 1, 1]]
    for r in range(2):
        trace_rows += [["KERNEL_DISPATCH", cov, 0, 70_000, 6_000_000, 1, 1],
                       ["KERNEL_DISPATCH", raster, 0, 9_000_000 + 1000 * r, 1 << 20, 1, 1],
                       ["KERNEL_DISPATCH", raster, 0, 10_000, 1 << 10, 1, 1]]   # another grid: a single view
    _write_csv(tmp_path / "trace" / "h" / "run_kernel_trace.csv", th, trace_rows)
    ch = ["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"]
    _write_csv(tmp_path / "fetch" / "run_counter_collection.csv", ch,
               [[copy, 524288, "FETCH_SIZE", 524288.0], [raster, 1 << 20, "FETCH_SIZE", 1_000_000.0],
                [cov, 6_000_000, "FETCH_SIZE", 100_000.0]])
    _write_csv(tmp_path / "write" / "run_counter_collection.csv", ch,
               [[copy, 524288, "WRITE_SIZE", 1048576.0], [raster, 1 << 20, "WRITE_SIZE", 2_000_000.0],
                [cov, 6_000_000, "WRITE_SIZE", 200_000.0]])
    _write_csv(tmp_path / "sq" / "run_counter_collection.csv", ch,
               [[raster, 1 << 20, "SQ_INSTS_VALU", 5.0e9], [raster, 1 << 20, "GRBM_GUI_ACTIVE", 8 * 2.0e7],
                [copy, 524288, "SQ_INSTS_VALU", 1.0e6], [copy, 524288, "GRBM_GUI_ACTIVE", 1.0e6]])
    bench.summarize_profile(*(str(tmp_path / d) for d in ("trace", "fetch", "write", "sq")))
    out = json.loads((tmp_path / "p.json").read_text())
    assert out["calibration"]["fetch_scale"] == 2.0 and out["calibration"]["write_scale"] == 1.0
    assert out["renders"] == 2
    r = out["kernels"]["sgs::batch_tile_raster_kernel<8u>"]
    assert r["grid_lanes"] == 1 << 20 and r["dispatches"] == 2
    assert abs(r["avg_ms"] - 9.0005) < 1e-9
    assert r["hbm_bytes"] == 1_000_000 * 1024 * 2.0 + 2_000_000 * 1024
    assert abs(r["valu_per_cu_cycle"] - 5.0e9 / (256 * 2.0e7)) < 1e-12
    assert list(out["kernels"]) == ["sgs::cov3d_kernel", "sgs::batch_tile_raster_kernel<8u>"]
    assert "batch_tile_raster_kernel<8u>" in (tmp_path / "p.md").read_text()
    assert bench.short_kernel_name("void a::k<b<1>, 2u>(x, y)") == "a::k<b<1>, 2u>"
