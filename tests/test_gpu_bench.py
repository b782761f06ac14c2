"""bench.py's output contract on the GPU: one JSON line with the metric, the whole-job value, the
timing fields, the roofline object (from the kernel's own launch time) and the CPU baseline."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_line_contract():
    print("bench.py --workload C2 --steps 20 --warmup 5 (short line)", flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--workload", "C2", "--steps", "20",
                        "--warmup", "5", "--no-storage-sweep", "--no-adjoint", "--cpu-seconds", "0.5"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.strip().splitlines() if ln.startswith("{")]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
                "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 20 and d["warmup"] == 5 and d["higher_is_better"] is True
    assert d["value"] > 0 and d["unit"] == "MLUPS" and "workload" in d["config"]
    cells = d["config"]["cells"]
    assert abs(d["value"] - cells / (d["ms_per_step"] * 1e3)) <= 1e-3 * d["value"]
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    # on 225 B per cell update the fraction can pass 1 (the row-compacted layout does not move the frozen
    # solid cells' NS populations); on the bytes moved it cannot
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3 and 0 < rf["frac"] < 1.6
    # FP32 runs in row-compacted NS planes by default: the counter summary of that kernel, and the same
    # time over the bytes the compact run moves
    assert rf["traffic"] > 0 and rf["traffic_source"].startswith("profiles/traffic_rc_c2_f32.json")
    assert d["config"]["ns_layout"].startswith("row-compacted") and rf["kernel"].startswith("k_ns_split + k_ad_split")
    cl = rf["compact_layout"]
    assert 0.3 < cl["active_fraction"] < 0.5 and cl["bytes_per_cell_update"] < rf["bytes_per_cell_update"]
    assert 0 < cl["frac"] < 1
    assert abs(cl["frac"] / rf["frac"] - cl["bytes_per_cell_update"] / rf["bytes_per_cell_update"]) < 1e-3
    cb = d["cpu_baseline"]
    assert cb["cores"] == 1 and cb["kind"] in ("port", "reference") and cb["value"] > 0 and cb["sample"]
