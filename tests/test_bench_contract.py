"""The bench line contract, checked on the CPU: bench.py refuses more than one
GPU before touching a device, and every committed bench line carries the keys
and the roofline / CPU-baseline objects the line format requires."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LINES = sorted((ROOT / "profiles").glob("bench_*.json"))

REQUIRED = {"metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
            "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"}
ROOFLINE = {"bound", "achieved", "peak", "unit", "frac", "traffic"}
CPU = {"value", "unit", "cores", "kind", "sample"}


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_committed_lines_exist():
    names = {p.name for p in LINES}
    assert "bench_pstm3d_precision.json" in names and "bench_lsm3d.json" in names


@pytest.mark.parametrize("path", LINES, ids=[p.stem for p in LINES])
def test_committed_bench_line_shape(path):
    lines = [ln for ln in path.read_text().splitlines() if ln.strip()]
    assert len(lines) == 1, "one JSON line per bench run"
    o = json.loads(lines[0])
    assert REQUIRED <= o.keys()
    assert o["n_gpus"] == 1 and o["value"] > 0 and o["ms_per_step"] > 0 and o["higher_is_better"] is True
    assert o["data"].startswith("synthetic") and isinstance(o["config"], dict) and "workload" in o["config"]
    rf = o["roofline"]
    assert ROOFLINE <= rf.keys() and rf["bound"] in ("hbm", "mfma")
    assert rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"], rel=1e-3)
    if o["cpu_baseline"] is not None:
        cb = o["cpu_baseline"]
        assert CPU <= cb.keys() and cb["kind"] in ("port", "reference") and cb["cores"] >= 1 and cb["value"] > 0


def test_default_line_has_measured_traffic_and_cpu_baseline():
    o = json.loads((ROOT / "profiles" / "bench_pstm3d_precision.json").read_text())
    assert o["config"]["workload"] == "pstm3d-precision"
    assert o["roofline"]["traffic"] and o["roofline"]["traffic"] > 0
    assert o["cpu_baseline"] is not None and o["steps"] >= 5
