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
    assert ROOFLINE <= rf.keys() and rf["bound"] in ("hbm", "mfma", "valu")
    if rf["achieved"] is not None:
        assert rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"], rel=1e-3)
    if o["cpu_baseline"] is not None:
        cb = o["cpu_baseline"]
        assert CPU <= cb.keys() and cb["kind"] in ("port", "reference") and cb["cores"] >= 1 and cb["value"] > 0


def test_counter_roofline_from_a_summary_record():
    # the roofline object computed from a counter summary record (no GPU)
    sys.path.insert(0, str(ROOT))
    import bench
    rec = {"contributions": 1.0e13, "calibration": {"fetch_factor": 2.0, "write_factor": 1.0},
           "counters": {"SQ_INSTS_VALU": 3.0e12, "SQ_INSTS_VALU_TRANS_F32": 3.5e11, "SQ_BUSY_CYCLES": 4.0e11,
                        "FETCH_SIZE_KiB": 1.0e9, "WRITE_SIZE_KiB": 5.0e8, "SQ_INSTS_LDS": 4.0e11,
                        "SQ_INSTS_LDS_LOAD_BANDWIDTH": 3.0e12, "SQ_INSTS_LDS_STORE_BANDWIDTH": 1.0e11,
                        "SQ_LDS_BANK_CONFLICT": 1.0e9, "SQ_LDS_IDX_ACTIVE": 1.0e11}, "source_sha": "x"}
    r = bench.roofline(1.0e13, 5.0, 4, (rec, "profiles/x.md"))
    assert r["bound"] == "valu"
    assert r["achieved"] == pytest.approx(3.0e12 * 64 / 5.0 / 1e12, rel=1e-3)
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"], rel=1e-3)
    assert r["traffic"] == pytest.approx(1.0e9 * 1024 * 2.0 + 5.0e8 * 1024, rel=1e-9)
    assert r["valu_per_contribution"] == pytest.approx(19.2)
    assert r["issue_cycles_per_valu"] == pytest.approx((4.0e11 / 32) / (3.0e12 / 1024), rel=1e-3)
    assert r["hbm_algorithmic"]["frac"] == pytest.approx(1.0e13 * 4 / 5.0 / 1e9 / 8000.0, rel=1e-3)
    # unpacked FP32 ceiling: half the published (packed) FP32 figure
    assert r["unpacked_fp32"]["peak"] == pytest.approx(r["peak"] / 2, rel=2e-3)
    assert r["unpacked_fp32"]["frac"] == pytest.approx(r["achieved"] / r["unpacked_fp32"]["peak"], rel=1e-3)
    stale = bench.roofline(1.0e13, 5.0, 4, (None, "no summary"))
    assert stale["achieved"] is None and stale["traffic"] is None and stale["counters"]["note"] == "no summary"


def test_default_line_has_measured_traffic_and_cpu_baseline():
    o = json.loads((ROOT / "profiles" / "bench_pstm3d_precision.json").read_text())
    assert o["config"]["workload"] == "pstm3d-precision"
    assert o["roofline"]["traffic"] and o["roofline"]["traffic"] > 0
    assert o["cpu_baseline"] is not None and o["steps"] >= 5
