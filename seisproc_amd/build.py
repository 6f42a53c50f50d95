"""Build libpstm_hip.so (gfx950) in-tree with hipcc.

    python -m seisproc_amd.build          # compile every translation unit, link

The library is built next to this file (``seisproc_amd/libpstm_hip.so``) so
that it travels with the source tree; nothing is installed.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT_DIR = HERE / "_obj"
LIB = HERE / "libpstm_hip.so"
SOURCES = ["pstm_abi.hip", "pstm_plan.hip", "pstm_fwd_f32.hip", "pstm_fwd_bf16.hip", "pstm_fwd_f16.hip",
           "pstm_adj.hip", "pstm_grad.hip", "pstm_static.hip", "pstm_integ.hip"]
ARCH = os.environ.get("PSTM_OFFLOAD_ARCH", "gfx950")
# Standalone probe (not part of the library): VALU issue rates of the
# instruction kinds the tile kernels use, ./seisproc_amd/valu_rates
PROBE_SRC = "valu_rates.hip"
PROBE = HERE / "valu_rates"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # keep the compiler from emitting packed-FP32 VALU instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def _compile(src: str, extra: list[str]) -> Path:
    obj = OUT_DIR / (Path(src).stem + ".o")
    cmd = [_hipcc(), *HIPCC_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False, extra: list[str] | None = None) -> Path:
    """Compile all sources and link ``libpstm_hip.so``. Always recompiles."""
    OUT_DIR.mkdir(exist_ok=True)
    extra = list(extra or [])
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, extra), SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *[str(o) for o in objs]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_probe(verbose)
    return LIB


def build_probe(verbose: bool = False) -> Path:
    """Compile the VALU issue-rate probe (an executable, same flags)."""
    tmp = PROBE.with_name(PROBE.name + ".tmp")
    cmd = [_hipcc(), *HIPCC_FLAGS, str(CSRC / PROBE_SRC), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {PROBE_SRC}:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, PROBE)
    if verbose:
        print(f"built {PROBE}")
    return PROBE


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
