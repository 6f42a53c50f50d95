"""FP64 CPU oracle for the Kirchhoff PSTM path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
path (``seisproc_amd``) never imports it and has no CPU fallback.

Contents
  pstm_oracle.c   FP64 C/OpenMP restatement of the reference method
                  (processors/migration/kirchhoff_kernel.py:459-854), plus the
                  exact adjoint and the analytic velocity gradient.
  numpy_ref.py    brute-force numpy restatement for tiny inputs (cross-check of
                  the C code's loops and index handling).
  ref_torch.py    FP32 torch-CPU restatement of the reference tile loop; timed as
                  the CPU baseline.

Parity status (how each oracle function is pinned; see tests/test_oracle_*.py):
  forward      known answers restated from the reference tests (DSR times
               1000 ms and 2*sqrt(1000^2+1500^2)/3000 s, interpolation exact /
               half-sample / clamp, aperture and angle masks, weight maximal at
               zero offset, fold normalisation), brute-force numpy equality,
               single-trace closed forms, diffractor focusing at tau0.
  antialias    L = 1 equals linear interpolation exactly; brute-force triangle
               convolution of the zero-extended trace.
  vti          eta = 0 equals the isotropic time; Alkhalifah closed form.
  adjoint      dot-product test <L m, d> = <m, L^T d> to 1e-12.
  vel_grad     central finite differences of forward with frozen masks.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_REPO = _HERE.parent
_LIB_PATH = _HERE / "liboracle_pstm.so"
_lib = None

if str(_REPO) not in sys.path:
    sys.path.insert(0, str(_REPO))


def build(force: bool = True) -> Path:
    """Compile pstm_oracle.c with gcc/OpenMP into oracle/liboracle_pstm.so."""
    src = _HERE / "pstm_oracle.c"
    if not force and _LIB_PATH.exists() and _LIB_PATH.stat().st_mtime >= src.stat().st_mtime:
        return _LIB_PATH
    tmp = _LIB_PATH.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c99", "-D_GNU_SOURCE",
           "-o", str(tmp), str(src), "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, _LIB_PATH)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build(force=True)
        lib = ctypes.CDLL(str(_LIB_PATH))
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        lib.oracle_forward.argtypes = [vp, dp, dp, dp, dp, dp, dp, vp, vp, dp, vp]
        lib.oracle_forward.restype = ctypes.c_int
        lib.oracle_adjoint.argtypes = [vp, dp, dp, dp, dp, dp, dp, vp, dp]
        lib.oracle_adjoint.restype = ctypes.c_int
        lib.oracle_vel_grad.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp]
        lib.oracle_vel_grad.restype = ctypes.c_int
        lib.oracle_leg_time_vti.argtypes = [ctypes.c_double] * 4
        lib.oracle_leg_time_vti.restype = ctypes.c_double
        lib.oracle_pair.argtypes = [vp] + [ctypes.c_double] * 7 + [dp, dp, dp]
        lib.oracle_pair.restype = ctypes.c_int
        lib.oracle_set_exhaustive.argtypes = [ctypes.c_int]
        _lib = lib
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _params_struct(params):
    # The params struct layout is shared with the product ABI; the oracle
    # imports the plain ctypes mirror, not the product loader.
    from seisproc_amd.abi_types import PstmParams
    return params if isinstance(params, PstmParams) else PstmParams.from_dict(params)


def set_exhaustive(on: bool) -> None:
    """Evaluate the mask at every tau instead of the monotone binary search."""
    _load().oracle_set_exhaustive(1 if on else 0)


def forward(params, traces, sx, sy, rx, ry, vel, eta=None, mask_vel=None):
    """FP64 forward migration. Returns (image, fold) as [il][xl][tau] arrays."""
    lib = _load()
    p = _params_struct(params)
    tr = _f64(traces).reshape(p.n_traces, p.n_samples)
    sx, sy, rx, ry = (_f64(a) for a in (sx, sy, rx, ry))
    vel = _f64(vel)
    eta_a = _f64(eta) if eta is not None else None
    mv = _f64(mask_vel) if mask_vel is not None else None
    img = np.zeros((p.n_il, p.n_xl, p.n_tau))
    fold = np.zeros((p.n_il, p.n_xl, p.n_tau))
    rc = lib.oracle_forward(ctypes.byref(p), _dp(tr), _dp(sx), _dp(sy), _dp(rx), _dp(ry), _dp(vel),
                            eta_a.ctypes.data if eta_a is not None else None,
                            mv.ctypes.data if mv is not None else None, _dp(img),
                            fold.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle_forward failed: {rc}")
    return img, fold


def adjoint(params, image, sx, sy, rx, ry, vel, eta=None):
    """FP64 adjoint (demigration). Returns traces [n_traces][n_samples]."""
    lib = _load()
    p = _params_struct(params)
    img = _f64(image).reshape(p.n_il, p.n_xl, p.n_tau)
    sx, sy, rx, ry = (_f64(a) for a in (sx, sy, rx, ry))
    vel = _f64(vel)
    eta_a = _f64(eta) if eta is not None else None
    out = np.zeros((p.n_traces, p.n_samples))
    rc = lib.oracle_adjoint(ctypes.byref(p), _dp(img), _dp(sx), _dp(sy), _dp(rx), _dp(ry), _dp(vel),
                            eta_a.ctypes.data if eta_a is not None else None, _dp(out))
    if rc != 0:
        raise RuntimeError(f"oracle_adjoint failed: {rc}")
    return out


def vel_grad(params, traces, sx, sy, rx, ry, vel, image_grad):
    """FP64 analytic d<image_grad, image>/dv per image sample (masks frozen)."""
    lib = _load()
    p = _params_struct(params)
    tr = _f64(traces).reshape(p.n_traces, p.n_samples)
    sx, sy, rx, ry = (_f64(a) for a in (sx, sy, rx, ry))
    vel = _f64(vel)
    g = _f64(image_grad).reshape(p.n_il, p.n_xl, p.n_tau)
    out = np.zeros((p.n_il, p.n_xl, p.n_tau))
    rc = lib.oracle_vel_grad(ctypes.byref(p), _dp(tr), _dp(sx), _dp(sy), _dp(rx), _dp(ry), _dp(vel),
                             _dp(g), _dp(out))
    if rc != 0:
        raise RuntimeError(f"oracle_vel_grad failed: {rc}")
    return out


def leg_time_vti(h2, z, v, eta) -> float:
    return _load().oracle_leg_time_vti(float(h2), float(z), float(v), float(eta))


def pair(params, dxs, dys, dxr, dyr, tau, v, eta=0.0):
    """One (trace, image sample) pair: (passes, weight, sample position, L)."""
    lib = _load()
    p = _params_struct(params)
    w = ctypes.c_double()
    u = ctypes.c_double()
    L = ctypes.c_double()
    ok = lib.oracle_pair(ctypes.byref(p), dxs, dys, dxr, dyr, tau, v, eta,
                         ctypes.byref(w), ctypes.byref(u), ctypes.byref(L))
    return bool(ok), w.value, u.value, L.value
