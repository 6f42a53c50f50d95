"""Kernel backend: the single HIP backend of the migration path.

The reference dispatches between a Python and a Metal/C++ backend with a
silent fallback (processors/kernel_backend.py:31-35, :207-222).  Here there is
exactly one backend, HIP on gfx950: ``libpstm_hip.so`` is loaded once with
ctypes and a missing or unloadable library is an error, never a fallback.
"""
from __future__ import annotations

import ctypes
import threading
from enum import Enum
from pathlib import Path

from ..abi_types import ABI_VERSION, PstmParams

LIB_PATH = Path(__file__).resolve().parent.parent / "libpstm_hip.so"

# Every symbol include/pstm_abi.h declares: name -> (restype, argtypes)
_P = ctypes.POINTER(PstmParams)
_V = ctypes.c_void_p
SYMBOLS = {
    "pstm_abi_version": (ctypes.c_int, []),
    "pstm_last_error": (ctypes.c_char_p, []),
    "pstm_validate": (ctypes.c_int, [_P]),
    "pstm_plan_bytes": (ctypes.c_size_t, [_P]),
    "pstm_plan_build": (ctypes.c_int, [_P, _V, _V, _V, ctypes.c_size_t, _V]),
    "pstm_plan_kept": (ctypes.c_int64, [_V, _V]),
    "pstm_quantize_traces": (ctypes.c_int, [_V, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                            _V, _V, _V]),
    "pstm_forward": (ctypes.c_int, [_P, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
    "pstm_plan_bands": (ctypes.c_int, [_P]),
    "pstm_forward_bands": (ctypes.c_int, [_P, _V, ctypes.c_int32, ctypes.c_int32, _V, _V, _V, _V, _V, _V, _V,
                                          _V]),
    "pstm_integrate_traces": (ctypes.c_int, [_P, _V, _V, _V, _V, _V]),
    "pstm_forward_integrated": (ctypes.c_int, [_P, _V, ctypes.c_int32, ctypes.c_int32, _V, _V, _V, _V, _V, _V,
                                               _V, _V]),
    "pstm_adjoint": (ctypes.c_int, [_P, _V, _V, _V, _V, _V, _V]),
    "pstm_vel_grad": (ctypes.c_int, [_P, _V, _V, _V, _V, _V, _V, _V]),
    "pstm_migrate_static": (ctypes.c_int, [_P, _V, _V, _V, _V, _V, _V, _V, _V, _V, _V]),
}


class KernelBackend(Enum):
    """Available backends.  Only HIP exists."""

    HIP = "hip"


class BackendUnavailable(RuntimeError):
    """libpstm_hip.so is missing or cannot be loaded."""


_lock = threading.Lock()
_lib = None


def get_effective_backend(preferred: KernelBackend | str | None = None) -> KernelBackend:
    """Mirror of the reference's selector: always HIP, validated."""
    if preferred is not None and KernelBackend(preferred) is not KernelBackend.HIP:
        raise ValueError(f"unsupported backend {preferred!r}; only 'hip' exists")
    return KernelBackend.HIP


def load_library(path: Path | str | None = None) -> ctypes.CDLL:
    """Load libpstm_hip.so once and bind every ABI symbol.  Raises
    BackendUnavailable if the library is missing, unloadable, lacks a symbol
    or has a different ABI version."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path is not None else LIB_PATH
        if not p.exists():
            raise BackendUnavailable(
                f"{p} not found: build it with `python -m seisproc_amd.build` "
                "(the migration path has no CPU fallback)")
        try:
            lib = ctypes.CDLL(str(p))
        except OSError as e:
            raise BackendUnavailable(f"cannot load {p}: {e}") from e
        for name, (res, args) in SYMBOLS.items():
            try:
                fn = getattr(lib, name)
            except AttributeError as e:
                raise BackendUnavailable(f"{p} does not export {name}") from e
            fn.restype = res
            fn.argtypes = args
        ver = lib.pstm_abi_version()
        if ver != ABI_VERSION:
            raise BackendUnavailable(f"{p} has ABI version {ver}, expected {ABI_VERSION}")
        if path is None:
            _lib = lib
        return lib


def last_error(lib=None) -> str:
    lib = lib or load_library()
    msg = lib.pstm_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str, lib=None) -> int:
    """Raise on a negative status from the C ABI (message from pstm_last_error)."""
    if rc < 0:
        raise RuntimeError(f"{what} failed (status {rc}): {last_error(lib)}")
    return rc
