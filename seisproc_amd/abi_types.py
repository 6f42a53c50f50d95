"""ctypes mirror of ``include/pstm_abi.h`` (``pstm_params_t`` and the enums).

Plain data only: importing this module loads no library and touches no GPU.
"""
from __future__ import annotations

import ctypes
from typing import Any, Dict

ABI_VERSION = 3

# pstm_status_t
PSTM_OK = 0
PSTM_ERR_INVALID = -1
PSTM_ERR_UNSUPPORTED = -2
PSTM_ERR_WORKSPACE = -3
PSTM_ERR_HIP = -4

# pstm_dtype_t
PSTM_F32, PSTM_BF16, PSTM_F16 = 0, 1, 2
DTYPE_CODES = {"f32": PSTM_F32, "float32": PSTM_F32, "bf16": PSTM_BF16,
               "bfloat16": PSTM_BF16, "f16": PSTM_F16, "float16": PSTM_F16}

# flags
FLAG_AA = 1 << 0
FLAG_VTI = 1 << 1
FLAG_KAHAN = 1 << 2
FLAG_TAPER = 1 << 3
FLAG_ZERO_OUTSIDE = 1 << 4
FLAG_TIME_APERTURE = 1 << 5

# weight modes
WEIGHT_FULL, WEIGHT_SPREADING, WEIGHT_NONE, WEIGHT_OBLIQUITY = 0, 1, 2, 3
WEIGHT_CODES = {"full": WEIGHT_FULL, "spreading": WEIGHT_SPREADING, "none": WEIGHT_NONE,
                "obliquity": WEIGHT_OBLIQUITY}

# velocity layouts
VEL_1D, VEL_3D = 0, 1


class PstmParams(ctypes.Structure):
    """Layout-identical to ``pstm_params_t``."""

    _fields_ = [
        ("n_traces", ctypes.c_int64),
        ("n_samples", ctypes.c_int32),
        ("n_il", ctypes.c_int32),
        ("n_xl", ctypes.c_int32),
        ("n_tau", ctypes.c_int32),
        ("dt", ctypes.c_double),
        ("t0", ctypes.c_double),
        ("dtau", ctypes.c_double),
        ("tau0", ctypes.c_double),
        ("origin_x", ctypes.c_double),
        ("origin_y", ctypes.c_double),
        ("il_dx", ctypes.c_double),
        ("il_dy", ctypes.c_double),
        ("xl_dx", ctypes.c_double),
        ("xl_dy", ctypes.c_double),
        ("aperture_m", ctypes.c_double),
        ("max_angle_deg", ctypes.c_double),
        ("taper_width", ctypes.c_double),
        ("aa_scale", ctypes.c_double),
        ("aa_max", ctypes.c_double),
        ("dtype", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("weight_mode", ctypes.c_int32),
        ("vel_mode", ctypes.c_int32),
        ("min_offset_m", ctypes.c_double),
        ("max_offset_m", ctypes.c_double),
    ]

    DEFAULTS: Dict[str, Any] = dict(
        n_traces=0, n_samples=0, n_il=1, n_xl=1, n_tau=1, dt=0.002, t0=0.0, dtau=0.002,
        tau0=0.0, origin_x=0.0, origin_y=0.0, il_dx=25.0, il_dy=0.0, xl_dx=0.0, xl_dy=25.0,
        aperture_m=5000.0, max_angle_deg=60.0, taper_width=0.1, aa_scale=1.0, aa_max=16.0,
        dtype=PSTM_F32, flags=0, weight_mode=WEIGHT_FULL, vel_mode=VEL_1D, min_offset_m=0.0,
        max_offset_m=0.0)

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "PstmParams":
        vals = dict(cls.DEFAULTS)
        unknown = set(d) - set(vals)
        if unknown:
            raise ValueError(f"unknown pstm parameter(s): {sorted(unknown)}")
        vals.update(d)
        return cls(**vals)

    def to_dict(self) -> Dict[str, Any]:
        return {name: getattr(self, name) for name, _ in self._fields_}

    def replace(self, **kw) -> "PstmParams":
        d = self.to_dict()
        d.update(kw)
        return PstmParams.from_dict(d)
