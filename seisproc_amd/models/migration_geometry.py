"""Survey geometry (reference models/migration_geometry.py:19-402).

Same fields and derived attributes as the reference, except that the
coordinates stay FP64: the reference casts them to FP32 on construction
(:55-58), which at UTM magnitudes (~6.7e6 m) quantises positions to 0.5 m.
The HIP plan reduces FP64 world coordinates to FP32 image-local ones on the
GPU, once per trace.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Dict, Optional

import numpy as np


@dataclass
class MigrationGeometry:
    source_x: np.ndarray
    source_y: np.ndarray
    receiver_x: np.ndarray
    receiver_y: np.ndarray
    source_z: Optional[np.ndarray] = None
    receiver_z: Optional[np.ndarray] = None
    trace_indices: Optional[np.ndarray] = None
    metadata: Dict[str, Any] = field(default_factory=dict)

    def __post_init__(self):
        for k in ("source_x", "source_y", "receiver_x", "receiver_y"):
            setattr(self, k, np.ascontiguousarray(np.asarray(getattr(self, k), dtype=np.float64)))
        n = len(self.source_x)
        for k in ("source_y", "receiver_x", "receiver_y"):
            if len(getattr(self, k)) != n:
                raise ValueError(f"{k} length {len(getattr(self, k))} != source_x length {n}")
        for k in ("source_z", "receiver_z"):
            v = getattr(self, k)
            if v is not None:
                v = np.asarray(v, dtype=np.float64)
                if len(v) != n:
                    raise ValueError(f"{k} length {len(v)} != n_traces {n}")
                setattr(self, k, v)
        if self.trace_indices is not None:
            self.trace_indices = np.asarray(self.trace_indices, dtype=np.int64)

    @property
    def n_traces(self) -> int:
        return len(self.source_x)

    @property
    def offset(self) -> np.ndarray:
        return np.hypot(self.receiver_x - self.source_x, self.receiver_y - self.source_y)

    @property
    def azimuth(self) -> np.ndarray:
        """Source-to-receiver azimuth, degrees clockwise from north in [0, 360)."""
        az = np.degrees(np.arctan2(self.receiver_x - self.source_x, self.receiver_y - self.source_y))
        return (az + 360.0) % 360.0

    @property
    def cdp_x(self) -> np.ndarray:
        return 0.5 * (self.source_x + self.receiver_x)

    @property
    def cdp_y(self) -> np.ndarray:
        return 0.5 * (self.source_y + self.receiver_y)

    def src_xy(self) -> np.ndarray:
        return np.stack([self.source_x, self.source_y], 1)

    def rcv_xy(self) -> np.ndarray:
        return np.stack([self.receiver_x, self.receiver_y], 1)

    def select(self, mask) -> "MigrationGeometry":
        idx = np.nonzero(np.asarray(mask))[0] if np.asarray(mask).dtype == bool else np.asarray(mask)
        ti = self.trace_indices[idx] if self.trace_indices is not None else idx
        return MigrationGeometry(self.source_x[idx], self.source_y[idx], self.receiver_x[idx],
                                 self.receiver_y[idx],
                                 self.source_z[idx] if self.source_z is not None else None,
                                 self.receiver_z[idx] if self.receiver_z is not None else None,
                                 ti, dict(self.metadata))

    def get_bounds(self) -> Dict[str, float]:
        xs = np.concatenate([self.source_x, self.receiver_x])
        ys = np.concatenate([self.source_y, self.receiver_y])
        return {"x_min": float(xs.min()), "x_max": float(xs.max()), "y_min": float(ys.min()),
                "y_max": float(ys.max())}

    def get_statistics(self) -> Dict[str, Any]:
        off = self.offset
        return {"n_traces": self.n_traces,
                "n_unique_sources": int(len(np.unique(np.stack([self.source_x, self.source_y], 1), axis=0)))
                if self.n_traces else 0,
                "offset_range": (float(off.min()), float(off.max())) if self.n_traces else (0.0, 0.0),
                "bounds": self.get_bounds() if self.n_traces else None}

    def to_dict(self) -> Dict[str, Any]:
        d = {"source_x": self.source_x.tolist(), "source_y": self.source_y.tolist(),
             "receiver_x": self.receiver_x.tolist(), "receiver_y": self.receiver_y.tolist(),
             "metadata": dict(self.metadata)}
        for k in ("source_z", "receiver_z", "trace_indices"):
            if getattr(self, k) is not None:
                d[k] = getattr(self, k).tolist()
        return d

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "MigrationGeometry":
        g = lambda k: np.asarray(d[k]) if k in d else None
        return cls(g("source_x"), g("source_y"), g("receiver_x"), g("receiver_y"), g("source_z"),
                   g("receiver_z"), g("trace_indices"), d.get("metadata", {}))

    @classmethod
    def from_headers(cls, headers: Dict[str, np.ndarray], sx_key: str = "SourceX", sy_key: str = "SourceY",
                     gx_key: str = "GroupX", gy_key: str = "GroupY",
                     coordinate_scalar: float = 1.0) -> "MigrationGeometry":
        """Geometry from SEG-Y style headers (reference :364-394), kept FP64."""
        c = float(coordinate_scalar)
        f = lambda k: np.asarray(headers[k], dtype=np.float64) * c
        return cls(f(sx_key), f(sy_key), f(gx_key), f(gy_key), metadata={"coordinate_scalar": coordinate_scalar})

    def __repr__(self) -> str:
        st = self.get_statistics()
        lo, hi = st["offset_range"]
        return f"<MigrationGeometry {self.n_traces} traces from {st['n_unique_sources']} sources, offsets {lo:g}..{hi:g} m>"
