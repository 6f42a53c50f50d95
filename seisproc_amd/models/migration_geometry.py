"""Survey geometry (reference models/migration_geometry.py:19-551).

Same fields and derived attributes as the reference, except that the
coordinates stay FP64: the reference casts them to FP32 on construction
(:55-58), which at UTM magnitudes (~6.7e6 m) quantises positions to 0.5 m.
The HIP plan reduces FP64 world coordinates to FP32 image-local ones on the
GPU, once per trace.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Dict, Optional, Tuple

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

    # -- per-trace positions
    def get_source_position(self, idx: int) -> Tuple[float, float]:
        return float(self.source_x[idx]), float(self.source_y[idx])

    def get_receiver_position(self, idx: int) -> Tuple[float, float]:
        return float(self.receiver_x[idx]), float(self.receiver_y[idx])

    def get_cdp_position(self, idx: int) -> Tuple[float, float]:
        return float(self.cdp_x[idx]), float(self.cdp_y[idx])

    # -- extents
    @staticmethod
    def _span(a: np.ndarray) -> Tuple[float, float]:
        return float(np.min(a)), float(np.max(a))

    def get_survey_bounds(self) -> Dict[str, Tuple[float, float]]:
        """(min, max) of x and y over sources and receivers, of offset and of azimuth."""
        b = self.get_bounds()
        return {"x": (b["x_min"], b["x_max"]), "y": (b["y_min"], b["y_max"]),
                "offset": self._span(self.offset), "azimuth": self._span(self.azimuth)}

    def get_offset_range(self) -> Tuple[float, float]:
        return self._span(self.offset)

    def get_azimuth_range(self) -> Tuple[float, float]:
        return self._span(self.azimuth)

    # -- offset / azimuth selections (inclusive bounds; an azimuth range with
    # min > max wraps through north, e.g. 350..10)
    def _bin_mask(self, min_offset=None, max_offset=None, min_azimuth=None, max_azimuth=None) -> np.ndarray:
        keep = np.ones(self.n_traces, dtype=bool)
        if min_offset is not None:
            off = self.offset
            keep &= (off >= min_offset) & (off <= max_offset)
        if min_azimuth is not None:
            az = self.azimuth
            if min_azimuth <= max_azimuth:
                keep &= (az >= min_azimuth) & (az <= max_azimuth)
            else:
                keep &= (az >= min_azimuth) | (az <= max_azimuth)
        return keep

    def filter_by_offset(self, min_offset: float, max_offset: float) -> "MigrationGeometry":
        return self.select(self._bin_mask(min_offset, max_offset))

    def filter_by_azimuth(self, min_azimuth: float, max_azimuth: float) -> "MigrationGeometry":
        return self.select(self._bin_mask(min_azimuth=min_azimuth, max_azimuth=max_azimuth))

    def filter_by_offset_azimuth(self, min_offset: float, max_offset: float, min_azimuth: float,
                                 max_azimuth: float) -> "MigrationGeometry":
        return self.select(self._bin_mask(min_offset, max_offset, min_azimuth, max_azimuth))

    def get_indices_for_bin(self, min_offset: float, max_offset: float, min_azimuth: float,
                            max_azimuth: float) -> np.ndarray:
        return np.nonzero(self._bin_mask(min_offset, max_offset, min_azimuth, max_azimuth))[0]

    # -- stations
    @staticmethod
    def _unique_xy(x: np.ndarray, y: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        if not len(x):
            return x[:0], y[:0]
        u = np.unique(np.stack([x, y], 1), axis=0)
        return u[:, 0], u[:, 1]

    def get_unique_sources(self) -> Tuple[np.ndarray, np.ndarray]:
        return self._unique_xy(self.source_x, self.source_y)

    def get_unique_receivers(self) -> Tuple[np.ndarray, np.ndarray]:
        return self._unique_xy(self.receiver_x, self.receiver_y)

    def get_statistics(self) -> Dict[str, Any]:
        """Trace and station counts, extents and offset moments (the
        reference's keys, :308-329)."""
        if not self.n_traces:
            return {"n_traces": 0, "n_unique_sources": 0, "n_unique_receivers": 0, "offset_range": (0.0, 0.0)}
        b = self.get_survey_bounds()
        off = self.offset
        return {"n_traces": self.n_traces, "n_unique_sources": len(self.get_unique_sources()[0]),
                "n_unique_receivers": len(self.get_unique_receivers()[0]), "x_range": b["x"], "y_range": b["y"],
                "offset_range": b["offset"], "offset_mean": float(off.mean()), "offset_std": float(off.std()),
                "azimuth_range": b["azimuth"]}

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


def create_synthetic_geometry(n_shots: int = 10, n_receivers_per_shot: int = 100, shot_spacing: float = 100.0,
                              receiver_spacing: float = 25.0, near_offset: float = 100.0,
                              shot_line_azimuth: float = 0.0) -> MigrationGeometry:
    """Streamer-style line (reference :409-471): shots every shot_spacing
    along the azimuth from the origin, each recorded by a receiver line
    trailing behind it from near_offset every receiver_spacing."""
    u = np.array([np.sin(np.radians(shot_line_azimuth)), np.cos(np.radians(shot_line_azimuth))])
    shot = np.arange(n_shots, dtype=np.float64)[:, None] * shot_spacing
    trail = near_offset + np.arange(n_receivers_per_shot, dtype=np.float64)[None, :] * receiver_spacing
    s = np.broadcast_to(shot, (n_shots, n_receivers_per_shot)).ravel()
    r = (shot - trail).ravel()
    return MigrationGeometry(s * u[0], s * u[1], r * u[0], r * u[1],
                             metadata={"synthetic": True, "n_shots": n_shots,
                                       "n_receivers_per_shot": n_receivers_per_shot,
                                       "shot_spacing": shot_spacing, "receiver_spacing": receiver_spacing})


def create_land_3d_geometry(n_source_lines: int = 5, n_sources_per_line: int = 20, n_receiver_lines: int = 10,
                            n_receivers_per_line: int = 50, source_line_spacing: float = 400.0,
                            source_spacing: float = 50.0, receiver_line_spacing: float = 200.0,
                            receiver_spacing: float = 25.0) -> MigrationGeometry:
    """Orthogonal land patch (reference :474-540): source lines along x
    (spaced in y), receiver lines along y (spaced in x), every source
    recorded by every receiver (source-major trace order)."""
    sx, sy = np.meshgrid(np.arange(n_sources_per_line) * source_spacing,
                         np.arange(n_source_lines) * source_line_spacing)
    rx, ry = np.meshgrid(np.arange(n_receiver_lines) * receiver_line_spacing,
                         np.arange(n_receivers_per_line) * receiver_spacing, indexing="ij")
    ns, nr = sx.size, rx.size
    return MigrationGeometry(np.repeat(sx.ravel(), nr), np.repeat(sy.ravel(), nr), np.tile(rx.ravel(), ns),
                             np.tile(ry.ravel(), ns),
                             metadata={"synthetic": True, "geometry_type": "land_3d_orthogonal",
                                       "n_source_lines": n_source_lines, "n_receiver_lines": n_receiver_lines})
