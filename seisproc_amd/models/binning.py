"""Offset-azimuth bins of a prestack migration job.

Same classes, fields, membership rules and JSON form as models/binning.py of
the reference (``OffsetAzimuthBin`` :34-190, ``BinningTable`` :194-497 and the
common-offset / uniform / OVT / narrow-azimuth / full-stack factories
:502-700, whose bin names name the per-bin output folders): a trace belongs to
a bin when offset_min <= offset < offset_max and its azimuth lies in
[azimuth_min, azimuth_max), the azimuth range wrapping through north when
azimuth_max < azimuth_min.  The job driver migrates the enabled bins one at
a time.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from enum import Enum
from typing import Any, Dict, Iterator, List, Optional, Tuple

import numpy as np


class BinningPreset(Enum):
    COMMON_OFFSET = "common_offset"
    OVT = "ovt"
    NARROW_AZIMUTH = "narrow_azimuth"
    FULL_STACK = "full_stack"
    CUSTOM = "custom"


@dataclass
class OffsetAzimuthBin:
    name: str
    offset_min: float
    offset_max: float
    azimuth_min: float = 0.0
    azimuth_max: float = 360.0
    enabled: bool = True
    metadata: Dict[str, Any] = field(default_factory=dict)

    def __post_init__(self):
        if not self.name:
            raise ValueError("Bin name cannot be empty")
        if self.offset_min < 0:
            raise ValueError(f"offset_min must be non-negative, got {self.offset_min}")
        if self.offset_max <= self.offset_min:
            raise ValueError(f"offset_max ({self.offset_max}) must be greater than offset_min ({self.offset_min})")
        if not 0 <= self.azimuth_min < 360:
            raise ValueError(f"azimuth_min must be in [0, 360), got {self.azimuth_min}")
        if not 0 < self.azimuth_max <= 360:
            raise ValueError(f"azimuth_max must be in (0, 360], got {self.azimuth_max}")

    @property
    def wraps(self) -> bool:
        return self.azimuth_max < self.azimuth_min

    @property
    def offset_center(self) -> float:
        return 0.5 * (self.offset_min + self.offset_max)

    @property
    def offset_width(self) -> float:
        return self.offset_max - self.offset_min

    @property
    def azimuth_width(self) -> float:
        return (360.0 - self.azimuth_min + self.azimuth_max) if self.wraps else self.azimuth_max - self.azimuth_min

    @property
    def azimuth_center(self) -> float:
        c = self.azimuth_min + 0.5 * self.azimuth_width
        return c - 360.0 if c >= 360.0 else c

    @property
    def is_full_azimuth(self) -> bool:
        return self.azimuth_width >= 359.9

    def contains_batch(self, offsets: np.ndarray, azimuths: np.ndarray) -> np.ndarray:
        offsets = np.asarray(offsets)
        azimuths = np.asarray(azimuths)
        in_off = (offsets >= self.offset_min) & (offsets < self.offset_max)
        if self.wraps:
            in_az = (azimuths >= self.azimuth_min) | (azimuths < self.azimuth_max)
        else:
            in_az = (azimuths >= self.azimuth_min) & (azimuths < self.azimuth_max)
        return in_off & in_az

    def contains(self, offset: float, azimuth: float) -> bool:
        return bool(self.contains_batch(np.array([offset]), np.array([azimuth]))[0])

    def to_dict(self) -> Dict[str, Any]:
        return {"name": self.name, "offset_min": self.offset_min, "offset_max": self.offset_max,
                "azimuth_min": self.azimuth_min, "azimuth_max": self.azimuth_max, "enabled": self.enabled,
                "metadata": dict(self.metadata)}

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "OffsetAzimuthBin":
        return cls(name=d["name"], offset_min=d["offset_min"], offset_max=d["offset_max"],
                   azimuth_min=d.get("azimuth_min", 0.0), azimuth_max=d.get("azimuth_max", 360.0),
                   enabled=d.get("enabled", True), metadata=dict(d.get("metadata", {})))


@dataclass
class BinningTable:
    name: str = "Custom Binning"
    bins: List[OffsetAzimuthBin] = field(default_factory=list)
    preset: BinningPreset = BinningPreset.CUSTOM
    metadata: Dict[str, Any] = field(default_factory=dict)

    def __post_init__(self):
        names = [b.name for b in self.bins]
        if len(names) != len(set(names)):
            raise ValueError("Bin names must be unique")

    @property
    def n_bins(self) -> int:
        return len(self.bins)

    @property
    def enabled_bins(self) -> List[OffsetAzimuthBin]:
        return [b for b in self.bins if b.enabled]

    @property
    def n_enabled_bins(self) -> int:
        return len(self.enabled_bins)

    @property
    def offset_range(self) -> Tuple[float, float]:
        if not self.bins:
            return (0.0, 0.0)
        return (min(b.offset_min for b in self.bins), max(b.offset_max for b in self.bins))

    def add_bin(self, bin_def: OffsetAzimuthBin) -> None:
        if self.get_bin(bin_def.name) is not None:
            raise ValueError(f"Bin name '{bin_def.name}' already exists")
        self.bins.append(bin_def)

    def remove_bin(self, name: str) -> bool:
        n = len(self.bins)
        self.bins = [b for b in self.bins if b.name != name]
        return len(self.bins) < n

    def get_bin(self, name: str) -> Optional[OffsetAzimuthBin]:
        return next((b for b in self.bins if b.name == name), None)

    def assign_trace(self, offset: float, azimuth: float, enabled_only: bool = True) -> List[str]:
        return [b.name for b in (self.enabled_bins if enabled_only else self.bins) if b.contains(offset, azimuth)]

    def assign_traces_batch(self, offsets: np.ndarray, azimuths: np.ndarray,
                            enabled_only: bool = True) -> Dict[str, np.ndarray]:
        return {b.name: b.contains_batch(offsets, azimuths) for b in (self.enabled_bins if enabled_only else self.bins)}

    # -- table checks (reference :327-438)
    def check_coverage(self, offsets: np.ndarray, azimuths: np.ndarray) -> Dict[str, Any]:
        """Traces per enabled bin, traces in no bin and in several bins."""
        masks = self.assign_traces_batch(offsets, azimuths)
        n = len(offsets)
        hits = np.zeros(n, dtype=np.int64)
        for m in masks.values():
            hits += m
        unassigned = int(np.count_nonzero(hits == 0))
        return {"n_traces": n, "n_unassigned": unassigned, "n_multi_bin": int(np.count_nonzero(hits > 1)),
                "bin_counts": {k: int(m.sum()) for k, m in masks.items()},
                "coverage_percent": 100.0 * (n - unassigned) / n if n else 0.0}

    @staticmethod
    def _arcs(b: OffsetAzimuthBin) -> List[Tuple[float, float]]:
        return [(b.azimuth_min, 360.0), (0.0, b.azimuth_max)] if b.wraps else [(b.azimuth_min, b.azimuth_max)]

    def check_overlaps(self) -> List[Tuple[str, str]]:
        """Pairs of bins whose [offset) ranges and azimuth sectors (wrapping
        through north where they do) intersect in an interval."""
        out = []
        for i, a in enumerate(self.bins):
            for b in self.bins[i + 1:]:
                if max(a.offset_min, b.offset_min) >= min(a.offset_max, b.offset_max):
                    continue
                if any(max(p, r) < min(q, s) for p, q in self._arcs(a) for r, s in self._arcs(b)):
                    out.append((a.name, b.name))
        return out

    def check_gaps(self, offset_step: float = 100.0) -> List[Tuple[float, float]]:
        """Offset intervals inside the table's offset range that no bin
        covers.  The intervals are exact; offset_step (the reference's scan
        resolution) is accepted and not needed."""
        if not self.bins:
            return []
        gaps, reach = [], None
        for lo, hi in sorted((b.offset_min, b.offset_max) for b in self.bins):
            if reach is not None and lo > reach:
                gaps.append((reach, lo))
            reach = hi if reach is None else max(reach, hi)
        return gaps

    def __iter__(self) -> Iterator[OffsetAzimuthBin]:
        return iter(self.bins)

    def __len__(self) -> int:
        return len(self.bins)

    def __getitem__(self, idx: int) -> OffsetAzimuthBin:
        return self.bins[idx]

    def to_dict(self) -> Dict[str, Any]:
        return {"name": self.name, "preset": self.preset.value, "bins": [b.to_dict() for b in self.bins],
                "metadata": dict(self.metadata)}

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "BinningTable":
        return cls(name=d.get("name", "Custom Binning"), bins=[OffsetAzimuthBin.from_dict(b) for b in d.get("bins", [])],
                   preset=BinningPreset(d.get("preset", "custom")), metadata=dict(d.get("metadata", {})))

    def to_json(self, indent: int = 2) -> str:
        return json.dumps(self.to_dict(), indent=indent)

    @classmethod
    def from_json(cls, s: str) -> "BinningTable":
        return cls.from_dict(json.loads(s))

    def save(self, filepath: str) -> None:
        with open(filepath, "w") as fh:
            fh.write(self.to_json())

    @classmethod
    def load(cls, filepath: str) -> "BinningTable":
        with open(filepath) as fh:
            return cls.from_json(fh.read())


def create_common_offset_binning(offset_ranges: List[Tuple[float, float]], name_prefix: str = "offset") -> BinningTable:
    """Full-azimuth bins over the given [min, max) offset ranges, named
    ``{prefix}_{min:04d}_{max:04d}``."""
    bins = [OffsetAzimuthBin(name=f"{name_prefix}_{int(lo):04d}_{int(hi):04d}", offset_min=lo, offset_max=hi)
            for lo, hi in offset_ranges]
    return BinningTable(name="Common Offset Binning", bins=bins, preset=BinningPreset.COMMON_OFFSET)


def create_uniform_offset_binning(offset_min: float = 0.0, offset_max: float = 5000.0, n_bins: int = 10) -> BinningTable:
    w = (offset_max - offset_min) / n_bins
    return create_common_offset_binning([(offset_min + i * w, offset_min + (i + 1) * w) for i in range(n_bins)])


def create_full_stack_binning(offset_max: float = 10000.0) -> BinningTable:
    return BinningTable(name="Full Stack", bins=[OffsetAzimuthBin("full_stack", 0.0, offset_max)],
                        preset=BinningPreset.FULL_STACK)


def create_ovt_binning(offset_ranges: List[Tuple[float, float]], n_azimuth_sectors: int = 4) -> BinningTable:
    """Offset-vector tiles: every offset range split into n equal azimuth
    sectors starting at north, named ``ovt_off{min:04d}_az{sector start:03d}``."""
    step = 360.0 / n_azimuth_sectors
    bins = [OffsetAzimuthBin(name=f"ovt_off{int(lo):04d}_az{int(k * step):03d}", offset_min=lo, offset_max=hi,
                             azimuth_min=k * step, azimuth_max=(k + 1) * step)
            for lo, hi in offset_ranges for k in range(n_azimuth_sectors)]
    return BinningTable(name="OVT Binning", bins=bins, preset=BinningPreset.OVT,
                        metadata={"n_azimuth_sectors": n_azimuth_sectors})


def create_narrow_azimuth_binning(offset_min: float = 0.0, offset_max: float = 5000.0, inline_azimuth: float = 0.0,
                                  azimuth_width: float = 20.0) -> BinningTable:
    """Four sectors of azimuth_width centred on the inline direction, its
    opposite and the two crossline directions (``inline_pos``, ``inline_neg``,
    ``xline_pos``, ``xline_neg``); sectors across north wrap, and a sector
    ending at north ends at 360 (the bin's upper bound must be positive)."""
    az0 = inline_azimuth % 360.0
    h = 0.5 * azimuth_width
    bins = [OffsetAzimuthBin(name=name, offset_min=offset_min, offset_max=offset_max,
                             azimuth_min=(az0 + turn - h) % 360.0, azimuth_max=(az0 + turn + h) % 360.0 or 360.0)
            for name, turn in (("inline_pos", 0.0), ("inline_neg", 180.0), ("xline_pos", 90.0),
                               ("xline_neg", 270.0))]
    return BinningTable(name="Narrow Azimuth Binning", bins=bins, preset=BinningPreset.NARROW_AZIMUTH,
                        metadata={"inline_azimuth": az0, "azimuth_width": azimuth_width})
