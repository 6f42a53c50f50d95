"""Named binning tables of a migration job (reference utils/binning_presets.py:
24-193), looked up by ``MigrationJobConfig.binning_preset``.

The preset names, their offset ranges and sector counts are the reference's,
so a job configured by preset name yields the same bins (and per-bin output
folders) here as there.
"""
from __future__ import annotations

from typing import Callable, Dict, List, Tuple

import numpy as np

from ..models.binning import (BinningTable, create_common_offset_binning, create_full_stack_binning,
                              create_narrow_azimuth_binning, create_ovt_binning)

LAND_3D_OFFSET_RANGES: List[Tuple[float, float]] = [
    (0, 200), (200, 400), (400, 600), (600, 800), (800, 1000),
    (1000, 1500), (1500, 2000), (2000, 2500), (2500, 3000), (3000, 5000)]
MARINE_OFFSET_RANGES: List[Tuple[float, float]] = [
    (0, 500), (500, 1000), (1000, 2000), (2000, 3000), (3000, 4500), (4500, 8000)]
WIDE_AZIMUTH_OFFSET_RANGES: List[Tuple[float, float]] = [(0, 1000), (1000, 2000), (2000, 3500), (3500, 6000)]


def _named(table: BinningTable, name: str) -> BinningTable:
    table.name = name
    return table


def get_land_3d_preset() -> BinningTable:
    return _named(create_common_offset_binning(LAND_3D_OFFSET_RANGES, name_prefix="land3d"),
                  "Land 3D - 10 Offset Bins")


def get_marine_preset() -> BinningTable:
    return _named(create_common_offset_binning(MARINE_OFFSET_RANGES, name_prefix="marine"),
                  "Marine - 6 Offset Bins")


def get_wide_azimuth_ovt_preset() -> BinningTable:
    return _named(create_ovt_binning(WIDE_AZIMUTH_OFFSET_RANGES, n_azimuth_sectors=4), "Wide Azimuth OVT - 16 Bins")


def get_narrow_azimuth_preset(inline_azimuth: float = 0.0) -> BinningTable:
    return _named(create_narrow_azimuth_binning(0.0, 8000.0, inline_azimuth=inline_azimuth, azimuth_width=30.0),
                  "Narrow Azimuth - 4 Sectors")


def get_full_stack_preset(offset_max: float = 10000.0) -> BinningTable:
    return _named(create_full_stack_binning(offset_max), "Full Stack - Single Bin")


# name -> (factory, description)
_PRESETS: Dict[str, Tuple[Callable[..., BinningTable], str]] = {
    "land_3d": (get_land_3d_preset, "10 offset bins for land 3D surveys (0-5000m)"),
    "marine": (get_marine_preset, "6 offset bins for marine streamer data (0-8000m)"),
    "wide_azimuth_ovt": (get_wide_azimuth_ovt_preset, "16 OVT bins (4 offset x 4 azimuth) for wide-azimuth"),
    "narrow_azimuth": (get_narrow_azimuth_preset, "4 bins separating inline/crossline directions"),
    "full_stack": (get_full_stack_preset, "Single bin for full stack migration"),
}


def list_presets() -> List[str]:
    return list(_PRESETS)


def get_preset(name: str, **kwargs) -> BinningTable:
    """A fresh table of the named preset (keyword arguments go to its
    factory: ``inline_azimuth`` for narrow_azimuth, ``offset_max`` for
    full_stack).  Unknown names raise ValueError listing the known ones."""
    if name not in _PRESETS:
        raise ValueError(f"Unknown preset '{name}'. Available: {', '.join(_PRESETS)}")
    return _PRESETS[name][0](**kwargs)


def get_preset_description(name: str) -> str:
    return _PRESETS[name][1] if name in _PRESETS else "No description available"


def create_custom_offset_binning(offset_min: float, offset_max: float, n_bins: int,
                                 logarithmic: bool = False) -> BinningTable:
    """n_bins full-azimuth offset bins named ``custom_*``: equal widths, or
    with logarithmic=True geometric edges from max(1, offset_min) to
    offset_max (narrower near offsets), the first edge reset to offset_min."""
    if logarithmic:
        lo = max(1.0, offset_min)
        edges = [lo * (offset_max / lo) ** (i / n_bins) for i in range(n_bins + 1)]
        edges[0] = offset_min
    else:
        w = (offset_max - offset_min) / n_bins
        edges = [offset_min + i * w for i in range(n_bins + 1)]
    return create_common_offset_binning(list(zip(edges[:-1], edges[1:])), name_prefix="custom")


def suggest_binning(offsets, azimuths, target_traces_per_bin: int = 10000) -> BinningTable:
    """Full-azimuth offset bins holding about target_traces_per_bin traces
    each: max(1, n // target) bins with edges at equally spaced percentiles
    of the offsets (the azimuths do not shape the suggestion)."""
    off = np.asarray(offsets, dtype=np.float64)
    n_bins = max(1, off.size // target_traces_per_bin)
    edges = np.percentile(off, np.linspace(0.0, 100.0, n_bins + 1))
    table = create_common_offset_binning([(float(a), float(b)) for a, b in zip(edges[:-1], edges[1:])],
                                         name_prefix="suggested")
    table.name = f"Suggested - {n_bins} Bins"
    table.metadata.update(source="auto_suggested", target_traces_per_bin=target_traces_per_bin)
    return table
