"""Migration job configuration (reference models/migration_job.py:27-491).

Same fields, defaults, validation messages, JSON save/load, derived objects
(get_output_grid / get_migration_config / get_velocity_model /
get_binning_table) and job templates.  The binning table is a BinningTable
(saved as its dict form); a preset name resolves to the preset's table (the
job driver's trace binning; the HIP plan bins traces on the GPU itself).
get_velocity_model
loads velocity_file through VelocityModel.load (the reference calls a load
its VelocityModel does not define, :235).
"""
from __future__ import annotations

import json
import logging
from dataclasses import dataclass, field
from datetime import datetime
from pathlib import Path
from typing import Any, Dict, List, Optional

from .binning import BinningTable
from .migration_config import InterpolationMode, MigrationConfig, OutputGrid, TraveltimeMode, WeightMode
from .velocity_model import VelocityModel

logger = logging.getLogger(__name__)

_HM_FIELDS = ("source_x", "source_y", "receiver_x", "receiver_y", "offset", "azimuth", "inline", "xline",
              "cdp_x", "cdp_y")


@dataclass
class HeaderMapping:
    source_x: Optional[str] = None
    source_y: Optional[str] = None
    receiver_x: Optional[str] = None
    receiver_y: Optional[str] = None
    offset: Optional[str] = None
    azimuth: Optional[str] = None
    inline: Optional[str] = None
    xline: Optional[str] = None
    cdp_x: Optional[str] = None
    cdp_y: Optional[str] = None

    def to_dict(self) -> Dict[str, str]:
        return {k: getattr(self, k) for k in _HM_FIELDS if getattr(self, k) is not None}

    @classmethod
    def from_dict(cls, d: Dict[str, str]) -> "HeaderMapping":
        return cls(**{k: v for k, v in d.items() if k in _HM_FIELDS})


@dataclass
class MigrationJobConfig:
    name: str
    input_file: str
    output_directory: str
    velocity_v0: float = 2500.0
    velocity_gradient: float = 0.0
    velocity_file: Optional[str] = None
    time_min_ms: float = 0.0
    time_max_ms: float = 4000.0
    dt_ms: float = 4.0
    inline_min: int = 1
    inline_max: int = 100
    inline_step: int = 1
    xline_min: int = 1
    xline_max: int = 100
    xline_step: int = 1
    origin_x: float = 0.0
    origin_y: float = 0.0
    inline_spacing: float = 25.0
    xline_spacing: float = 25.0
    grid_azimuth_deg: float = 0.0
    max_aperture_m: float = 3000.0
    max_angle_deg: float = 60.0
    max_offset_m: float = 10000.0
    taper_width_samples: int = 20
    near_offset_taper_m: float = 0.0
    anti_alias_filter: bool = True
    binning_preset: Optional[str] = None
    binning_table: Optional[BinningTable] = None
    header_mapping: HeaderMapping = field(default_factory=HeaderMapping)
    processing_mode: str = "sequential"
    max_workers: int = 1
    memory_limit_gb: float = 8.0
    enable_checkpointing: bool = True
    output_format: str = "segy"
    create_stack_volume: bool = True
    create_fold_volume: bool = False
    compress_output: bool = False
    created_at: Optional[str] = None
    description: str = ""
    metadata: Dict[str, Any] = field(default_factory=dict)

    def __post_init__(self):
        if self.created_at is None:
            self.created_at = datetime.now().isoformat()

    def validate(self) -> List[str]:
        errors = []
        if not Path(self.input_file).exists():
            errors.append(f"Input file does not exist: {self.input_file}")
        out = Path(self.output_directory)
        if not out.exists():
            try:
                out.mkdir(parents=True, exist_ok=True)
            except Exception as e:  # noqa: BLE001 - reported as a validation message
                errors.append(f"Cannot create output directory: {e}")
        if self.velocity_v0 <= 0:
            errors.append(f"Invalid velocity v0: {self.velocity_v0}")
        if self.time_max_ms <= self.time_min_ms:
            errors.append("time_max_ms must be greater than time_min_ms")
        if self.dt_ms <= 0:
            errors.append("dt_ms must be positive")
        if self.inline_max < self.inline_min:
            errors.append("inline_max must be >= inline_min")
        if self.xline_max < self.xline_min:
            errors.append("xline_max must be >= xline_min")
        if self.max_aperture_m <= 0:
            errors.append("max_aperture_m must be positive")
        if not 0 < self.max_angle_deg <= 90:
            errors.append("max_angle_deg must be between 0 and 90")
        if self.binning_preset is None and self.binning_table is None:
            errors.append("Either binning_preset or binning_table must be specified")
        if self.processing_mode not in ("sequential", "parallel"):
            errors.append(f"Invalid processing_mode: {self.processing_mode}")
        return errors

    def is_valid(self) -> bool:
        return not self.validate()

    def get_output_grid(self) -> OutputGrid:
        return OutputGrid(
            n_time=int((self.time_max_ms - self.time_min_ms) / self.dt_ms) + 1,
            n_inline=(self.inline_max - self.inline_min) // self.inline_step + 1,
            n_xline=(self.xline_max - self.xline_min) // self.xline_step + 1,
            dt=self.dt_ms / 1000.0, d_inline=self.inline_spacing, d_xline=self.xline_spacing,
            t0=self.time_min_ms / 1000.0, inline_start=self.inline_min, xline_start=self.xline_min,
            x_origin=self.origin_x, y_origin=self.origin_y, inline_azimuth=self.grid_azimuth_deg)

    def get_migration_config(self) -> MigrationConfig:
        return MigrationConfig(
            output_grid=self.get_output_grid(), max_aperture_m=self.max_aperture_m,
            max_angle_deg=self.max_angle_deg, max_offset_m=self.max_offset_m,
            min_offset_m=self.near_offset_taper_m, taper_width=self.taper_width_samples / 100.0,
            traveltime_mode=TraveltimeMode.STRAIGHT_RAY, weight_mode=WeightMode.SPREADING,
            interpolation_mode=InterpolationMode.LINEAR, antialias_enabled=self.anti_alias_filter)

    def get_velocity_model(self) -> VelocityModel:
        if self.velocity_file:
            return VelocityModel.load(self.velocity_file)
        if self.velocity_gradient:
            import numpy as np
            g = self.get_output_grid()
            t = g.time_axis
            return VelocityModel(data=(self.velocity_v0 + self.velocity_gradient * t).astype("float32"),
                                 z_axis=t, gradient=self.velocity_gradient, v0=self.velocity_v0)
        return VelocityModel(data=float(self.velocity_v0))

    def get_binning_table(self) -> BinningTable:
        """The explicit table, else a fresh table of the named preset."""
        if self.binning_table is not None:
            return self.binning_table
        if self.binning_preset:
            from ..utils.binning_presets import get_preset
            return get_preset(self.binning_preset)
        raise ValueError("No binning configuration available")

    def to_dict(self) -> Dict[str, Any]:
        d = {k: getattr(self, k) for k in (
            "name", "input_file", "output_directory", "velocity_v0", "velocity_gradient", "velocity_file",
            "time_min_ms", "time_max_ms", "dt_ms", "inline_min", "inline_max", "inline_step", "xline_min",
            "xline_max", "xline_step", "origin_x", "origin_y", "inline_spacing", "xline_spacing",
            "grid_azimuth_deg", "max_aperture_m", "max_angle_deg", "max_offset_m", "taper_width_samples",
            "near_offset_taper_m", "anti_alias_filter", "binning_preset", "processing_mode", "max_workers",
            "memory_limit_gb", "enable_checkpointing", "output_format", "create_stack_volume",
            "create_fold_volume", "compress_output", "created_at", "description", "metadata")}
        d["header_mapping"] = self.header_mapping.to_dict()
        if self.binning_table is not None:
            d["binning_table"] = self.binning_table.to_dict()
        return d

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "MigrationJobConfig":
        d = dict(d)
        hm = HeaderMapping.from_dict(d.pop("header_mapping", {}) or {})
        bt = d.pop("binning_table", None)
        return cls(header_mapping=hm, binning_table=BinningTable.from_dict(bt) if bt is not None else None,
                   **{k: v for k, v in d.items() if k in cls.__dataclass_fields__})

    def save(self, filepath: str) -> None:
        with open(filepath, "w") as f:
            json.dump(self.to_dict(), f, indent=2)
        logger.info("Saved migration job config to %s", filepath)

    @classmethod
    def load(cls, filepath: str) -> "MigrationJobConfig":
        with open(filepath) as f:
            return cls.from_dict(json.load(f))


# ---------------------------------------------------------------------------
# job templates (reference models/migration_job.py:341-491): a preset binning
# and survey-typical velocity, aperture and angle per survey kind

_TEMPLATES: Dict[str, Dict[str, Any]] = {
    "land_3d": dict(name="Land 3D Migration", binning_preset="land_3d", max_aperture_m=3000.0, max_angle_deg=60.0,
                    description="Standard land 3D survey migration with 10 offset bins"),
    "marine": dict(name="Marine Migration", binning_preset="marine", max_aperture_m=5000.0, max_angle_deg=50.0,
                   description="Marine streamer migration with 6 offset bins"),
    "wide_azimuth_ovt": dict(name="Wide Azimuth OVT Migration", binning_preset="wide_azimuth_ovt",
                             max_aperture_m=4000.0, max_angle_deg=65.0,
                             description="Wide-azimuth migration with 16 OVT bins (4 offset x 4 azimuth)"),
    "full_stack": dict(name="Full Stack Migration", binning_preset="full_stack", max_aperture_m=5000.0,
                       max_angle_deg=70.0, create_stack_volume=True,
                       description="Single full-stack migration output"),
}


def _from_template(kind: str, input_file: str, output_directory: str, **values) -> MigrationJobConfig:
    return MigrationJobConfig(input_file=input_file, output_directory=output_directory, **_TEMPLATES[kind], **values)


def create_land_3d_template(input_file: str, output_directory: str, velocity_v0: float = 3000.0) -> MigrationJobConfig:
    return _from_template("land_3d", input_file, output_directory, velocity_v0=velocity_v0)


def create_marine_template(input_file: str, output_directory: str, velocity_v0: float = 1500.0,
                           velocity_gradient: float = 0.3) -> MigrationJobConfig:
    return _from_template("marine", input_file, output_directory, velocity_v0=velocity_v0,
                          velocity_gradient=velocity_gradient)


def create_wide_azimuth_template(input_file: str, output_directory: str,
                                 velocity_v0: float = 2800.0) -> MigrationJobConfig:
    return _from_template("wide_azimuth_ovt", input_file, output_directory, velocity_v0=velocity_v0)


def create_full_stack_template(input_file: str, output_directory: str,
                               velocity_v0: float = 2500.0) -> MigrationJobConfig:
    return _from_template("full_stack", input_file, output_directory, velocity_v0=velocity_v0)


JOB_TEMPLATES = {"land_3d": create_land_3d_template, "marine": create_marine_template,
                 "wide_azimuth_ovt": create_wide_azimuth_template, "full_stack": create_full_stack_template}


def list_templates() -> List[str]:
    return list(JOB_TEMPLATES)


def create_from_template(template_name: str, input_file: str, output_directory: str, **kwargs) -> MigrationJobConfig:
    """The named template; keyword arguments go to its factory (velocity_v0,
    and velocity_gradient for marine)."""
    if template_name not in JOB_TEMPLATES:
        raise ValueError(f"Unknown template '{template_name}'. Available: {', '.join(JOB_TEMPLATES)}")
    return JOB_TEMPLATES[template_name](input_file, output_directory, **kwargs)
