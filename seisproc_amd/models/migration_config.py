"""Migration configuration objects (reference models/migration_config.py:20-359).

Same enums, dataclasses, field names, defaults, validation messages and
to_dict/from_dict as the reference.  ``MigrationConfig.to_pstm_params`` maps
a configuration onto the C-ABI job description of the HIP kernels.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from enum import Enum
from typing import Any, Dict, Tuple

import numpy as np


class TraveltimeMode(Enum):
    STRAIGHT_RAY = "straight_ray"
    CURVED_RAY = "curved_ray"
    VTI = "vti"


class AnisotropyMethod(Enum):
    ANELLIPTIC = "anelliptic"
    EXACT = "exact"
    WEAK = "weak"


class InterpolationMode(Enum):
    LINEAR = "linear"
    SINC = "sinc"


class WeightMode(Enum):
    NONE = "none"
    SPREADING = "spreading"
    OBLIQUITY = "obliquity"
    FULL = "full"


@dataclass
class OutputGrid:
    """Output image grid.  dt and t0 in seconds, spacings in metres,
    inline_azimuth in degrees from north."""

    n_time: int = 1000
    n_inline: int = 100
    n_xline: int = 100
    dt: float = 0.004
    d_inline: float = 25.0
    d_xline: float = 25.0
    t0: float = 0.0
    inline_start: int = 1
    xline_start: int = 1
    x_origin: float = 0.0
    y_origin: float = 0.0
    inline_azimuth: float = 0.0

    def __post_init__(self):
        for name in ("n_time", "n_inline", "n_xline"):
            if getattr(self, name) <= 0:
                raise ValueError(f"{name} must be positive, got {getattr(self, name)}")
        for name in ("dt", "d_inline", "d_xline"):
            if getattr(self, name) <= 0:
                raise ValueError(f"{name} must be positive, got {getattr(self, name)}")

    @property
    def t_max(self) -> float:
        return self.t0 + (self.n_time - 1) * self.dt

    @property
    def time_axis(self) -> np.ndarray:
        return np.arange(self.n_time) * self.dt + self.t0

    @property
    def inline_axis(self) -> np.ndarray:
        return np.arange(self.n_inline) + self.inline_start

    @property
    def xline_axis(self) -> np.ndarray:
        return np.arange(self.n_xline) + self.xline_start

    @property
    def x_extent(self) -> float:
        return (self.n_inline - 1) * self.d_inline

    @property
    def y_extent(self) -> float:
        return (self.n_xline - 1) * self.d_xline

    @property
    def total_samples(self) -> int:
        return self.n_time * self.n_inline * self.n_xline

    @property
    def memory_gb(self) -> float:
        return self.total_samples * 4 / (1024 ** 3)

    def get_coordinates(self, inline_idx, xline_idx) -> Tuple[float, float]:
        """World (X, Y) of an image column; inline axis at `inline_azimuth`
        from north, crossline axis 90 degrees clockwise from it
        (reference models/migration_config.py:135-163)."""
        az = math.radians(self.inline_azimuth)
        di, dx = inline_idx * self.d_inline, xline_idx * self.d_xline
        x = self.x_origin + di * math.sin(az) + dx * math.cos(az)
        y = self.y_origin + di * math.cos(az) - dx * math.sin(az)
        return x, y

    def step_vectors(self, honor_azimuth: bool = False):
        """World step per inline and per crossline index.  The reference's
        migration kernels ignore the azimuth (x = x0 + il*d_il, y = y0 +
        xl*d_xl: migration_engine.py:292-301, kirchhoff_migrator.py:404-414);
        honor_azimuth=True uses get_coordinates' rotation instead."""
        if not honor_azimuth:
            return (self.d_inline, 0.0), (0.0, self.d_xline)
        az = math.radians(self.inline_azimuth)
        return ((self.d_inline * math.sin(az), self.d_inline * math.cos(az)),
                (self.d_xline * math.cos(az), -self.d_xline * math.sin(az)))

    def to_dict(self) -> Dict[str, Any]:
        return {k: getattr(self, k) for k in (
            "n_time", "n_inline", "n_xline", "dt", "d_inline", "d_xline", "t0", "inline_start",
            "xline_start", "x_origin", "y_origin", "inline_azimuth")}

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "OutputGrid":
        return cls(**d)


@dataclass
class MigrationConfig:
    """Kirchhoff PSTM configuration (aperture, traveltime, amplitude,
    antialias, interpolation, processing)."""

    output_grid: OutputGrid = field(default_factory=OutputGrid)
    max_aperture_m: float = 5000.0
    max_angle_deg: float = 60.0
    max_offset_m: float = 10000.0
    min_offset_m: float = 0.0
    taper_width: float = 0.1
    traveltime_mode: TraveltimeMode = TraveltimeMode.STRAIGHT_RAY
    use_anisotropy: bool = False
    anisotropy_method: AnisotropyMethod = AnisotropyMethod.ANELLIPTIC
    weight_mode: WeightMode = WeightMode.SPREADING
    preserve_amplitudes: bool = False
    antialias_enabled: bool = True
    antialias_fmax: float = 80.0
    interpolation_mode: InterpolationMode = InterpolationMode.SINC
    min_fold: int = 1
    normalize_by_fold: bool = True
    metadata: Dict[str, Any] = field(default_factory=dict)

    def __post_init__(self):
        self._validate_aperture()
        self._validate_processing()

    def _validate_aperture(self):
        if self.max_aperture_m <= 0:
            raise ValueError(f"max_aperture_m must be positive, got {self.max_aperture_m}")
        if not 0 < self.max_angle_deg <= 90:
            raise ValueError(f"max_angle_deg must be in (0, 90], got {self.max_angle_deg}")
        if self.max_offset_m <= 0:
            raise ValueError(f"max_offset_m must be positive, got {self.max_offset_m}")
        if self.min_offset_m < 0:
            raise ValueError(f"min_offset_m must be non-negative, got {self.min_offset_m}")
        if self.min_offset_m >= self.max_offset_m:
            raise ValueError(f"min_offset_m ({self.min_offset_m}) must be less than "
                             f"max_offset_m ({self.max_offset_m})")
        if not 0 <= self.taper_width <= 1:
            raise ValueError(f"taper_width must be in [0, 1], got {self.taper_width}")

    def _validate_processing(self):
        if self.min_fold < 1:
            raise ValueError(f"min_fold must be >= 1, got {self.min_fold}")
        if self.antialias_fmax <= 0:
            raise ValueError(f"antialias_fmax must be positive, got {self.antialias_fmax}")

    @property
    def max_angle_rad(self) -> float:
        return math.radians(self.max_angle_deg)

    def get_summary(self) -> str:
        """One ``name: value`` line per setting of the migration."""
        g = self.output_grid
        rows = [
            ("output", f"{g.n_inline} x {g.n_xline} columns x {g.n_time} samples"),
            ("aperture", f"{self.max_aperture_m:g} m, angle {self.max_angle_deg:g} deg"),
            ("offsets", f"{self.min_offset_m:g}..{self.max_offset_m:g} m"),
            ("traveltime", self.traveltime_mode.value),
            ("weights", self.weight_mode.value),
            ("antialias", "on" if self.antialias_enabled else "off"),
        ]
        if self.use_anisotropy:
            rows.append(("anisotropy", f"VTI, {self.anisotropy_method.value}"))
        return "PSTM configuration\n" + "\n".join(f"  {k:<10} {v}" for k, v in rows)

    def to_dict(self) -> Dict[str, Any]:
        return {
            "output_grid": self.output_grid.to_dict(), "max_aperture_m": self.max_aperture_m,
            "max_angle_deg": self.max_angle_deg, "max_offset_m": self.max_offset_m,
            "min_offset_m": self.min_offset_m, "taper_width": self.taper_width,
            "traveltime_mode": self.traveltime_mode.value, "use_anisotropy": self.use_anisotropy,
            "anisotropy_method": self.anisotropy_method.value, "weight_mode": self.weight_mode.value,
            "preserve_amplitudes": self.preserve_amplitudes, "antialias_enabled": self.antialias_enabled,
            "antialias_fmax": self.antialias_fmax, "interpolation_mode": self.interpolation_mode.value,
            "min_fold": self.min_fold, "normalize_by_fold": self.normalize_by_fold,
            "metadata": dict(self.metadata),
        }

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "MigrationConfig":
        d = dict(d)
        grid = OutputGrid.from_dict(d.pop("output_grid"))
        return cls(output_grid=grid,
                   traveltime_mode=TraveltimeMode(d.pop("traveltime_mode", "straight_ray")),
                   weight_mode=WeightMode(d.pop("weight_mode", "spreading")),
                   interpolation_mode=InterpolationMode(d.pop("interpolation_mode", "sinc")),
                   anisotropy_method=AnisotropyMethod(d.pop("anisotropy_method", "anelliptic")),
                   **d)

    def copy(self) -> "MigrationConfig":
        return MigrationConfig.from_dict(self.to_dict())

    # ------------------------------------------------------------------
    def to_pstm_params(self, n_traces: int, n_samples: int, dt_s: float, t0_s: float = 0.0,
                       honor_azimuth: bool = False, kahan: bool = True, aa_scale: float = 1.0,
                       aa_max: float = 16.0, zero_outside: bool = False,
                       time_dependent_aperture: bool = False, taper: bool = False):
        """C-ABI job description for this configuration.

        The offset window [min_offset_m, max_offset_m] is a property of the
        plan (traces outside it are dropped); ``taper`` turns on the
        reference's taper composition (processors/migration/aperture.py:
        147-270: distance and angle cosine tapers of both legs and the offset
        window's near/far tapers, each over taper_width of its range).

        Weight modes map one to one (FULL = cos*cos/(r*r) of
        kirchhoff_kernel.py:691-693, SPREADING = 1/(r_s r_r) of
        kirchhoff_migrator.py:378).  The traveltime is the straight-ray DSR,
        or the anelliptic VTI time when use_anisotropy or
        traveltime_mode == VTI.  Interpolation is linear (the only one the
        reference kernels use, kirchhoff_kernel.py:708-721).
        """
        from ..abi_types import (FLAG_AA, FLAG_KAHAN, FLAG_TAPER, FLAG_TIME_APERTURE, FLAG_VTI,
                                 FLAG_ZERO_OUTSIDE, WEIGHT_CODES, PstmParams)
        if self.traveltime_mode is TraveltimeMode.CURVED_RAY:
            raise NotImplementedError("curved-ray traveltimes are not part of the PSTM path")
        g = self.output_grid
        (ilx, ily), (xlx, xly) = g.step_vectors(honor_azimuth)
        flags = 0
        if self.antialias_enabled:
            flags |= FLAG_AA
        if self.use_anisotropy or self.traveltime_mode is TraveltimeMode.VTI:
            flags |= FLAG_VTI
        if kahan:
            flags |= FLAG_KAHAN
        if zero_outside:
            flags |= FLAG_ZERO_OUTSIDE
        if time_dependent_aperture:
            flags |= FLAG_TIME_APERTURE
        if taper and 0 < self.taper_width < 1:
            flags |= FLAG_TAPER
        return PstmParams.from_dict(dict(
            n_traces=int(n_traces), n_samples=int(n_samples), n_il=g.n_inline, n_xl=g.n_xline,
            n_tau=g.n_time, dt=float(dt_s), t0=float(t0_s), dtau=g.dt, tau0=g.t0,
            origin_x=g.x_origin, origin_y=g.y_origin, il_dx=ilx, il_dy=ily, xl_dx=xlx, xl_dy=xly,
            aperture_m=self.max_aperture_m, max_angle_deg=self.max_angle_deg,
            taper_width=self.taper_width if 0 < self.taper_width < 1 else 0.1,
            aa_scale=aa_scale, aa_max=aa_max, flags=flags,
            weight_mode=WEIGHT_CODES[self.weight_mode.value],
            min_offset_m=self.min_offset_m, max_offset_m=self.max_offset_m))


def create_default_config(n_time: int = 1000, n_inline: int = 100, n_xline: int = 100,
                          dt_ms: float = 4.0, d_inline_m: float = 25.0,
                          d_xline_m: float = 25.0) -> MigrationConfig:
    return MigrationConfig(output_grid=OutputGrid(n_time=n_time, n_inline=n_inline, n_xline=n_xline,
                                                  dt=dt_ms / 1000.0, d_inline=d_inline_m,
                                                  d_xline=d_xline_m))


def create_high_resolution_config(n_time: int = 2000, n_inline: int = 200, n_xline: int = 200,
                                  max_aperture_m: float = 8000.0,
                                  max_angle_deg: float = 70.0) -> MigrationConfig:
    return MigrationConfig(
        output_grid=OutputGrid(n_time=n_time, n_inline=n_inline, n_xline=n_xline, dt=0.002,
                               d_inline=12.5, d_xline=12.5),
        max_aperture_m=max_aperture_m, max_angle_deg=max_angle_deg,
        interpolation_mode=InterpolationMode.SINC, antialias_enabled=True, weight_mode=WeightMode.FULL)
