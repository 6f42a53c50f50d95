"""Wizard configuration -> MigrationEngine parameters
(reference processors/migration/config_adapter.py:21-395).

Same keys, fallbacks and counts: n_il = (inline_max - inline_min) //
inline_step + 1 (:118-123), n_samples = (time_max - time_min) / dt + 1
(:126-131), bin sizes output_bin_* -> input_bin_* -> 25 m, and the same
validation messages.  ``get_engine_params`` returns the kwargs of
``MigrationEngine.migrate_bin_full``.
"""
from __future__ import annotations

import logging
from dataclasses import dataclass
from typing import Any, Callable, Dict, Optional, Tuple

import numpy as np

logger = logging.getLogger(__name__)


@dataclass
class MigrationParams:
    origin_x: float
    origin_y: float
    il_spacing: float
    xl_spacing: float
    azimuth_deg: float
    n_il: int
    n_xl: int
    dt_ms: float
    t_min_ms: float
    n_samples: int
    velocity_mps: float
    max_aperture_m: float
    max_angle_deg: float
    tile_size: int = 100
    use_time_domain: bool = False
    use_kdtree: bool = False
    sample_batch_size: int = 200
    max_traces_per_tile: int = 5000


class ConfigAdapter:
    def __init__(self, config: Dict[str, Any]):
        self.config = config
        self._params: Optional[MigrationParams] = None

    @property
    def params(self) -> MigrationParams:
        if self._params is None:
            self._params = self._extract_params()
        return self._params

    def _extract_params(self) -> MigrationParams:
        cfg = self.config
        il_sp = cfg.get("output_bin_il", cfg.get("input_bin_il", 25.0))
        xl_sp = cfg.get("output_bin_xl", cfg.get("input_bin_xl", 25.0))
        n_il = max(1, (cfg.get("inline_max", 100) - cfg.get("inline_min", 1)) // cfg.get("inline_step", 1) + 1)
        n_xl = max(1, (cfg.get("xline_max", 100) - cfg.get("xline_min", 1)) // cfg.get("xline_step", 1) + 1)
        dt_ms = cfg.get("dt_ms", 4.0)
        if dt_ms <= 0:
            dt_ms = 4.0
        t_min = cfg.get("time_min_ms", 0.0)
        t_max = cfg.get("time_max_ms", 6000.0)
        n_samples = max(1, int((t_max - t_min) / dt_ms) + 1)
        p = MigrationParams(
            origin_x=cfg.get("x_origin", 0.0), origin_y=cfg.get("y_origin", 0.0), il_spacing=il_sp,
            xl_spacing=xl_sp, azimuth_deg=cfg.get("grid_azimuth_deg", 0.0), n_il=n_il, n_xl=n_xl,
            dt_ms=dt_ms, t_min_ms=t_min, n_samples=n_samples, velocity_mps=cfg.get("velocity_v0", 2500.0),
            max_aperture_m=cfg.get("max_aperture_m", 5000.0), max_angle_deg=cfg.get("max_angle_deg", 60.0),
            tile_size=cfg.get("tile_size", 100), use_time_domain=cfg.get("use_time_domain", False),
            use_kdtree=cfg.get("use_kdtree", False), sample_batch_size=cfg.get("sample_batch_size", 200),
            max_traces_per_tile=cfg.get("max_traces_per_tile", 5000))
        logger.info("ConfigAdapter: grid=%dx%d, dt=%sms, v=%sm/s, aperture=%sm, angle=%s deg",
                    p.n_il, p.n_xl, p.dt_ms, p.velocity_mps, p.max_aperture_m, p.max_angle_deg)
        return p

    def get_engine_params(self, traces: np.ndarray, source_x: np.ndarray, source_y: np.ndarray,
                          receiver_x: np.ndarray, receiver_y: np.ndarray, normalize: bool = True,
                          min_fold: int = 1,
                          progress_callback: Optional[Callable[[float, str], None]] = None,
                          enable_profiling: bool = False,
                          use_time_dependent_aperture: bool = False) -> Dict[str, Any]:
        p = self.params
        return {
            "traces": traces, "source_x": source_x, "source_y": source_y, "receiver_x": receiver_x,
            "receiver_y": receiver_y, "origin_x": p.origin_x, "origin_y": p.origin_y,
            "il_spacing": p.il_spacing, "xl_spacing": p.xl_spacing, "azimuth_deg": p.azimuth_deg,
            "n_il": p.n_il, "n_xl": p.n_xl, "dt_ms": p.dt_ms, "t_min_ms": p.t_min_ms,
            "n_times": p.n_samples, "velocity_mps": p.velocity_mps, "max_aperture_m": p.max_aperture_m,
            "max_angle_deg": p.max_angle_deg, "normalize": normalize, "min_fold": min_fold,
            "progress_callback": progress_callback, "enable_profiling": enable_profiling,
            "use_time_dependent_aperture": use_time_dependent_aperture,
        }

    def get_time_domain_params(self, *args, **kwargs) -> Dict[str, Any]:
        raise NotImplementedError(
            "the time-domain kernels of the reference (migrate_bin_time_domain[_rms]) map input to output "
            "time with t_out = sqrt(t_in^2 + 4h^2/v^2) -- the modelling direction -- and are not a DSR "
            "migration; use get_engine_params / migrate_bin_full")

    def validate(self) -> Tuple[bool, str]:
        issues = []
        cfg = self.config
        if cfg.get("dt_ms", 4.0) <= 0:
            issues.append(f"Invalid sample interval: {cfg.get('dt_ms')}ms")
        if cfg.get("inline_max", 100) < cfg.get("inline_min", 1):
            issues.append(f"Invalid inline range: max ({cfg.get('inline_max')}) < min ({cfg.get('inline_min')})")
        if cfg.get("xline_max", 100) < cfg.get("xline_min", 1):
            issues.append(f"Invalid crossline range: max ({cfg.get('xline_max')}) < min ({cfg.get('xline_min')})")
        p = self.params
        if p.n_il <= 0:
            issues.append(f"Invalid inline count: {p.n_il}")
        if p.n_xl <= 0:
            issues.append(f"Invalid crossline count: {p.n_xl}")
        if p.n_il * p.n_xl > 1e8:
            issues.append(f"Grid too large: {p.n_il}x{p.n_xl} = {p.n_il * p.n_xl:.0e} points")
        if p.n_samples <= 0:
            issues.append(f"Invalid sample count: {p.n_samples}")
        if p.velocity_mps <= 0:
            issues.append(f"Invalid velocity: {p.velocity_mps} m/s")
        elif p.velocity_mps < 1000 or p.velocity_mps > 8000:
            issues.append(f"Unusual velocity: {p.velocity_mps} m/s (typical range: 1000-8000)")
        if p.max_aperture_m <= 0:
            issues.append(f"Invalid aperture: {p.max_aperture_m}m")
        if p.max_angle_deg <= 0 or p.max_angle_deg > 90:
            issues.append(f"Invalid max angle: {p.max_angle_deg}°")
        if issues:
            return False, "; ".join(issues)
        return True, "Configuration valid"

    def get_summary(self) -> str:
        """A log-ready description of the extracted parameters (the monitor
        dialog logs it, reference views/migration_monitor_dialog.py:140): one
        ``name: value`` line each for the grid, origin, bins, azimuth, time
        axis, velocity and aperture."""
        p = self.params
        t_end = p.t_min_ms + (p.n_samples - 1) * p.dt_ms
        rows = (
            ("grid", f"{p.n_il} x {p.n_xl} (inline x crossline), {p.n_samples} samples"),
            ("origin", f"x {p.origin_x:.1f} m, y {p.origin_y:.1f} m"),
            ("bins", f"{p.il_spacing:g} m inline, {p.xl_spacing:g} m crossline"),
            ("azimuth", f"{p.azimuth_deg:g} deg"),
            ("time", f"{p.t_min_ms:g}..{t_end:g} ms every {p.dt_ms:g} ms"),
            ("velocity", f"{p.velocity_mps:g} m/s"),
            ("aperture", f"{p.max_aperture_m:g} m, angle {p.max_angle_deg:g} deg"),
        )
        return "PSTM parameters\n" + "\n".join(f"  {k:<9} {v}" for k, v in rows)


def create_adapter_from_wizard(config: Dict[str, Any]) -> ConfigAdapter:
    return ConfigAdapter(config)
