"""KirchhoffMigrator: the class API of the reference
(processors/migration/kirchhoff_migrator.py:41-592) on the HIP backend.

``migrate_gather(gather, geometry, output_image=None, output_fold=None)``
accumulates one gather into (n_time, n_inline, n_xline) GPU tensors like the
reference (:173-256); ``migrate_dataset`` loops over gathers and applies the
reference's fold normalisation (:496-501) into a MigrationResult.  The grid
follows the reference's _grid_to_world (x = x0 + il*d_il, y = y0 + xl*d_xl,
:404-414); the weight is the configuration's (default SPREADING =
1/(r_s r_r), the reference's :378).  The migrator is also a BaseProcessor
(``process(SeismicData)``), so the registry entry 'KirchhoffMigrator'
reconstructs through BaseProcessor.from_dict.
"""
from __future__ import annotations

import logging
import time
from dataclasses import dataclass
from typing import Any, Callable, Dict, List, Optional, Tuple

import numpy as np
import torch

from ...models.migration_config import MigrationConfig
from ...models.migration_geometry import MigrationGeometry
from ...models.seismic_data import SeismicData
from ...models.velocity_model import VelocityModel
from ..base_processor import BaseProcessor
from .base_migrator import BaseMigrator, MigrationResult
from .pstm_processor import PSTMProcessor, _as_anisotropy, _as_config, _as_velocity, migrate_arrays

logger = logging.getLogger(__name__)


class KirchhoffMigrator(BaseMigrator, BaseProcessor):
    def __init__(self, velocity=None, config=None, device: Optional[torch.device] = None,
                 use_traveltime_lut: bool = True, anisotropy=None, trace_dtype: str = "f32", **params):
        vel = _as_velocity(velocity if velocity is not None else params.pop("velocity_model", 2500.0))
        cfg = _as_config(config if config is not None else MigrationConfig())
        self.velocity = vel
        self.anisotropy = _as_anisotropy(anisotropy)
        self.trace_dtype = trace_dtype
        if not torch.cuda.is_available():
            raise RuntimeError("KirchhoffMigrator needs a ROCm GPU; there is no CPU fallback")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.params = dict(params)
        self.params.update(velocity=vel.to_dict(), config=cfg.to_dict(), trace_dtype=trace_dtype,
                           anisotropy=self.anisotropy.to_dict() if self.anisotropy else None)
        self._progress_callback = None
        BaseMigrator.__init__(self, vel, cfg)

    # BaseProcessor and BaseMigrator both validate; one method serves both.
    def _validate_params(self):
        self._validate_inputs()

    def _validate_inputs(self):
        if self.velocity.v0 is None or float(np.min(np.atleast_1d(self.velocity.data))) <= 0:
            raise ValueError("Velocity model must have positive v0")
        if self.config.output_grid.n_time <= 0:
            raise ValueError("Output grid must have positive n_time")

    def _empty(self) -> torch.Tensor:
        g = self.config.output_grid
        return torch.zeros((g.n_time, g.n_inline, g.n_xline), dtype=torch.float32, device=self.device)

    def migrate_gather(self, gather: SeismicData, geometry: MigrationGeometry,
                       output_image: Optional[torch.Tensor] = None,
                       output_fold: Optional[torch.Tensor] = None) -> Tuple[torch.Tensor, torch.Tensor]:
        """Accumulate one gather; returns (image, fold) of shape (n_time, n_inline, n_xline)."""
        g = self.config.output_grid
        img_k = output_image.permute(1, 2, 0).contiguous() if output_image is not None else None
        fold_k = output_fold.permute(1, 2, 0).contiguous() if output_fold is not None else None
        image, comp, fold, _ = migrate_arrays(self.config, self.velocity, gather.traces, gather.sample_rate,
                                              geometry, anisotropy=self.anisotropy, trace_dtype=self.trace_dtype,
                                              image=img_k, fold=fold_k, device=self.device)
        if comp is not None:
            image = image + comp
        return image.permute(2, 0, 1).contiguous(), fold.permute(2, 0, 1).contiguous()

    def migrate_dataset(self, gathers: List[SeismicData], geometries: List[MigrationGeometry],
                        progress_callback: Optional[Callable[[float, str], None]] = None,
                        batch_size: int = 1) -> MigrationResult:
        t0 = time.time()
        image, fold = self._empty(), self._empty()
        n = len(gathers)
        for i, (gth, geo) in enumerate(zip(gathers, geometries)):
            if progress_callback:
                progress_callback(i / max(n, 1) * 100, f"Migrating gather {i + 1}/{n}")
            image, fold = self.migrate_gather(gth, geo, image, fold)
            self._report_progress(i + 1, n, f"gather {i + 1}/{n}")
        if self.config.normalize_by_fold:
            safe = torch.where(fold > 0, fold, torch.ones_like(fold))
            image = torch.where(fold >= self.config.min_fold, image / safe, torch.zeros_like(image))
        elapsed = time.time() - t0
        if progress_callback:
            progress_callback(100.0, "Migration complete")
        return MigrationResult(image=image.cpu().numpy(), fold=fold.cpu().numpy(), config=self.config,
                               metadata={"n_gathers": n, "elapsed_seconds": elapsed,
                                         "traces_per_second": sum(g.n_traces for g in gathers) / max(elapsed, 1e-9),
                                         "velocity_v0": self.velocity.v0})

    def estimate_memory_gb(self, n_traces: int, n_samples: int) -> float:
        g = self.config.output_grid
        per_trace = n_samples * 4 + 48 + 16
        grid = g.n_time * g.n_inline * g.n_xline * 4 * 3  # image, compensation, fold
        return (n_traces * per_trace + grid) / 1024 ** 3

    def get_description(self) -> str:
        g = self.config.output_grid
        return (f"Kirchhoff PSTM (HIP/gfx950)\n  Grid: {g.n_inline}x{g.n_xline}x{g.n_time}\n"
                f"  Aperture: {self.config.max_aperture_m}m, {self.config.max_angle_deg}°\n"
                f"  Weights: {self.config.weight_mode.value}")

    def process(self, data: SeismicData) -> SeismicData:
        proc = PSTMProcessor(config=self.config.to_dict(), velocity=self.velocity.to_dict(),
                             anisotropy=self.anisotropy.to_dict() if self.anisotropy else None,
                             trace_dtype=self.trace_dtype)
        proc.set_progress_callback(self._progress_callback)
        return proc.process(data)

    def to_dict(self) -> Dict[str, Any]:
        """The processor form (class_name, module, params: what
        BaseProcessor.from_dict rebuilds the migrator from) with the
        migrator form's velocity_model and config beside it."""
        d = BaseProcessor.to_dict(self)
        d.update(velocity_model=self.velocity.to_dict(), config=self.config.to_dict())
        return d


def create_kirchhoff_migrator(velocity, config: MigrationConfig, device=None, **kw) -> KirchhoffMigrator:
    return KirchhoffMigrator(velocity, config, device=device, **kw)


@dataclass
class OptimizationStats:
    """Statistics of the last migrate_gather call, with the reference's fields
    (processors/migration/optimized_kirchhoff_migrator.py:35-48).  This
    backend fills the trace counts and time_accumulation (the whole call);
    it has no traveltime table or reciprocity pairs, so those stay 0."""

    total_traces: int = 0
    traces_processed: int = 0
    trace_reduction_ratio: float = 1.0
    depth_groups_used: int = 0
    reciprocal_pairs: int = 0
    traveltime_lut_hits: int = 0
    memory_peak_mb: float = 0.0
    time_traveltime: float = 0.0
    time_aperture: float = 0.0
    time_interpolation: float = 0.0
    time_accumulation: float = 0.0


class OptimizedKirchhoffMigrator(KirchhoffMigrator):
    """The interactive path's migrator (reference
    processors/migration/optimized_kirchhoff_migrator.py:50-719, called from
    main_window.py:4642-4653) on the HIP backend.

    Same constructor flags.  They name torch-side strategies of the
    reference that the HIP kernel carries out by construction, so they change
    nothing here and are kept for the interface and its description:
    ``enable_spatial_index`` and ``enable_adaptive_aperture`` are the plan's
    midpoint cells and each tile's aperture/angle reach, ``enable_gpu_tiling``
    the tile kernel itself, ``enable_lut`` is not needed (the DSR time costs
    two reciprocal square roots), and ``enable_symmetry`` is off in the
    reference too.  The sum is the reference tile loop's
    (:356-562): SPREADING weight 1/(r_s r_r) (the configuration's weight
    mode), hard aperture and angle masks, linear interpolation.
    """

    def __init__(self, velocity, config, device: Optional[torch.device] = None, enable_lut: bool = True,
                 enable_adaptive_aperture: bool = True, enable_spatial_index: bool = True,
                 enable_symmetry: bool = False, enable_gpu_tiling: bool = True, **kw):
        self._opts = dict(lut=bool(enable_lut), adaptive_aperture=bool(enable_adaptive_aperture),
                          spatial_index=bool(enable_spatial_index), symmetry=bool(enable_symmetry),
                          gpu_tiling=bool(enable_gpu_tiling))
        self._stats = OptimizationStats()
        self._profile_tiles = False
        self._bin_traces = None
        super().__init__(velocity, config, device=device, **kw)

    def migrate_gather(self, gather: SeismicData, geometry: MigrationGeometry,
                       output_image: Optional[torch.Tensor] = None,
                       output_fold: Optional[torch.Tensor] = None) -> Tuple[torch.Tensor, torch.Tensor]:
        t0 = time.time()
        image, fold = super().migrate_gather(gather, geometry, output_image, output_fold)
        torch.cuda.synchronize(self.device)
        off = geometry.offset
        used = int(np.count_nonzero((off >= self.config.min_offset_m) & (off <= self.config.max_offset_m)))
        n = int(gather.n_traces)
        self._stats = OptimizationStats(total_traces=n, traces_processed=used,
                                        trace_reduction_ratio=used / n if n else 1.0,
                                        memory_peak_mb=torch.cuda.max_memory_allocated(self.device) / 2**20,
                                        time_accumulation=time.time() - t0)
        return image, fold

    def initialize_for_bin(self, all_geometry: MigrationGeometry) -> None:
        """Start a bin made of several gathers (reference :260-284, where it
        builds the bin's trace spatial index once).  Each call's HIP plan
        bins its own traces by midpoint on the GPU, so nothing is prebuilt;
        the bin's trace count is kept for the statistics."""
        self._bin_traces = int(all_geometry.n_traces)

    def reset_for_new_bin(self) -> None:
        """Forget the previous bin and its statistics (reference :286-293)."""
        self._bin_traces = None
        self._stats = OptimizationStats()

    def get_optimization_stats(self) -> OptimizationStats:
        """Statistics of the last migrate_gather call."""
        return self._stats

    def enable_profiling(self, enable: bool = True) -> None:
        """Kept for the interface: per-tile timing is taken with rocprofv3."""
        self._profile_tiles = bool(enable)

    def get_description(self) -> str:
        o = self._opts
        return (super().get_description() +
                f"\n  Optimizations: LUT={o['lut']}, AdaptiveAperture={o['adaptive_aperture']}, "
                f"SpatialIndex={o['spatial_index']}, Symmetry={o['symmetry']}, GPUTiling={o['gpu_tiling']}")


_OPT_LEVELS = {
    "low": dict(enable_lut=True, enable_adaptive_aperture=False, enable_spatial_index=False, enable_symmetry=False,
                enable_gpu_tiling=False),
    "medium": dict(enable_lut=True, enable_adaptive_aperture=True, enable_spatial_index=True, enable_symmetry=False,
                   enable_gpu_tiling=False),
    "high": dict(enable_lut=True, enable_adaptive_aperture=True, enable_spatial_index=True, enable_symmetry=False,
                 enable_gpu_tiling=True),
}


def create_optimized_kirchhoff_migrator(velocity, config: MigrationConfig, prefer_gpu: bool = True,
                                        optimization_level: str = "high") -> OptimizedKirchhoffMigrator:
    """Factory of the interactive path (reference
    optimized_kirchhoff_migrator.py:721-775): same keywords and levels
    ('low', 'medium', anything else = 'high').  The migration runs on the
    GPU only: ``prefer_gpu=False`` is logged and the GPU is used; without a
    ROCm GPU the constructor raises (there is no CPU path)."""
    if not prefer_gpu:
        logger.warning("prefer_gpu=False: the Kirchhoff migration runs on the GPU only (HIP backend)")
    opts = _OPT_LEVELS.get(optimization_level, _OPT_LEVELS["high"])
    return OptimizedKirchhoffMigrator(velocity, config, **opts)
