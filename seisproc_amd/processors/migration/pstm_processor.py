"""PSTM as a processor: ``PSTMProcessor(BaseProcessor).process(SeismicData)``.

The reference registers a migrator in its processor registry
(processors/__init__.py:51,74) whose class lacks ``process``
(base_migrator.py:24), so ``BaseProcessor.from_dict`` could not build it.  This
module supplies the processor the registry implies:

  * input: SeismicData with traces (n_samples, n_traces), sample_rate in ms
    and SEG-Y style coordinate headers (SourceX/SourceY/GroupX/GroupY by
    default, scaled by coordinate_scalar; models/migration_geometry.py:364-394);
  * output: a new SeismicData whose traces are the migrated image il-major,
    (n_time, n_inline * n_xline) (the layout of
    utils/migration_output_storage.py:236-238), sample_rate = the output
    grid's dt, inline/crossline/CDP_X/CDP_Y headers, and the fold (same
    layout), contribution count and configuration in metadata.

``migrate_arrays`` is the shared host driver (also used by
KirchhoffMigrator): offset-range selection, trace upload (FP64 coordinates,
optional BF16/FP16 quantisation on the GPU), v_rms / eta on the image grid,
one plan and one forward launch accumulating into caller-held tensors.
"""
from __future__ import annotations

import logging
from typing import Any, Dict, Optional, Tuple

import numpy as np
import torch

from ...abi_types import FLAG_VTI
from ...models.anisotropy_model import AnisotropyModel
from ...models.migration_config import InterpolationMode, MigrationConfig
from ...models.migration_geometry import MigrationGeometry
from ...models.seismic_data import SeismicData
from ...models.velocity_model import VelocityModel
from ..base_processor import BaseProcessor

logger = logging.getLogger(__name__)
_TORCH_DTYPE = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16}


def _as_config(c) -> MigrationConfig:
    if isinstance(c, MigrationConfig):
        return c
    if isinstance(c, dict):
        return MigrationConfig.from_dict(dict(c))
    raise TypeError(f"config must be a MigrationConfig or dict, got {type(c)}")


def _as_velocity(v) -> VelocityModel:
    if isinstance(v, VelocityModel):
        return v
    if isinstance(v, (int, float)):
        return VelocityModel(data=float(v))
    if isinstance(v, dict):
        return VelocityModel.from_dict(dict(v))
    raise TypeError(f"velocity must be a VelocityModel, number or dict, got {type(v)}")


def _as_anisotropy(a) -> Optional[AnisotropyModel]:
    if a is None or isinstance(a, AnisotropyModel):
        return a
    if isinstance(a, dict):
        return AnisotropyModel.from_dict(dict(a))
    raise TypeError(f"anisotropy must be an AnisotropyModel or dict, got {type(a)}")


def migrate_arrays(config: MigrationConfig, velocity: VelocityModel, traces, sample_rate_ms: float,
                   geometry: MigrationGeometry, *, anisotropy: Optional[AnisotropyModel] = None,
                   trace_dtype: str = "f32", image: Optional[torch.Tensor] = None,
                   comp: Optional[torch.Tensor] = None, fold: Optional[torch.Tensor] = None,
                   device=None, honor_azimuth: bool = False, kahan: bool = True,
                   time_dependent_aperture: bool = False, taper: bool = False, zero_outside: bool = False,
                   aa_scale: float = 1.0, aa_max: float = 16.0, input_t0_s: float = 0.0):
    """Migrate traces (n_samples, n_traces) into [il][xl][tau] accumulators.
    Returns (image, comp, fold, n_used_traces)."""
    from .hip_ops import PstmPlan, quantize_traces
    if not torch.cuda.is_available():
        raise RuntimeError("migration needs a ROCm GPU; there is no CPU fallback")
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    g = config.output_grid
    off = geometry.offset
    keep = np.nonzero((off >= config.min_offset_m) & (off <= config.max_offset_m))[0]
    geo = geometry if len(keep) == geometry.n_traces else geometry.select(keep)
    tr = torch.as_tensor(traces)
    if tr.dim() != 2 or tr.shape[1] != geometry.n_traces:
        raise ValueError(f"traces must be (n_samples, {geometry.n_traces}), got {tuple(tr.shape)}")
    if len(keep) != geometry.n_traces:
        tr = tr[:, torch.as_tensor(keep, device=tr.device)]
    n_samples = tr.shape[0]
    if config.interpolation_mode is InterpolationMode.SINC:
        logger.info("interpolation_mode=SINC requested; the Kirchhoff sum reads traces with linear "
                    "interpolation as the reference's kernels do (kirchhoff_kernel.py:708-721)")
    p = config.to_pstm_params(len(keep), n_samples, sample_rate_ms / 1000.0, input_t0_s,
                              honor_azimuth=honor_azimuth, kahan=kahan, aa_scale=aa_scale, aa_max=aa_max,
                              zero_outside=zero_outside, time_dependent_aperture=time_dependent_aperture,
                              taper=taper)
    if (p.flags & FLAG_VTI) and anisotropy is None:
        raise ValueError("the configuration asks for VTI traveltimes but no anisotropy model was given")
    shape = (p.n_il, p.n_xl, p.n_tau)
    if image is None:
        image = torch.zeros(shape, dtype=torch.float32, device=dev)
    if comp is None and kahan:
        comp = torch.zeros(shape, dtype=torch.float32, device=dev)
    if fold is None:
        fold = torch.zeros(shape, dtype=torch.float32, device=dev)
    if len(keep) == 0:
        return image, comp, fold, 0
    plan = PstmPlan(p, geo.src_xy(), geo.rcv_xy(), device=dev)
    trd = tr.to(dtype=torch.float32).t().contiguous().to(dev)
    scale = None
    if trace_dtype != "f32":
        trd, scale = quantize_traces(trd, _TORCH_DTYPE[trace_dtype])
    vel = velocity.vrms_on_grid(p, device=dev)
    eta = anisotropy.eta_on_grid(p, device=dev, full=vel.dim() == 3) if (p.flags & FLAG_VTI) else None
    plan.forward(trd, vel, scale=scale, eta=eta, image=image, comp=comp, fold=fold)
    return image, comp, fold, len(keep)


class PSTMProcessor(BaseProcessor):
    """Kirchhoff prestack time migration of a SeismicData gather/dataset.

    Parameters (all serialisable):
      config        MigrationConfig dict (or object)
      velocity      VelocityModel dict, object or a constant in m/s
      anisotropy    AnisotropyModel dict/object (VTI), optional
      trace_dtype   'f32' | 'bf16' | 'f16' storage of traces on the GPU
      header_keys   (sx, sy, gx, gy) header names
      coordinate_scalar, honor_azimuth, kahan, time_dependent_aperture,
      taper, zero_outside, aa_scale, aa_max
    """

    DEFAULTS = dict(anisotropy=None, trace_dtype="f32", header_keys=("SourceX", "SourceY", "GroupX", "GroupY"),
                    coordinate_scalar=1.0, honor_azimuth=False, kahan=True, time_dependent_aperture=False,
                    taper=False, zero_outside=False, aa_scale=1.0, aa_max=16.0)

    def _validate_params(self):
        if "config" not in self.params or "velocity" not in self.params:
            raise ValueError("PSTMProcessor needs 'config' and 'velocity'")
        cfg = self.params["config"]
        if isinstance(cfg, MigrationConfig):
            self.params["config"] = cfg.to_dict()
        vel = self.params["velocity"]
        if isinstance(vel, VelocityModel):
            self.params["velocity"] = vel.to_dict()
        an = self.params.get("anisotropy")
        if isinstance(an, AnisotropyModel):
            self.params["anisotropy"] = an.to_dict()
        for k, v in self.DEFAULTS.items():
            self.params.setdefault(k, v)
        if self.params["trace_dtype"] not in _TORCH_DTYPE:
            raise ValueError(f"trace_dtype must be one of {sorted(_TORCH_DTYPE)}")
        self.config = _as_config(self.params["config"])
        self.velocity = _as_velocity(self.params["velocity"])
        self.anisotropy = _as_anisotropy(self.params.get("anisotropy"))
        if len(tuple(self.params["header_keys"])) != 4:
            raise ValueError("header_keys must name (sx, sy, gx, gy)")

    def get_description(self) -> str:
        g = self.config.output_grid
        return (f"Kirchhoff PSTM (HIP/gfx950): {g.n_inline}x{g.n_xline}x{g.n_time} image, aperture "
                f"{self.config.max_aperture_m} m, max angle {self.config.max_angle_deg} deg, weights "
                f"{self.config.weight_mode.value}, antialias {'on' if self.config.antialias_enabled else 'off'}, "
                f"traces {self.params['trace_dtype']}")

    def geometry_of(self, data: SeismicData) -> MigrationGeometry:
        if not data.headers:
            raise ValueError("PSTMProcessor needs source/receiver coordinate headers")
        sx, sy, gx, gy = self.params["header_keys"]
        missing = [k for k in (sx, sy, gx, gy) if k not in data.headers]
        if missing:
            raise ValueError(f"missing coordinate headers: {missing}")
        return MigrationGeometry.from_headers(data.headers, sx, sy, gx, gy, self.params["coordinate_scalar"])

    def migrate(self, data: SeismicData):
        """(image, fold) torch tensors [il][xl][tau] on the GPU (unnormalised)."""
        pp = self.params
        self._report_progress(0, 1, "migrating")
        image, comp, fold, used = migrate_arrays(
            self.config, self.velocity, data.traces, data.sample_rate, self.geometry_of(data),
            anisotropy=self.anisotropy, trace_dtype=pp["trace_dtype"], honor_azimuth=pp["honor_azimuth"],
            kahan=pp["kahan"], time_dependent_aperture=pp["time_dependent_aperture"], taper=pp["taper"],
            zero_outside=pp["zero_outside"], aa_scale=pp["aa_scale"], aa_max=pp["aa_max"])
        if comp is not None:
            image = image + comp
        self._report_progress(1, 1, f"migrated {used} traces")
        return image, fold

    def process(self, data: SeismicData) -> SeismicData:
        image, fold = self.migrate(data)
        cfg = self.config
        g = cfg.output_grid
        if cfg.normalize_by_fold:
            image = torch.where(fold >= cfg.min_fold, image / torch.clamp(fold, min=1.0), torch.zeros_like(image))
        img = image.permute(2, 0, 1).reshape(g.n_time, g.n_inline * g.n_xline).cpu().numpy()
        fld = fold.permute(2, 0, 1).reshape(g.n_time, g.n_inline * g.n_xline).cpu().numpy()
        il = np.repeat(g.inline_axis, g.n_xline)
        xl = np.tile(g.xline_axis, g.n_inline)
        (ilx, ily), (xlx, xly) = g.step_vectors(self.params["honor_azimuth"])
        ii = np.repeat(np.arange(g.n_inline), g.n_xline)
        jj = np.tile(np.arange(g.n_xline), g.n_inline)
        headers = {"inline": il.astype(np.int32), "crossline": xl.astype(np.int32),
                   "CDP_X": g.x_origin + ii * ilx + jj * xlx, "CDP_Y": g.y_origin + ii * ily + jj * xly}
        meta = dict(data.metadata)
        meta.update({"migration": "kirchhoff_pstm", "image_shape": (g.n_time, g.n_inline, g.n_xline),
                     "fold": fld, "contributions": float(fld.astype(np.float64).sum()),
                     "fold_normalized": bool(cfg.normalize_by_fold), "migration_config": cfg.to_dict()})
        return SeismicData(traces=img.astype(np.float32), sample_rate=g.dt * 1000.0, headers=headers, metadata=meta)
