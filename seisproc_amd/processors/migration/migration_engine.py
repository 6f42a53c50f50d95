"""MigrationEngine on the HIP backend (reference processors/migration/migration_engine.py:34-876).

``migrate_bin_full`` keeps the reference's keyword arguments and return value
((image, fold) numpy arrays of shape (n_times, n_il, n_xl)):
  * grid x = origin_x + il * il_spacing, y = origin_y + xl * xl_spacing with
    the azimuth ignored, as the reference does (:292-301);
  * output times t_min_ms + k * dt_ms, k < n_times (default: input length);
  * optional fold normalisation with normalize_by_fold semantics (:347-348).
Extensions (keyword-only): an RMS velocity model (v_rms(t) or a 3D v_rms
grid) instead of the constant velocity_mps, an anisotropy model (per-leg
anelliptic eta), BF16/FP16 trace storage, the triangle antialias filter and
the weight mode.  Coordinates stay FP64 until the on-GPU reduction to local
coordinates (the reference casts them to FP32 first, :286-290).

There is no CPU path: constructing the engine without a ROCm GPU raises.
"""
from __future__ import annotations

import logging
import time
from typing import Any, Callable, Dict, Optional, Tuple

import numpy as np
import torch

from ...abi_types import (DTYPE_CODES, FLAG_AA, FLAG_KAHAN, FLAG_TIME_APERTURE, FLAG_VTI, VEL_1D,
                          WEIGHT_CODES, PstmParams)
from .kirchhoff_kernel import normalize_by_fold

logger = logging.getLogger(__name__)

_TORCH_DTYPE = {"f32": torch.float32, "bf16": torch.bfloat16, "f16": torch.float16}


def _dtype_key(trace_dtype) -> str:
    if isinstance(trace_dtype, torch.dtype):
        return {torch.float32: "f32", torch.bfloat16: "bf16", torch.float16: "f16"}[trace_dtype]
    k = str(trace_dtype).lower()
    code = DTYPE_CODES.get(k)
    if code is None:
        raise ValueError(f"unknown trace dtype {trace_dtype!r}")
    return ["f32", "bf16", "f16"][code]


class MigrationEngine:
    """Kirchhoff PSTM engine.  device: a ROCm GPU (default: current)."""

    def __init__(self, device: Optional[torch.device] = None):
        if not torch.cuda.is_available():
            raise RuntimeError("MigrationEngine needs a ROCm GPU; the migration path has no CPU fallback")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if self.device.type != "cuda":
            raise ValueError(f"MigrationEngine runs on the GPU only (got {self.device})")
        from ..kernel_backend import load_library
        load_library()
        logger.info("MigrationEngine initialized on %s", self.device)

    # ------------------------------------------------------------------
    def _vel_table(self, p: PstmParams, velocity_mps: float, velocity_model) -> torch.Tensor:
        if velocity_model is None:
            return torch.full((p.n_tau,), float(velocity_mps), dtype=torch.float32, device=self.device)
        if isinstance(velocity_model, torch.Tensor):
            v = velocity_model.to(self.device, torch.float32).contiguous()
            if v.shape not in ((p.n_tau,), (p.n_il, p.n_xl, p.n_tau)):
                raise ValueError(f"velocity grid must be ({p.n_tau},) or {(p.n_il, p.n_xl, p.n_tau)}")
            return v
        if hasattr(velocity_model, "vrms_on_grid"):          # models.VelocityModel
            return velocity_model.vrms_on_grid(p, device=self.device)
        if hasattr(velocity_model, "vrms_table"):            # engine-side VelocityModel
            t_ms = (p.tau0 + p.dtau * np.arange(p.n_tau)) * 1000.0
            return velocity_model.vrms_table(t_ms, device=self.device)
        raise TypeError(f"unsupported velocity model {type(velocity_model)}")

    def build_params(self, n_traces: int, n_samples: int, origin_x: float, origin_y: float, il_spacing: float,
                     xl_spacing: float, n_il: int, n_xl: int, dt_ms: float, t_min_ms: float,
                     n_times: Optional[int], max_aperture_m: float, max_angle_deg: float,
                     use_time_dependent_aperture: bool = False, antialias: bool = False,
                     vti: bool = False, kahan: bool = True, weight_mode: str = "full",
                     aa_max: float = 16.0) -> PstmParams:
        flags = (FLAG_TIME_APERTURE if use_time_dependent_aperture else 0) | (FLAG_AA if antialias else 0) \
            | (FLAG_VTI if vti else 0) | (FLAG_KAHAN if kahan else 0)
        return PstmParams.from_dict(dict(
            n_traces=int(n_traces), n_samples=int(n_samples), n_il=int(n_il), n_xl=int(n_xl),
            n_tau=int(n_times if n_times is not None else n_samples), dt=dt_ms / 1000.0, t0=t_min_ms / 1000.0,
            dtau=dt_ms / 1000.0, tau0=t_min_ms / 1000.0, origin_x=float(origin_x), origin_y=float(origin_y),
            il_dx=float(il_spacing), il_dy=0.0, xl_dx=0.0, xl_dy=float(xl_spacing),
            aperture_m=float(max_aperture_m), max_angle_deg=float(max_angle_deg), flags=flags,
            weight_mode=WEIGHT_CODES[weight_mode], vel_mode=VEL_1D, aa_max=aa_max))

    def migrate_bin_full(self, traces: np.ndarray, source_x: np.ndarray, source_y: np.ndarray,
                         receiver_x: np.ndarray, receiver_y: np.ndarray, origin_x: float, origin_y: float,
                         il_spacing: float, xl_spacing: float, azimuth_deg: float, n_il: int, n_xl: int,
                         dt_ms: float, t_min_ms: float, n_times: Optional[int] = None,
                         velocity_mps: float = 2500.0, max_aperture_m: float = 5000.0,
                         max_angle_deg: float = 60.0, normalize: bool = True, min_fold: int = 1,
                         progress_callback: Optional[Callable[[float, str], None]] = None,
                         enable_profiling: bool = False, use_time_dependent_aperture: bool = False, *,
                         velocity_model=None, anisotropy_model=None, trace_dtype="f32",
                         antialias: bool = False, kahan: bool = True, weight_mode: str = "full",
                         return_tensors: bool = False) -> Tuple[Any, Any]:
        """Full Kirchhoff migration of one bin.  traces: (n_samples, n_traces)."""
        t0 = time.perf_counter()
        traces_t = torch.as_tensor(traces)
        if traces_t.dim() != 2:
            raise ValueError(f"traces must be (n_samples, n_traces), got {tuple(traces_t.shape)}")
        n_samples, n_traces = traces_t.shape
        if progress_callback:
            progress_callback(0.0, "Preparing data...")
        p = self.build_params(n_traces, n_samples, origin_x, origin_y, il_spacing, xl_spacing, n_il, n_xl,
                              dt_ms, t_min_ms, n_times, max_aperture_m, max_angle_deg,
                              use_time_dependent_aperture, antialias, anisotropy_model is not None, kahan,
                              weight_mode)
        from .hip_ops import PstmPlan, quantize_traces
        src = np.stack([np.asarray(source_x, np.float64), np.asarray(source_y, np.float64)], 1)
        rcv = np.stack([np.asarray(receiver_x, np.float64), np.asarray(receiver_y, np.float64)], 1)
        plan = PstmPlan(p, src, rcv, device=self.device)
        tr = traces_t.to(dtype=torch.float32).t().contiguous().to(self.device, non_blocking=False)
        scale = None
        key = _dtype_key(trace_dtype)
        if key != "f32":
            tr, scale = quantize_traces(tr, _TORCH_DTYPE[key])
        vel = self._vel_table(p, velocity_mps, velocity_model)
        eta = None
        if anisotropy_model is not None:
            eta = anisotropy_model.eta_on_grid(p, device=self.device, full=vel.dim() == 3)
        if progress_callback:
            progress_callback(0.1, "Running Kirchhoff migration...")
        image = torch.zeros(plan.image_shape, dtype=torch.float32, device=self.device)
        comp = torch.zeros_like(image) if kahan else None
        fold = torch.zeros_like(image)
        band_cb = None
        if progress_callback:
            band_cb = lambda f, m: progress_callback(0.1 + 0.8 * f, m)  # noqa: E731
        plan.forward(tr, vel, scale=scale, eta=eta, image=image, comp=comp, fold=fold, progress_callback=band_cb)
        if comp is not None:
            image += comp
        image = image.permute(2, 0, 1).contiguous()
        fold = fold.permute(2, 0, 1).contiguous()
        if progress_callback:
            progress_callback(0.9, "Normalizing...")
        if normalize:
            image = normalize_by_fold(image, fold, min_fold)
        if enable_profiling:
            torch.cuda.synchronize()
            logger.info("migrate_bin_full: %d traces -> %dx%dx%d in %.3f s", n_traces, n_il, n_xl, p.n_tau,
                        time.perf_counter() - t0)
        if progress_callback:
            progress_callback(1.0, "Complete")
        if return_tensors:
            return image, fold
        return image.cpu().numpy(), fold.cpu().numpy()

    def migrate_bin_batched(self, traces: np.ndarray, source_x, source_y, receiver_x, receiver_y,
                            origin_x: float, origin_y: float, il_spacing: float, xl_spacing: float,
                            azimuth_deg: float, n_il: int, n_xl: int, dt_ms: float, t_min_ms: float,
                            velocity_mps: float, max_aperture_m: float = 5000.0, max_angle_deg: float = 60.0,
                            normalize: bool = True, min_fold: int = 1, batch_size: int = 10000,
                            progress_callback: Optional[Callable[[float, str], None]] = None):
        """Full Kirchhoff migration in trace batches accumulated into one image
        (compensated across batches)."""
        traces = np.asarray(traces)
        n_samples, n_traces = traces.shape
        p = self.build_params(n_traces, n_samples, origin_x, origin_y, il_spacing, xl_spacing, n_il, n_xl,
                              dt_ms, t_min_ms, None, max_aperture_m, max_angle_deg)
        from .hip_ops import PstmPlan
        vel = torch.full((p.n_tau,), float(velocity_mps), dtype=torch.float32, device=self.device)
        shape = (p.n_il, p.n_xl, p.n_tau)
        image = torch.zeros(shape, dtype=torch.float32, device=self.device)
        comp = torch.zeros_like(image)
        fold = torch.zeros_like(image)
        n_batches = max(1, (n_traces + batch_size - 1) // batch_size)
        for b in range(n_batches):
            a, e = b * batch_size, min(n_traces, (b + 1) * batch_size)
            if a >= e:
                break
            pb = p.replace(n_traces=e - a)
            src = np.stack([np.asarray(source_x[a:e], np.float64), np.asarray(source_y[a:e], np.float64)], 1)
            rcv = np.stack([np.asarray(receiver_x[a:e], np.float64), np.asarray(receiver_y[a:e], np.float64)], 1)
            plan = PstmPlan(pb, src, rcv, device=self.device)
            tr = torch.as_tensor(np.ascontiguousarray(traces[:, a:e].T), dtype=torch.float32).to(self.device)
            plan.forward(tr, vel, image=image, comp=comp, fold=fold)
            if progress_callback:
                progress_callback((b + 1) / n_batches, f"Batch {b + 1}/{n_batches}")
        image = (image + comp).permute(2, 0, 1).contiguous()
        fold = fold.permute(2, 0, 1).contiguous()
        if normalize:
            image = normalize_by_fold(image, fold, min_fold)
        return image.cpu().numpy(), fold.cpu().numpy()

    # The reference's other entry points are not Kirchhoff (DSR) migrations.
    def migrate_bin(self, *args, **kwargs):
        raise NotImplementedError(
            "the reference's migrate_bin (migration_engine.py:77) adds each trace to its own midpoint "
            "column only (an NMO-like stack, not a Kirchhoff sum); use migrate_bin_full")

    def migrate_bin_time_domain(self, *args, **kwargs):
        raise NotImplementedError(
            "migrate_bin_time_domain (migration_engine.py:467) uses t_out = sqrt(t_in^2 + 4h^2/v^2), the "
            "modelling direction; use migrate_bin_full")

    def migrate_bin_time_domain_rms(self, *args, **kwargs):
        raise NotImplementedError(
            "migrate_bin_time_domain_rms (migration_engine.py:633) shares the inverted time mapping; use "
            "migrate_bin_full(velocity_model=...) for RMS velocities")

    def benchmark(self, n_traces: int = 10000, n_samples: int = 1501, n_il: int = 100, n_xl: int = 100,
                  seed: int = 42) -> Dict[str, float]:
        """Random zero-offset traces over the grid (reference :804-876), full
        Kirchhoff migration; returns timing and Gcontrib/s."""
        rng = np.random.default_rng(seed)
        traces = rng.standard_normal((n_samples, n_traces)).astype(np.float32)
        spacing, origin = 50.0, 2500.0
        hi = origin + (n_il - 1) * spacing
        sx = rng.uniform(origin, hi, n_traces)
        sy = rng.uniform(origin, hi, n_traces)
        t0 = time.perf_counter()
        image, fold = self.migrate_bin_full(traces, sx, sy, sx, sy, origin, origin, spacing, spacing, 0.0, n_il,
                                            n_xl, 2.0, 0.0, velocity_mps=2500.0, normalize=False,
                                            return_tensors=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        contrib = float(fold.double().sum())
        return {"total_time_s": dt, "traces_per_second": n_traces / dt, "contributions": contrib,
                "gcontrib_per_second": contrib / dt / 1e9}
