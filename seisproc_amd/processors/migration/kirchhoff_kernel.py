"""Functional Kirchhoff API of the reference (processors/migration/kirchhoff_kernel.py)
on the HIP backend.

``migrate_kirchhoff_full`` keeps the reference signature and output layout
(image and fold as (n_depths, n_il, n_xl) tensors on the traces' device)
and runs the gfx950 forward kernel.  The reference evaluates any list of
output points; the kernel works on a regular grid, so the il-major
output_x/output_y the engine builds (migration_engine.py:292-301) are checked
to be one and the depth axis to be regular in time (tau = 2 z / v).
``normalize_by_fold`` is the reference's semantics (:437-456): image/fold
where fold >= min_fold, else unchanged.
"""
from __future__ import annotations

import logging
import math
import time
from typing import Callable, Optional, Tuple

import numpy as np
import torch

from ...abi_types import FLAG_TIME_APERTURE, VEL_1D, WEIGHT_FULL, PstmParams

logger = logging.getLogger(__name__)


def normalize_by_fold(image: torch.Tensor, fold: torch.Tensor, min_fold: int = 1) -> torch.Tensor:
    out = image.clone()
    mask = fold >= min_fold
    out[mask] = image[mask] / fold[mask]
    return out


def _grid_from_points(output_x: torch.Tensor, output_y: torch.Tensor, n_il: int, n_xl: int):
    ox = output_x.detach().double().cpu().numpy().reshape(n_il, n_xl)
    oy = output_y.detach().double().cpu().numpy().reshape(n_il, n_xl)
    x0, y0 = float(ox[0, 0]), float(oy[0, 0])
    il = (float(ox[1, 0] - x0), float(oy[1, 0] - y0)) if n_il > 1 else None
    xl = (float(ox[0, 1] - x0), float(oy[0, 1] - y0)) if n_xl > 1 else None
    if il is None and xl is None:
        il, xl = (1.0, 0.0), (0.0, 1.0)
    elif il is None:
        il = (-xl[1], xl[0])
    elif xl is None:
        xl = (-il[1], il[0])
    i = np.arange(n_il)[:, None]
    j = np.arange(n_xl)[None, :]
    rx = x0 + i * il[0] + j * xl[0]
    ry = y0 + i * il[1] + j * xl[1]
    scale = max(1.0, float(np.abs(ox).max()), float(np.abs(oy).max()))
    if not (np.allclose(rx, ox, atol=1e-6 * scale + 1e-3) and np.allclose(ry, oy, atol=1e-6 * scale + 1e-3)):
        raise ValueError("the HIP backend migrates onto a regular il-major output grid; output_x/output_y "
                         "are not one (build them as in MigrationEngine.migrate_bin_full)")
    return x0, y0, il, xl


def migrate_kirchhoff_full(traces: torch.Tensor, source_x: torch.Tensor, source_y: torch.Tensor,
                           receiver_x: torch.Tensor, receiver_y: torch.Tensor, output_x: torch.Tensor,
                           output_y: torch.Tensor, depth_axis: torch.Tensor, velocity: float, dt_ms: float,
                           t_min_ms: float, max_aperture_m: float, max_angle_deg: float, n_il: int, n_xl: int,
                           progress_callback: Optional[Callable[[float, str], None]] = None,
                           enable_profiling: bool = False,
                           use_time_dependent_aperture: bool = False) -> Tuple[torch.Tensor, torch.Tensor]:
    """Full Kirchhoff migration on the GPU.  traces: (n_samples, n_traces)."""
    from .hip_ops import PstmPlan

    n_samples, n_traces = traces.shape
    z = depth_axis.detach().double().cpu().numpy()
    tau = 2.0 * z / float(velocity)
    n_tau = len(tau)
    dtau = float(tau[-1] - tau[0]) / (n_tau - 1) if n_tau > 1 else 1.0
    if n_tau > 2 and np.abs(tau - (tau[0] + dtau * np.arange(n_tau))).max() > 1e-3 * abs(dtau):
        raise ValueError("depth_axis must be regular in two-way time (z = v t / 2)")
    x0, y0, il, xl = _grid_from_points(output_x, output_y, n_il, n_xl)
    p = PstmParams.from_dict(dict(
        n_traces=n_traces, n_samples=n_samples, n_il=n_il, n_xl=n_xl, n_tau=n_tau, dt=dt_ms / 1000.0,
        t0=t_min_ms / 1000.0, dtau=dtau if dtau > 0 else 1.0, tau0=float(tau[0]), origin_x=x0, origin_y=y0,
        il_dx=il[0], il_dy=il[1], xl_dx=xl[0], xl_dy=xl[1], aperture_m=max_aperture_m,
        max_angle_deg=max_angle_deg, flags=FLAG_TIME_APERTURE if use_time_dependent_aperture else 0,
        weight_mode=WEIGHT_FULL, vel_mode=VEL_1D))
    dev = traces.device if traces.device.type == "cuda" else torch.device("cuda")
    t0 = time.perf_counter()
    if progress_callback:
        progress_callback(0.0, "Kirchhoff migration on the GPU")
    src = torch.stack([source_x.detach().double(), source_y.detach().double()], 1)
    rcv = torch.stack([receiver_x.detach().double(), receiver_y.detach().double()], 1)
    plan = PstmPlan(p, src, rcv, device=dev)
    tr = traces.detach().to(device=dev, dtype=torch.float32).t().contiguous()
    vel = torch.full((n_tau,), float(velocity), dtype=torch.float32, device=dev)
    band_cb = None
    if progress_callback:
        band_cb = lambda f, m: progress_callback(0.99 * f, m)  # noqa: E731
    img, fold = plan.forward(tr, vel, progress_callback=band_cb)
    image = img.permute(2, 0, 1).contiguous().to(traces.device)
    fold = fold.permute(2, 0, 1).contiguous().to(traces.device)
    if enable_profiling or progress_callback:
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        contrib = float(fold.double().sum())
        msg = f"{contrib:.3e} contributions in {dt:.3f} s ({contrib / max(dt, 1e-12) / 1e9:.1f} Gcontrib/s)"
        if enable_profiling:
            logger.info("migrate_kirchhoff_full: %s", msg)
        if progress_callback:
            progress_callback(1.0, msg)
    return image, fold


def migrate_kirchhoff_time_domain(*args, **kwargs):
    raise NotImplementedError(
        "the reference's time-domain kernel (kirchhoff_kernel.py:857-1212) maps each input sample to "
        "t_out = sqrt(t_in^2 + 4 h_eff^2 / v^2), the modelling direction with a non-DSR h_eff; it is not a "
        "migration and is not provided.  Use migrate_kirchhoff_full.")


def migrate_kirchhoff_time_domain_rms(*args, **kwargs):
    raise NotImplementedError(
        "the reference's RMS time-domain kernel (kirchhoff_kernel.py:1313-1783) shares the inverted "
        "time mapping and scatters to the trace midpoint; it is not provided.  Use migrate_kirchhoff_full "
        "with an RMS velocity (MigrationEngine.migrate_bin_full(velocity_model=...)).")
