"""FP32 torch-CPU restatement of the reference's CPU migration method --
TEST / BASELINE INFRASTRUCTURE ONLY (timed as ``cpu_baseline`` by bench.py,
checked against the FP64 oracle by tests/test_ref_torch.py).

Restates processors/migration/kirchhoff_kernel.py:459-730
(migrate_kirchhoff_full) as the reference runs it on ``device='cpu'``:
  * output points in tiles of 100 (:522), il-major (:569-570);
  * conservative midpoint pre-filter per tile, radius aperture + tile
    radius + 100 m (:579-590);
  * per tile: horizontal distances (:611-617), static aperture mask (:635);
  * depth batches sized from a 500 MB budget, at most 50 depths (:642-643);
  * per batch: r = sqrt(h^2 + z^2), atan2 angle masks, DSR time in ms,
    obliquity*spreading weight (:661-693);
  * per depth: clamped floor/frac interpolation by advanced-index gathers and
    a sum over traces (:700-730).
Constant velocity, FP32, no progress/profiling code.

``demigrate_kirchhoff_full_cpu`` is the transpose of the same loop (the
reference has no adjoint): per depth the image values times weight and mask
are scattered into the two interpolation samples of every trace.  It is
the CPU side of the least-squares iteration baseline.
"""
from __future__ import annotations

import math

import torch


def migrate_kirchhoff_full_cpu(traces: torch.Tensor, source_x, source_y, receiver_x, receiver_y,
                               output_x, output_y, depth_axis, velocity: float, dt_ms: float,
                               t_min_ms: float, max_aperture_m: float, max_angle_deg: float,
                               n_il: int, n_xl: int, tile_size: int = 100):
    """traces: (n_samples, n_traces) float32; outputs (n_depths, n_il, n_xl)."""
    n_samples, n_traces = traces.shape
    n_depths = depth_axis.shape[0]
    n_out = n_il * n_xl
    image = torch.zeros(n_depths, n_il, n_xl)
    fold = torch.zeros(n_depths, n_il, n_xl)
    theta = math.radians(max_angle_deg)
    inv_v_ms = 1000.0 / velocity
    mid_x = (source_x + receiver_x) / 2
    mid_y = (source_y + receiver_y) / 2
    for start in range(0, n_out, tile_size):
        stop = min(start + tile_size, n_out)
        n_tile = stop - start
        tx, ty = output_x[start:stop], output_y[start:stop]
        flat = torch.arange(start, stop)
        t_il, t_xl = flat // n_xl, flat % n_xl
        cx = (tx.min() + tx.max()) / 2
        cy = (ty.min() + ty.max()) / 2
        radius = torch.sqrt(((tx - cx) ** 2 + (ty - cy) ** 2).max())
        keep = torch.sqrt((mid_x - cx) ** 2 + (mid_y - cy) ** 2) < (max_aperture_m + radius + 100)
        idx = torch.where(keep)[0]
        n_f = idx.numel()
        if n_f == 0:
            continue
        sx, sy, rx, ry = source_x[idx], source_y[idx], receiver_x[idx], receiver_y[idx]
        tr = traces[:, idx]
        hs2 = (tx[:, None] - sx[None, :]) ** 2 + (ty[:, None] - sy[None, :]) ** 2
        hr2 = (tx[:, None] - rx[None, :]) ** 2 + (ty[:, None] - ry[None, :]) ** 2
        hs, hr = torch.sqrt(hs2), torch.sqrt(hr2)
        ap_static = (hs < max_aperture_m) & (hr < max_aperture_m)
        per_depth = n_tile * n_f * 4 * 8
        batch = max(1, min(50, int(500_000_000 / max(per_depth, 1))))
        col = torch.arange(n_f).unsqueeze(0).expand(n_tile, -1).reshape(-1)
        for d0 in range(0, n_depths, batch):
            d1 = min(d0 + batch, n_depths)
            z = depth_axis[d0:d1].view(-1, 1, 1)
            rs = torch.sqrt(hs2[None] + z * z)
            rr = torch.sqrt(hr2[None] + z * z)
            ang_ok = (torch.atan2(hs[None], z + 1e-6) < theta) & (torch.atan2(hr[None], z + 1e-6) < theta)
            mask = ap_static[None] & ang_ok
            t_ms = (rs + rr) * inv_v_ms
            w = (z / (rs + 1e-6)) * (z / (rr + 1e-6)) / (rs * rr + 1e-6)
            for j in range(d1 - d0):
                u = torch.clamp((t_ms[j] - t_min_ms) / dt_ms, 0, n_samples - 1.001)
                i0 = u.long()
                i1 = (i0 + 1).clamp(max=n_samples - 1)
                f = u - i0.float()
                a0 = tr[i0.reshape(-1), col].reshape(n_tile, n_f)
                a1 = tr[i1.reshape(-1), col].reshape(n_tile, n_f)
                amp = a0 + f * (a1 - a0)
                m = mask[j].float()
                image[d0 + j, t_il, t_xl] = (amp * w[j] * m).sum(1)
                fold[d0 + j, t_il, t_xl] = m.sum(1)
    return image, fold


def demigrate_kirchhoff_full_cpu(image: torch.Tensor, n_samples: int, source_x, source_y, receiver_x, receiver_y,
                                 output_x, output_y, depth_axis, velocity: float, dt_ms: float, t_min_ms: float,
                                 max_aperture_m: float, max_angle_deg: float, n_il: int, n_xl: int,
                                 tile_size: int = 100):
    """Transpose of migrate_kirchhoff_full_cpu: image (n_depths, n_il, n_xl)
    -> traces (n_samples, n_traces)."""
    n_traces = source_x.shape[0]
    n_depths = depth_axis.shape[0]
    n_out = n_il * n_xl
    out = torch.zeros(n_samples * n_traces)
    theta = math.radians(max_angle_deg)
    inv_v_ms = 1000.0 / velocity
    mid_x = (source_x + receiver_x) / 2
    mid_y = (source_y + receiver_y) / 2
    for start in range(0, n_out, tile_size):
        stop = min(start + tile_size, n_out)
        n_tile = stop - start
        tx, ty = output_x[start:stop], output_y[start:stop]
        flat = torch.arange(start, stop)
        t_il, t_xl = flat // n_xl, flat % n_xl
        cx = (tx.min() + tx.max()) / 2
        cy = (ty.min() + ty.max()) / 2
        radius = torch.sqrt(((tx - cx) ** 2 + (ty - cy) ** 2).max())
        keep = torch.sqrt((mid_x - cx) ** 2 + (mid_y - cy) ** 2) < (max_aperture_m + radius + 100)
        idx = torch.where(keep)[0]
        n_f = idx.numel()
        if n_f == 0:
            continue
        sx, sy, rx, ry = source_x[idx], source_y[idx], receiver_x[idx], receiver_y[idx]
        hs2 = (tx[:, None] - sx[None, :]) ** 2 + (ty[:, None] - sy[None, :]) ** 2
        hr2 = (tx[:, None] - rx[None, :]) ** 2 + (ty[:, None] - ry[None, :]) ** 2
        hs, hr = torch.sqrt(hs2), torch.sqrt(hr2)
        ap_static = (hs < max_aperture_m) & (hr < max_aperture_m)
        per_depth = n_tile * n_f * 4 * 8
        batch = max(1, min(50, int(500_000_000 / max(per_depth, 1))))
        col = idx.unsqueeze(0).expand(n_tile, -1).reshape(-1)
        for d0 in range(0, n_depths, batch):
            d1 = min(d0 + batch, n_depths)
            z = depth_axis[d0:d1].view(-1, 1, 1)
            rs = torch.sqrt(hs2[None] + z * z)
            rr = torch.sqrt(hr2[None] + z * z)
            ang_ok = (torch.atan2(hs[None], z + 1e-6) < theta) & (torch.atan2(hr[None], z + 1e-6) < theta)
            mask = ap_static[None] & ang_ok
            t_ms = (rs + rr) * inv_v_ms
            w = (z / (rs + 1e-6)) * (z / (rr + 1e-6)) / (rs * rr + 1e-6)
            for j in range(d1 - d0):
                u = torch.clamp((t_ms[j] - t_min_ms) / dt_ms, 0, n_samples - 1.001)
                i0 = u.long()
                i1 = (i0 + 1).clamp(max=n_samples - 1)
                f = u - i0.float()
                g = image[d0 + j, t_il, t_xl][:, None] * w[j] * mask[j].float()
                out.index_add_(0, (i0 * n_traces + col.view(n_tile, n_f)).reshape(-1), (g * (1 - f)).reshape(-1))
                out.index_add_(0, (i1 * n_traces + col.view(n_tile, n_f)).reshape(-1), (g * f).reshape(-1))
    return out.view(n_samples, n_traces)
