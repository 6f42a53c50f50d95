"""Seeded synthetic surveys for the PSTM workloads.

Traces are sums of Ricker wavelets placed at exact traveltimes of point
diffractors (double-square-root time with the RMS velocity at the
diffractor), flat reflectors (hyperbolic moveout at the midpoint) and dipping
planes (zero-offset time linear in the midpoint, hyperbolic offset moveout,
the NMO approximation of tests/fixtures/synthetic_prestack.py:370-378 of the
reference), plus optional Gaussian noise.  The wavelet is
(1 - 2 pi^2 f^2 t^2) exp(-pi^2 f^2 t^2) (synthetic_prestack.py:28-46).

Everything is generated with torch on the requested device (GPU for the
full-size workloads, CPU for tests) from ``torch.Generator`` seeds, so a
workload is reproducible from (name, seed, sizes).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np
import torch

from .abi_types import FLAG_AA, FLAG_KAHAN, FLAG_VTI, PstmParams


def ricker(t: torch.Tensor, f: float) -> torch.Tensor:
    a = (math.pi * f * t) ** 2
    return (1.0 - 2.0 * a) * torch.exp(-a)


@dataclass
class SyntheticSurvey:
    name: str
    params: PstmParams           # image grid + options (n_traces/n_samples filled)
    src_xy: torch.Tensor         # (n, 2) float64 world coordinates
    rcv_xy: torch.Tensor         # (n, 2) float64
    traces: torch.Tensor         # (n, n_samples) float32, trace-major
    vel: torch.Tensor            # (n_tau,) or (n_il, n_xl, n_tau) float32 v_rms on the image grid
    eta: Optional[torch.Tensor] = None
    class_sizes: List[int] = field(default_factory=list)
    info: Dict = field(default_factory=dict)

    @property
    def n_traces(self) -> int:
        return int(self.params.n_traces)

    def numpy_geometry(self):
        s = self.src_xy.double().cpu().numpy()
        r = self.rcv_xy.double().cpu().numpy()
        return s[:, 0], s[:, 1], r[:, 0], r[:, 1]


# ---------------------------------------------------------------------------
# velocity fields

class VrmsField:
    """v_rms(x, y, tau) = (v0 + k tau) * (1 + amp * g(x, y)), g in [-1, 1] a
    sum of Gaussian bumps; x, y world coordinates, tau two-way time (s)."""

    def __init__(self, v0: float, k: float = 0.0, amp: float = 0.0, bumps=None):
        self.v0, self.k, self.amp = float(v0), float(k), float(amp)
        self.bumps = list(bumps or [])  # (x, y, sigma, sign)

    def lateral(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        if self.amp == 0.0 or not self.bumps:
            return torch.ones_like(x)
        g = torch.zeros_like(x)
        for bx, by, sg, sign in self.bumps:
            g = g + sign * torch.exp(-((x - bx) ** 2 + (y - by) ** 2) / (2 * sg * sg))
        g = g.clamp(-1.0, 1.0)
        return 1.0 + self.amp * g

    def __call__(self, x: torch.Tensor, y: torch.Tensor, tau: torch.Tensor) -> torch.Tensor:
        return (self.v0 + self.k * tau) * self.lateral(x, y)

    @property
    def laterally_constant(self) -> bool:
        return self.amp == 0.0 or not self.bumps

    def on_grid(self, p: PstmParams, device) -> torch.Tensor:
        tau = p.tau0 + p.dtau * torch.arange(p.n_tau, dtype=torch.float64, device=device)
        if self.laterally_constant:
            return (self.v0 + self.k * tau).float()
        il = torch.arange(p.n_il, dtype=torch.float64, device=device)
        xl = torch.arange(p.n_xl, dtype=torch.float64, device=device)
        X = p.origin_x + il[:, None] * p.il_dx + xl[None, :] * p.xl_dx
        Y = p.origin_y + il[:, None] * p.il_dy + xl[None, :] * p.xl_dy
        lat = self.lateral(X, Y)
        return ((self.v0 + self.k * tau)[None, None, :] * lat[:, :, None]).float().contiguous()


def _smooth_random_field(gen: torch.Generator, n_il: int, n_xl: int, n_tau: int, lo: float, hi: float,
                         device) -> torch.Tensor:
    """Smooth random field in [lo, hi] on the image grid: a few random
    low-order cosines in each direction, normalised to the range."""
    il = torch.linspace(0, 1, n_il, dtype=torch.float64)
    xl = torch.linspace(0, 1, n_xl, dtype=torch.float64)
    tt = torch.linspace(0, 1, n_tau, dtype=torch.float64)
    f = torch.zeros(n_il, n_xl, n_tau, dtype=torch.float64)
    for _ in range(4):
        c = torch.rand(6, generator=gen, dtype=torch.float64)
        f += (torch.cos(math.pi * (1 + 2 * c[0]) * il + 6.28 * c[1])[:, None, None]
              * torch.cos(math.pi * (1 + 2 * c[2]) * xl + 6.28 * c[3])[None, :, None]
              * torch.cos(math.pi * (1 + 2 * c[4]) * tt + 6.28 * c[5])[None, None, :])
    f = (f - f.min()) / max(float(f.max() - f.min()), 1e-12)
    return (lo + (hi - lo) * f).float().to(device).contiguous()


# ---------------------------------------------------------------------------
# trace synthesis

def _leg_time(h2: torch.Tensor, tau0, v, eta=None) -> torch.Tensor:
    t2 = 0.25 * tau0 * tau0 + h2 / (v * v)
    if eta is not None:
        p = h2 / (v * v)
        t02 = 0.25 * tau0 * tau0
        t2 = t2 - 2.0 * eta * p * p / (t02 + (1.0 + 2.0 * eta) * p)
    return torch.sqrt(torch.clamp(t2, min=1e-12))


def synthesize(src: torch.Tensor, rcv: torch.Tensor, n_samples: int, dt: float, f_peak: float,
               diffractors: Optional[torch.Tensor] = None, flats: Optional[torch.Tensor] = None,
               planes: Optional[torch.Tensor] = None, vfield: Optional[VrmsField] = None,
               eta_diff: Optional[torch.Tensor] = None, noise: float = 0.0,
               gen: Optional[torch.Generator] = None, chunk: int = 32768) -> torch.Tensor:
    """Dense synthesis of (n, n_samples) float32 traces on src.device.

    diffractors: (E, 4) [x, y, tau0, amp];  flats: (F, 2) [tau0, amp];
    planes: (P, 6) [x_c, y_c, tau_c, p_x, p_y, amp] (zero-offset time
    tau_c + p_x (mx - x_c) + p_y (my - y_c)).  Times in s.
    noise: Gaussian sigma relative to the largest absolute sample.
    """
    device = src.device
    n = src.shape[0]
    t_axis = dt * torch.arange(n_samples, dtype=torch.float64, device=device)
    out = torch.zeros((n, n_samples), dtype=torch.float32, device=device)
    vfield = vfield or VrmsField(2000.0)
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        s, r = src[c0:c1], rcv[c0:c1]
        mx, my = 0.5 * (s[:, 0] + r[:, 0]), 0.5 * (s[:, 1] + r[:, 1])
        off = torch.sqrt(((s - r) ** 2).sum(1))
        arr, amp = [], []
        if diffractors is not None and len(diffractors):
            d = diffractors.to(device=device, dtype=torch.float64)
            v = vfield(d[:, 0], d[:, 1], d[:, 2])
            hs2 = (s[:, None, 0] - d[None, :, 0]) ** 2 + (s[:, None, 1] - d[None, :, 1]) ** 2
            hr2 = (r[:, None, 0] - d[None, :, 0]) ** 2 + (r[:, None, 1] - d[None, :, 1]) ** 2
            e = eta_diff.to(device=device, dtype=torch.float64)[None, :] if eta_diff is not None else None
            t = _leg_time(hs2, d[None, :, 2], v[None, :], e) + _leg_time(hr2, d[None, :, 2], v[None, :], e)
            arr.append(t)
            amp.append(d[None, :, 3].expand_as(t))
        if flats is not None and len(flats):
            fl = flats.to(device=device, dtype=torch.float64)
            v = vfield(mx[:, None], my[:, None], fl[None, :, 0])
            t = torch.sqrt(fl[None, :, 0] ** 2 + (off[:, None] / v) ** 2)
            arr.append(t)
            amp.append(fl[None, :, 1].expand_as(t))
        if planes is not None and len(planes):
            pl = planes.to(device=device, dtype=torch.float64)
            t0 = pl[None, :, 2] + pl[None, :, 3] * (mx[:, None] - pl[None, :, 0]) + pl[None, :, 4] * (my[:, None] - pl[None, :, 1])
            v = vfield(mx[:, None], my[:, None], torch.clamp(t0, min=0.0))
            t = torch.sqrt(torch.clamp(t0, min=0.0) ** 2 + (off[:, None] / v) ** 2)
            a = torch.where(t0 > 0.05, pl[None, :, 5].expand_as(t), torch.zeros_like(t))
            arr.append(t)
            amp.append(a)
        if not arr:
            continue
        T = torch.cat(arr, 1)  # (c, E)
        A = torch.cat(amp, 1)
        blk = out[c0:c1]
        for e0 in range(0, T.shape[1], 16):
            tt = T[:, e0:e0 + 16, None]
            aa = A[:, e0:e0 + 16, None]
            blk += (aa * ricker(t_axis[None, None, :] - tt, f_peak)).sum(1).float()
    if noise > 0.0:
        peak = float(out.abs().max()) if out.numel() else 0.0
        seed = int(torch.randint(0, 2**31 - 1, (1,), generator=gen)) if gen is not None else 0
        ng = torch.Generator(device=device).manual_seed(seed)
        out += (noise * peak) * torch.randn(out.shape, generator=ng, device=device, dtype=torch.float32)
    return out


# ---------------------------------------------------------------------------
# geometries

def _cmp_grid_geometry(gen: torch.Generator, nx: int, ny: int, d: float, offsets: List[float],
                       x0: float = 0.0, y0: float = 0.0, random_azimuth: bool = True):
    """One trace per (offset class, CMP bin); azimuth uniform in [0, 180) deg."""
    ix = torch.arange(nx, dtype=torch.float64)
    iy = torch.arange(ny, dtype=torch.float64)
    mx = (x0 + d * ix[:, None] + 0 * iy[None, :]).reshape(-1)
    my = (y0 + 0 * ix[:, None] + d * iy[None, :]).reshape(-1)
    S, R, sizes = [], [], []
    for o in offsets:
        if random_azimuth:
            az = math.pi * torch.rand(mx.shape[0], generator=gen, dtype=torch.float64)
        else:
            az = torch.zeros(mx.shape[0], dtype=torch.float64)
        hx, hy = 0.5 * o * torch.cos(az), 0.5 * o * torch.sin(az)
        S.append(torch.stack([mx - hx, my - hy], 1))
        R.append(torch.stack([mx + hx, my + hy], 1))
        sizes.append(mx.shape[0])
    return torch.cat(S), torch.cat(R), sizes


def _image_params(n_il, n_xl, n_tau, d_il, d_xl, n_samples, dt, aperture, angle, flags=0,
                  origin=(0.0, 0.0), n_traces=0, **kw) -> PstmParams:
    return PstmParams.from_dict(dict(
        n_traces=n_traces, n_samples=n_samples, n_il=n_il, n_xl=n_xl, n_tau=n_tau, dt=dt, t0=0.0,
        dtau=dt, tau0=0.0, origin_x=origin[0], origin_y=origin[1], il_dx=d_il, il_dy=0.0,
        xl_dx=0.0, xl_dy=d_xl, aperture_m=aperture, max_angle_deg=angle, flags=flags, **kw))


def _events_line(gen, extent_x, extent_y, t_max, n_diff=5, n_flat=3):
    dx = torch.rand(n_diff, generator=gen, dtype=torch.float64)
    dy = torch.rand(n_diff, generator=gen, dtype=torch.float64)
    dtau = torch.rand(n_diff, generator=gen, dtype=torch.float64)
    diff = torch.stack([0.15 * extent_x + 0.7 * extent_x * dx,
                        0.15 * extent_y + 0.7 * extent_y * dy if extent_y > 0 else torch.zeros(n_diff, dtype=torch.float64),
                        0.25 * t_max + 0.5 * t_max * dtau,
                        torch.ones(n_diff, dtype=torch.float64)], 1)
    flats = torch.stack([torch.linspace(0.3 * t_max, 0.85 * t_max, n_flat, dtype=torch.float64),
                         0.5 * torch.ones(n_flat, dtype=torch.float64)], 1)
    return diff, flats


def line2d_survey(seed: int = 1, n_cmp: int = 400, n_off: int = 48, n_samples: int = 1001,
               device="cpu") -> SyntheticSurvey:
    """2D line: n_cmp CMPs at 12.5 m, offsets 0:50:(50 n_off - 50) m, 1001
    samples at 2 ms, 25 Hz Ricker at 5 diffractors + 3 flat reflectors,
    v = 2000 m/s, noise sigma 0.05 of peak; image n_cmp x 1 x 1001,
    aperture 1500 m, 60 deg, triangle antialias."""
    gen = torch.Generator().manual_seed(seed)
    dt, v = 0.002, 2000.0
    cmp = 12.5 * torch.arange(n_cmp, dtype=torch.float64)
    offs = 50.0 * torch.arange(n_off, dtype=torch.float64)
    mx = cmp.repeat_interleave(n_off)
    oo = offs.repeat(n_cmp)
    src = torch.stack([mx - 0.5 * oo, torch.zeros_like(mx)], 1)
    rcv = torch.stack([mx + 0.5 * oo, torch.zeros_like(mx)], 1)
    diff, flats = _events_line(gen, 12.5 * (n_cmp - 1), 0.0, dt * (n_samples - 1))
    vf = VrmsField(v)
    tr = synthesize(src.to(device), rcv.to(device), n_samples, dt, 25.0, diff, flats, vfield=vf,
                    noise=0.05, gen=gen)
    p = _image_params(n_cmp, 1, n_samples, 12.5, 12.5, n_samples, dt, 1500.0, 60.0, FLAG_AA,
                      n_traces=src.shape[0], aa_scale=1.0, aa_max=16.0)
    return SyntheticSurvey("line2d", p, src.to(device), rcv.to(device), tr, vf.on_grid(p, device),
                           info=dict(diffractors=diff, flats=flats, v=v))


def single_offset_3d(seed: int = 2, n_bins: int = 200, n_samples: int = 1501, n_diff: int = 200,
               device="cpu", aa: bool = True) -> SyntheticSurvey:
    """3D single offset class: n_bins^2 CMP bins at 25 m, offset 500 m, random
    azimuth; 1501 samples at 2 ms; v_rms = (1500 + 600 tau)(1 +- 5 % smooth
    lateral); 30 Hz Ricker from 200 random diffractors (N(0,1) amplitudes) and
    4 dipping planes; image n_bins^2 x 1501 at 25 m, aperture 3000 m,
    triangle antialias."""
    gen = torch.Generator().manual_seed(seed)
    dt, d = 0.002, 25.0
    ext = d * (n_bins - 1)
    src, rcv, sizes = _cmp_grid_geometry(gen, n_bins, n_bins, d, [500.0])
    vf = VrmsField(1500.0, 600.0, 0.05, [(0.3 * ext, 0.4 * ext, 0.25 * ext + 1, 1.0),
                                         (0.75 * ext, 0.7 * ext, 0.2 * ext + 1, -1.0)])
    tmax = dt * (n_samples - 1)
    u = torch.rand(n_diff, 3, generator=gen, dtype=torch.float64)
    amp = torch.randn(n_diff, generator=gen, dtype=torch.float64)
    diff = torch.stack([ext * u[:, 0], ext * u[:, 1], 0.1 * tmax + 0.8 * tmax * u[:, 2], amp], 1)
    pu = torch.rand(4, 3, generator=gen, dtype=torch.float64)
    planes = torch.stack([torch.full((4,), 0.5 * ext, dtype=torch.float64),
                          torch.full((4,), 0.5 * ext, dtype=torch.float64),
                          0.25 * tmax + 0.5 * tmax * pu[:, 0],
                          (pu[:, 1] - 0.5) * 2e-4, (pu[:, 2] - 0.5) * 2e-4,
                          0.5 * torch.ones(4, dtype=torch.float64)], 1)
    tr = synthesize(src.to(device), rcv.to(device), n_samples, dt, 30.0, diff, None, planes, vfield=vf)
    p = _image_params(n_bins, n_bins, n_samples, d, d, n_samples, dt, 3000.0, 60.0, FLAG_AA if aa else 0,
                      n_traces=src.shape[0], aa_scale=1.0, aa_max=16.0)
    return SyntheticSurvey("pstm3d-single-offset", p, src.to(device), rcv.to(device), tr, vf.on_grid(p, device),
                           class_sizes=sizes, info=dict(diffractors=diff, planes=planes))


def precision_3d(seed: int = 3, n_bins: int = 200, n_classes: int = 25, n_samples: int = 1001,
                      n_img: int = 400, device="cpu", vt: bool = False) -> SyntheticSurvey:
    """Precision form: n_classes offset classes 100:100:2500 m, one trace per
    25 m CMP bin per class (n_bins^2 each, random azimuth), 1001 samples at
    2 ms; the 2D line's generator (5 diffractors + 3 flat reflectors, 25 Hz, v = 2000,
    noise 0.05 of peak); image n_img^2 x 1001 at 12.5 m; aperture 1500 m, 60
    deg; compensated accumulation, no antialias.  vt: v_rms = 1500 + 600 t
    (the single-offset form's v(t), laterally constant) instead of 2000 m/s,
    for traces and migration alike."""
    gen = torch.Generator().manual_seed(seed)
    dt, d = 0.002, 25.0
    ext = d * (n_bins - 1)
    offsets = [100.0 * (i + 1) for i in range(n_classes)]
    src, rcv, sizes = _cmp_grid_geometry(gen, n_bins, n_bins, d, offsets)
    diff, flats = _events_line(gen, ext, ext, dt * (n_samples - 1))
    vf = VrmsField(1500.0, 600.0) if vt else VrmsField(2000.0)
    tr = synthesize(src.to(device), rcv.to(device), n_samples, dt, 25.0, diff, flats, vfield=vf,
                    noise=0.05, gen=gen)
    d_img = ext / (n_img - 1) if n_img > 1 else 12.5
    p = _image_params(n_img, n_img, n_samples, d_img, d_img, n_samples, dt, 1500.0, 60.0, FLAG_KAHAN,
                      n_traces=src.shape[0])
    return SyntheticSurvey("pstm3d-precision", p, src.to(device), rcv.to(device), tr, vf.on_grid(p, device),
                           class_sizes=sizes, info=dict(diffractors=diff, flats=flats))


def vti_largest_3d(seed: int = 5, n_bins: int = 200, n_classes: int = 60, class_min: int = 5000,
                class_max: int = 40000, n_samples: int = 1251, n_img: int = 512, d_img: float = 12.5,
                n_diff: int = 60, device="cpu") -> SyntheticSurvey:
    """Largest single-GPU form: VTI with a smooth random eta in [0, 0.15], 60
    offset classes of uneven size U{class_min..class_max} (random CMP bins of a
    n_bins^2 grid at 25 m, random azimuth), offsets 0..4000 m, 1251 samples;
    image n_img^2 x 1251 at d_img; aperture 4000 m, 60 deg, triangle
    antialias; traces migrate as FP16 with per-trace scales."""
    gen = torch.Generator().manual_seed(seed)
    dt = 0.002
    d = 25.0
    ext_img = d_img * (n_img - 1)
    ext = ext_img
    sizes = torch.randint(class_min, class_max + 1, (n_classes,), generator=gen).tolist()
    offsets = torch.linspace(0.0, 4000.0, n_classes, dtype=torch.float64).tolist()
    S, R = [], []
    for n_c, o in zip(sizes, offsets):
        mx = ext * torch.rand(n_c, generator=gen, dtype=torch.float64)
        my = ext * torch.rand(n_c, generator=gen, dtype=torch.float64)
        az = math.pi * torch.rand(n_c, generator=gen, dtype=torch.float64)
        hx, hy = 0.5 * o * torch.cos(az), 0.5 * o * torch.sin(az)
        S.append(torch.stack([mx - hx, my - hy], 1))
        R.append(torch.stack([mx + hx, my + hy], 1))
    src, rcv = torch.cat(S), torch.cat(R)
    vf = VrmsField(1500.0, 600.0)
    tmax = dt * (n_samples - 1)
    u = torch.rand(n_diff, 4, generator=gen, dtype=torch.float64)
    diff = torch.stack([ext * u[:, 0], ext * u[:, 1], 0.15 * tmax + 0.7 * tmax * u[:, 2],
                        torch.randn(n_diff, generator=gen, dtype=torch.float64)], 1)
    p = _image_params(n_img, n_img, n_samples, d_img, d_img, n_samples, dt, 4000.0, 60.0,
                      FLAG_AA | FLAG_VTI | FLAG_KAHAN, n_traces=src.shape[0], aa_scale=1.0, aa_max=16.0)
    eta = _smooth_random_field(gen, n_img, n_img, n_samples, 0.0, 0.15, device)
    # eta at each diffractor (nearest image sample)
    ii = torch.clamp((diff[:, 0] / d_img).round().long(), 0, n_img - 1)
    jj = torch.clamp((diff[:, 1] / d_img).round().long(), 0, n_img - 1)
    kk = torch.clamp((diff[:, 2] / dt).round().long(), 0, n_samples - 1)
    eta_d = eta.cpu()[ii, jj, kk].double()
    tr = synthesize(src.to(device), rcv.to(device), n_samples, dt, 30.0, diff, None, None, vfield=vf,
                    eta_diff=eta_d)
    vel = vf.on_grid(p, device)  # laterally constant: broadcast to eta's layout once
    vel = vel[None, None, :].expand(n_img, n_img, n_samples).contiguous()
    return SyntheticSurvey("pstm3d-vti-largest", p, src.to(device), rcv.to(device), tr, vel, eta=eta,
                           class_sizes=sizes, info=dict(diffractors=diff))


def reflectivity(gen: torch.Generator, n_il: int, n_xl: int, n_tau: int, dt: float, p_spike: float = 0.01,
                 f_peak: float = 25.0, device="cpu") -> torch.Tensor:
    """Sparse Bernoulli(p_spike) spikes with N(0, 1) amplitudes on an
    [il][xl][tau] grid, convolved along time with a zero-phase Ricker."""
    spikes = torch.rand((n_il, n_xl, n_tau), generator=gen) < p_spike
    amp = (torch.randn((n_il, n_xl, n_tau), generator=gen) * spikes).to(device)
    half = int(math.ceil(1.5 / (f_peak * dt)))
    w = ricker(dt * torch.arange(-half, half + 1, dtype=torch.float64), f_peak).tolist()
    r = torch.zeros_like(amp)
    for k, wk in enumerate(w):  # r[t] = sum_k w[k] amp[t - half + k] (w is symmetric)
        s = half - k
        if s >= 0:
            r[..., s:] += wk * amp[..., :n_tau - s]
        else:
            r[..., :n_tau + s] += wk * amp[..., -s:]
    return r


def lsm_3d(seed: int = 4, n_img: int = 256, n_tau: int = 751, d_img: float = 20.0, n_classes: int = 16,
                aperture: float = 2000.0, snr: float = 10.0, p_spike: float = 0.01,
                device="cuda") -> SyntheticSurvey:
    """Differentiation form: reflectivity n_img^2 x n_tau at d_img m and 2 ms
    (Bernoulli(p_spike) N(0,1) spikes convolved with a 25 Hz Ricker);
    n_classes offset classes 0:100:(100 (n_classes - 1)) m, one trace per image
    bin per class (random azimuth), n_tau samples at 2 ms; v_rms rising
    linearly from 1500 to 3500 m/s over the record; data = demigration
    (adjoint kernel) of the reflectivity plus Gaussian noise at an RMS
    signal-to-noise ratio of snr; aperture 2000 m, 60 deg.  Needs the GPU
    (the data are modelled by the adjoint kernel)."""
    from .processors.migration.hip_ops import PstmPlan
    gen = torch.Generator().manual_seed(seed)
    dt = 0.002
    offsets = [100.0 * i for i in range(n_classes)]
    src, rcv, sizes = _cmp_grid_geometry(gen, n_img, n_img, d_img, offsets)
    p = _image_params(n_img, n_img, n_tau, d_img, d_img, n_tau, dt, aperture, 60.0, 0, n_traces=src.shape[0])
    tmax = dt * (n_tau - 1)
    vf = VrmsField(1500.0, 2000.0 / tmax)
    vel = vf.on_grid(p, device)
    m_true = reflectivity(gen, n_img, n_img, n_tau, dt, p_spike, 25.0, device)
    plan = PstmPlan(p, src.to(device), rcv.to(device), device=torch.device(device))
    data = plan.adjoint(m_true, vel, flags=0)
    rms = float(data.double().pow(2).mean().sqrt())
    ng = torch.Generator(device=device).manual_seed(seed + 1)
    data += (rms / snr) * torch.randn(data.shape, generator=ng, device=device, dtype=torch.float32)
    return SyntheticSurvey("lsm3d", p, src.to(device), rcv.to(device), data, vel, class_sizes=sizes,
                           info=dict(reflectivity=m_true, plan=plan, snr=snr))


def tiny_survey(seed: int = 0, n_bins: int = 12, n_classes: int = 3, n_samples: int = 301,
                n_img: int = 10, d_img: float = 25.0, aperture: float = 400.0, flags: int = 0,
                lateral: bool = False, device="cpu", utm: bool = False) -> SyntheticSurvey:
    """Small 3D survey for parity tests (oracle-sized)."""
    gen = torch.Generator().manual_seed(seed)
    dt, d = 0.004, 25.0
    x0, y0 = (500000.0, 6700000.0) if utm else (0.0, 0.0)
    offsets = [150.0 * (i + 1) for i in range(n_classes)]
    src, rcv, sizes = _cmp_grid_geometry(gen, n_bins, n_bins, d, offsets, x0=x0 - 25.0, y0=y0 - 25.0)
    ext = d * (n_bins - 1)
    vf = VrmsField(1800.0, 500.0, 0.05 if lateral else 0.0,
                   [(x0 + 0.3 * ext, y0 + 0.5 * ext, 0.3 * ext, 1.0)] if lateral else None)
    tmax = dt * (n_samples - 1)
    u = torch.rand(4, 3, generator=gen, dtype=torch.float64)
    diff = torch.stack([x0 + ext * u[:, 0], y0 + ext * u[:, 1], 0.3 * tmax + 0.5 * tmax * u[:, 2],
                        torch.randn(4, generator=gen, dtype=torch.float64)], 1)
    flats = torch.tensor([[0.55 * tmax, 0.5]], dtype=torch.float64)
    tr = synthesize(src, rcv, n_samples, dt, 20.0, diff, flats, vfield=vf, noise=0.02, gen=gen)
    p = _image_params(n_img, n_img, n_samples, d_img, d_img, n_samples, dt, aperture, 60.0, flags,
                      origin=(x0, y0), n_traces=src.shape[0], aa_scale=1.0, aa_max=8.0)
    vel = vf.on_grid(p, "cpu")
    if not lateral:
        vel = vel.contiguous()
    eta = None
    if flags & FLAG_VTI:
        eta = _smooth_random_field(gen, n_img, n_img, n_samples, 0.0, 0.15, "cpu")
        if vel.dim() == 1:
            vel = vel[None, None, :].expand(n_img, n_img, n_samples).contiguous()
    return SyntheticSurvey("tiny", p, src.to(device), rcv.to(device), tr.to(device), vel.to(device),
                           eta=None if eta is None else eta.to(device), class_sizes=sizes)
