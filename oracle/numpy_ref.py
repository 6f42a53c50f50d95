"""Brute-force numpy restatement of the oracle forward operator for tiny
inputs -- TEST INFRASTRUCTURE ONLY (cross-checks the loops, index handling and
mask search of pstm_oracle.c).  Follows the same reference lines as
pstm_oracle.c (processors/migration/kirchhoff_kernel.py:611-730); every
(column, trace, tau) triple is evaluated, with no pre-filter and no search.
"""
from __future__ import annotations

import numpy as np

EPS = 1e-6


def _integrate(d: np.ndarray, pad: int) -> np.ndarray:
    ext = np.concatenate([np.zeros(pad), d, np.zeros(pad)])
    c = np.cumsum(ext)
    return np.cumsum(c[::-1])[::-1]


def _lerp(a: np.ndarray, x: float) -> float:
    i = int(np.floor(x))
    f = x - i
    return a[i] + f * (a[i + 1] - a[i])


def _edge(x: float, xmax: float, tw: float) -> float:
    """Cosine edge taper (processors/migration/aperture.py:226-270)."""
    s = xmax * (1 - tw)
    if x > xmax:
        return 0.0
    return 0.5 * (1 + np.cos(np.pi * (x - s) / (xmax - s))) if x > s else 1.0


def _offset_taper(p: dict, off: float) -> float:
    """Offset window with near/far cosine tapers (aperture.py:184-224)."""
    lo, hi = p.get("min_offset_m", 0.0), p.get("max_offset_m", 0.0)
    if not hi > 0:
        return 1.0
    if not lo <= off <= hi:
        return 0.0
    tr = p["taper_width"] * (hi - lo)
    near = 0.5 * (1 - np.cos(np.pi * (off - lo) / tr)) if off < lo + tr else 1.0
    far = 0.5 * (1 + np.cos(np.pi * (off - (hi - tr)) / tr)) if off > hi - tr else 1.0
    return near * far


def forward(p: dict, traces, sx, sy, rx, ry, vel, eta=None):
    n_il, n_xl, nt, ns = p["n_il"], p["n_xl"], p["n_tau"], p["n_samples"]
    flags = p.get("flags", 0)
    aa, vti, taper, zo, tap_time = (flags & 1, flags & 2, flags & 8, flags & 16, flags & 32)
    th = np.radians(p["max_angle_deg"])
    A = p["aperture_m"]
    pad = int(np.ceil(p.get("aa_max", 16.0))) + 2
    img = np.zeros((n_il, n_xl, nt))
    fold = np.zeros((n_il, n_xl, nt))
    vel = np.asarray(vel, dtype=np.float64)
    aint = [_integrate(np.asarray(traces[j], np.float64), pad) for j in range(len(sx))] if aa else None
    off = np.hypot(np.asarray(rx) - np.asarray(sx), np.asarray(ry) - np.asarray(sy))
    lo, hi = p.get("min_offset_m", 0.0), p.get("max_offset_m", 0.0)
    kept = (off >= lo) & (off <= hi) if hi > 0 else np.ones(len(sx), bool)
    for il in range(n_il):
        for xl in range(n_xl):
            x = p["origin_x"] + il * p["il_dx"] + xl * p["xl_dx"]
            y = p["origin_y"] + il * p["il_dy"] + xl * p["xl_dy"]
            for j in range(len(sx)):
                if not kept[j]:
                    continue
                dxs, dys, dxr, dyr = x - sx[j], y - sy[j], x - rx[j], y - ry[j]
                hs, hr = np.hypot(dxs, dys), np.hypot(dxr, dyr)
                for k in range(nt):
                    tau = p["tau0"] + k * p["dtau"]
                    v = vel[k] if vel.ndim == 1 else vel[il, xl, k]
                    z = v * tau / 2
                    ap = min(A, z * np.tan(th)) if tap_time else A
                    if not (hs < ap and hr < ap):
                        continue
                    if not (np.arctan2(hs, z + EPS) < th and np.arctan2(hr, z + EPS) < th):
                        continue
                    rs, rr = np.hypot(hs, z), np.hypot(hr, z)
                    if vti:
                        e = eta[k] if np.ndim(eta) == 1 else eta[il, xl, k]
                        t = 0.0
                        for h2 in (hs * hs, hr * hr):
                            zs = max(abs(z), 1e-6)
                            t0 = zs / v
                            den = max(v * v * (t0 * t0 * v * v + (1 + 2 * e) * h2), 1e-10)
                            t += np.sqrt(max(t0 * t0 + h2 / (v * v) - 2 * e * h2 * h2 / den, 1e-12))
                    else:
                        t = (rs + rr) / v
                    wm = p.get("weight_mode", 0)
                    w = {0: (z / (rs + EPS)) * (z / (rr + EPS)) / (rs * rr + EPS),
                         1: 1.0 / (rs * rr + EPS), 2: 1.0,
                         3: (z / (rs + EPS)) * (z / (rr + EPS))}[wm]
                    if taper:
                        tw = p["taper_width"]
                        w *= (_edge(hs, ap, tw) * _edge(hr, ap, tw) * _edge(np.arctan2(hs, abs(z)), th, tw)
                              * _edge(np.arctan2(hr, abs(z)), th, tw) * _offset_taper(p, off[j]))
                    u = (t - p["t0"]) / p["dt"]
                    if zo and (u < 0 or u > ns - 1):
                        continue
                    u = min(max(u, 0.0), ns - 1.001)
                    if aa:
                        i_s = 1.0 / (v * rs) if rs > 0 else 0.0
                        i_r = 1.0 / (v * rr) if rr > 0 else 0.0
                        gi = (dxs * p["il_dx"] + dys * p["il_dy"]) * i_s + (dxr * p["il_dx"] + dyr * p["il_dy"]) * i_r
                        gx = (dxs * p["xl_dx"] + dys * p["xl_dy"]) * i_s + (dxr * p["xl_dx"] + dyr * p["xl_dy"]) * i_r
                        L = min(max(p["aa_scale"] * (abs(gi) + abs(gx)) / p["dt"], 1.0), p["aa_max"])
                        a_ = aint[j]
                        xx = u + pad
                        amp = (2 * _lerp(a_, xx) - _lerp(a_, xx - L) - _lerp(a_, xx + L)) / (L * L)
                    else:
                        i0 = int(u)
                        f = u - i0
                        amp = traces[j][i0] + f * (traces[j][min(i0 + 1, ns - 1)] - traces[j][i0])
                    img[il, xl, k] += w * amp
                    fold[il, xl, k] += 1
    return img, fold
