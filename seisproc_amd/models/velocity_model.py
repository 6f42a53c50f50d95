"""Velocity models (reference models/velocity_model.py:22-513 and
processors/migration/velocity_model.py:30-345).

``VelocityModel`` keeps the reference's fields (data, z_axis, x_axis,
y_axis, is_time, gradient, v0, metadata) and type inference.  For time
migration (is_time=True) the values are RMS velocities v_rms(t) (the
reference's DSR kernel takes velocity_mps as the RMS velocity); a depth
model (is_time=False) holds interval velocities v(z) and is converted to
v_rms(t) by the same trapezoid integration as the reference
(processors/migration/velocity_model.py:234-261).

``vrms_on_grid`` evaluates v_rms at every output sample of an image grid, as a
torch tensor in the kernels' layout: (n_tau,) when the model has no lateral
variation, (n_il, n_xl, n_tau) otherwise.

``BSplineVelocity`` is a cubic B-spline v_rms(x, y, tau) on a coarse lattice
(e.g. 32 x 32 x 64) whose grid evaluation is a chain of three small matrix
products, so torch autograd carries image-sample gradients back to the
coefficients.
"""
from __future__ import annotations

import logging
import math
from dataclasses import dataclass, field
from enum import Enum
from pathlib import Path
from typing import Any, Dict, Optional, Tuple, Union

import numpy as np
import torch

logger = logging.getLogger(__name__)


class VelocityType(Enum):
    CONSTANT = "constant"
    V_OF_Z = "v_of_z"
    V_OF_XZ = "v_of_xz"
    V_OF_XYZ = "v_of_xyz"


@dataclass
class VelocityModel:
    data: Union[float, np.ndarray]
    z_axis: Optional[np.ndarray] = None
    x_axis: Optional[np.ndarray] = None
    y_axis: Optional[np.ndarray] = None
    is_time: bool = True
    gradient: Optional[float] = None
    v0: Optional[float] = None
    metadata: Dict[str, Any] = field(default_factory=dict)

    def __post_init__(self):
        if isinstance(self.data, (int, float, np.floating)):
            self.data = float(self.data)
            self._type = VelocityType.CONSTANT
        elif isinstance(self.data, np.ndarray):
            nd = self.data.ndim
            if nd == 1:
                self._type = VelocityType.V_OF_Z
                if self.z_axis is None:
                    raise ValueError("z_axis required for 1D velocity model")
                if len(self.z_axis) != len(self.data):
                    raise ValueError(f"z_axis length ({len(self.z_axis)}) must match data length ({len(self.data)})")
            elif nd == 2:
                self._type = VelocityType.V_OF_XZ
                if self.z_axis is None or self.x_axis is None:
                    raise ValueError("z_axis and x_axis required for 2D velocity model")
            elif nd == 3:
                self._type = VelocityType.V_OF_XYZ
                if self.z_axis is None or self.x_axis is None or self.y_axis is None:
                    raise ValueError("z_axis, x_axis, y_axis required for 3D velocity model")
            else:
                raise ValueError(f"Unsupported velocity data dimension: {nd}")
            if self.data.dtype != np.float32:
                self.data = self.data.astype(np.float32)
        else:
            raise ValueError(f"Unsupported data type: {type(self.data)}")
        for ax in ("z_axis", "x_axis", "y_axis"):
            a = getattr(self, ax)
            if a is not None:
                setattr(self, ax, np.asarray(a, dtype=np.float64))
        if self.v0 is None:
            if self._type is VelocityType.CONSTANT:
                self.v0 = self.data
            elif self._type is VelocityType.V_OF_Z:
                self.v0 = float(self.data[0]) if self.z_axis[0] == 0 else float(np.interp(0.0, self.z_axis, self.data))
        if self.gradient is None and self._type is VelocityType.V_OF_Z and len(self.z_axis) >= 2:
            self.gradient = float(np.polyfit(self.z_axis, self.data, 1)[0])
        self._validate_velocities()

    def _validate_velocities(self):
        vmin = 0.1
        if self._type is VelocityType.CONSTANT:
            if self.data < vmin:
                raise ValueError(f"Velocity must be >= {vmin} m/s, got {self.data}")
            if self.data < 300 or self.data > 10000:
                logger.warning("Velocity %s m/s outside typical range [300, 10000] m/s", self.data)
        else:
            if np.min(self.data) < vmin:
                logger.warning("Clipping %d velocity values below %s m/s", int(np.sum(self.data < vmin)), vmin)
                self.data = np.clip(self.data, vmin, None)

    # ------------------------------------------------------------------
    @property
    def velocity_type(self) -> VelocityType:
        return self._type

    @property
    def is_constant(self) -> bool:
        return self._type is VelocityType.CONSTANT

    @property
    def has_gradient(self) -> bool:
        return self.gradient is not None and abs(self.gradient) > 1e-6

    @property
    def z_unit(self) -> str:
        return "s" if self.is_time else "m"

    @property
    def gradient_unit(self) -> str:
        return "m/s/s" if self.is_time else "1/s"

    def get_velocity_at(self, z, x=None, y=None):
        """Velocity at (z[, x[, y]]) with linear interpolation (and linear
        extrapolation on the lateral axes, like RegularGridInterpolator with
        fill_value=None in the reference)."""
        if self._type is VelocityType.CONSTANT:
            return np.full_like(np.asarray(z, dtype=np.float32), self.data) if isinstance(z, np.ndarray) else self.data
        if self._type is VelocityType.V_OF_Z:
            return np.interp(z, self.z_axis, self.data).astype(np.float32)
        z = np.atleast_1d(np.asarray(z, dtype=np.float64))
        if x is None or (self._type is VelocityType.V_OF_XYZ and y is None):
            raise ValueError("lateral coordinates required for this velocity model")
        x = np.broadcast_to(np.atleast_1d(np.asarray(x, dtype=np.float64)), z.shape)
        pts = [(self.z_axis, z), (self.x_axis, x)]
        if self._type is VelocityType.V_OF_XYZ:
            y = np.broadcast_to(np.atleast_1d(np.asarray(y, dtype=np.float64)), z.shape)
            pts.append((self.y_axis, y))
        return _multilinear(self.data.astype(np.float64), pts).astype(np.float32)

    def get_effective_velocity(self, z):
        if self._type is VelocityType.CONSTANT:
            return np.full_like(np.asarray(z, dtype=np.float32), self.data) if isinstance(z, np.ndarray) else self.data
        if self._type is not VelocityType.V_OF_Z:
            raise NotImplementedError(f"Effective velocity not implemented for {self._type}")
        z = np.atleast_1d(z)
        out = np.zeros_like(z, dtype=np.float32)
        for i, zi in enumerate(z):
            if zi <= 0:
                out[i] = self.v0
            else:
                zz = np.linspace(0, zi, 100)
                s = np.trapezoid(1.0 / np.interp(zz, self.z_axis, self.data), zz)
                out[i] = zi / s if s > 0 else self.v0
        return out[0] if len(out) == 1 else out

    # ------------------------------------------------------------------
    def rms_velocity_at_time(self, t_s: np.ndarray) -> np.ndarray:
        """v_rms at two-way times t_s (s) for a 1D or constant model."""
        t_s = np.asarray(t_s, dtype=np.float64)
        if self._type is VelocityType.CONSTANT:
            return np.full(t_s.shape, self.data, dtype=np.float64)
        if self._type is not VelocityType.V_OF_Z:
            raise ValueError("rms_velocity_at_time needs a 1D model; use vrms_on_grid")
        if self.is_time:
            return np.interp(t_s, self.z_axis, self.data.astype(np.float64))
        tt, vr = depth_interval_to_vrms_table(self.z_axis, self.data.astype(np.float64))
        return np.interp(t_s, tt, vr)

    def vrms_on_grid(self, p, device="cuda", honor_lateral: bool = True, inline_start: int = 1,
                     xline_start: int = 1) -> torch.Tensor:
        """v_rms at every output sample of the image described by the
        PstmParams `p`.  Lateral axes are world coordinates (from the grid's
        origin and step vectors), or, for tables read by lateral number
        (``metadata['lateral_axes']``), the image columns' inline/xline
        numbers (inline_start + il, xline_start + xl) or CDP numbers along a
        line (inline_start + il)."""
        tau = p.tau0 + p.dtau * np.arange(p.n_tau)
        if self._type in (VelocityType.CONSTANT, VelocityType.V_OF_Z):
            return torch.as_tensor(self.rms_velocity_at_time(tau), dtype=torch.float32, device=device)
        if not self.is_time:
            raise NotImplementedError("laterally varying depth models must be converted to v_rms(t) first")
        il = np.arange(p.n_il, dtype=np.float64)
        xl = np.arange(p.n_xl, dtype=np.float64)
        lateral = self.metadata.get("lateral_axes")
        if lateral in ("inline_xline", "cdp"):
            X = (inline_start + il[:, None]) + 0.0 * xl[None, :]
            Y = (xline_start + xl[None, :]) + 0.0 * il[:, None]
        else:
            X = p.origin_x + il[:, None] * p.il_dx + xl[None, :] * p.xl_dx
            Y = p.origin_y + il[:, None] * p.il_dy + xl[None, :] * p.xl_dy
        out = np.empty((p.n_il, p.n_xl, p.n_tau), dtype=np.float32)
        for i in range(p.n_il):
            zz = np.broadcast_to(tau[None, :], (p.n_xl, p.n_tau))
            xx = np.broadcast_to(X[i][:, None], (p.n_xl, p.n_tau))
            yy = np.broadcast_to(Y[i][:, None], (p.n_xl, p.n_tau))
            out[i] = self.get_velocity_at(zz.ravel(), xx.ravel(), yy.ravel()).reshape(p.n_xl, p.n_tau)
        return torch.as_tensor(out, device=device)

    # ------------------------------------------------------------------
    def to_dict(self) -> Dict[str, Any]:
        d = {"velocity_type": self._type.value, "is_time": self.is_time, "gradient": self.gradient,
             "v0": self.v0, "metadata": dict(self.metadata)}
        if self._type is VelocityType.CONSTANT:
            d["data"] = self.data
        else:
            d["data"] = self.data.tolist()
            d["z_axis"] = self.z_axis.tolist()
            if self.x_axis is not None:
                d["x_axis"] = self.x_axis.tolist()
            if self.y_axis is not None:
                d["y_axis"] = self.y_axis.tolist()
        return d

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "VelocityModel":
        vt = VelocityType(d["velocity_type"])
        data = d["data"] if vt is VelocityType.CONSTANT else np.array(d["data"], dtype=np.float32)
        ax = lambda k: np.array(d[k]) if d.get(k) else None
        return cls(data=data, z_axis=ax("z_axis"), x_axis=ax("x_axis"), y_axis=ax("y_axis"),
                   is_time=d.get("is_time", True), gradient=d.get("gradient"), v0=d.get("v0"),
                   metadata=d.get("metadata", {}))

    @classmethod
    def load(cls, path: Union[str, Path]) -> "VelocityModel":
        """Load a model from JSON (to_dict) or a two-column text file
        (z or t, v).  The reference's job config calls a `load` that its
        VelocityModel lacks (models/migration_job.py:235)."""
        path = Path(path)
        if path.suffix.lower() == ".json":
            import json
            return cls.from_dict(json.loads(path.read_text()))
        arr = np.loadtxt(path)
        return cls(data=arr[:, 1].astype(np.float32), z_axis=arr[:, 0])

    def copy(self) -> "VelocityModel":
        return VelocityModel.from_dict(self.to_dict())

    def __repr__(self) -> str:
        if self._type is VelocityType.CONSTANT:
            return f"<VelocityModel constant {self.data:g} m/s>"
        if self._type is VelocityType.V_OF_Z:
            return (f"<VelocityModel v_of_z, {len(self.data)} samples in {self.z_unit}, {self.v0:g} m/s at the top, "
                    f"mean gradient {self.gradient or 0.0:.3g} {self.gradient_unit}>")
        return f"<VelocityModel {self._type.value} {'x'.join(map(str, self.data.shape))}>"


def _multilinear(data: np.ndarray, axes_pts) -> np.ndarray:
    """Linear interpolation (linear extrapolation outside) on a tensor grid."""
    idx, wts = [], []
    for ax, x in axes_pts:
        ax = np.asarray(ax, dtype=np.float64)
        if len(ax) == 1:
            idx.append((np.zeros_like(x, dtype=np.int64), np.zeros_like(x, dtype=np.int64)))
            wts.append(np.zeros_like(x))
            continue
        i = np.clip(np.searchsorted(ax, x, side="right") - 1, 0, len(ax) - 2)
        f = (x - ax[i]) / (ax[i + 1] - ax[i])
        idx.append((i, i + 1))
        wts.append(f)
    out = np.zeros_like(axes_pts[0][1], dtype=np.float64)
    nd = len(axes_pts)
    for corner in range(1 << nd):
        w = np.ones_like(out)
        sel = []
        for d in range(nd):
            b = (corner >> d) & 1
            w = w * (wts[d] if b else (1.0 - wts[d]))
            sel.append(idx[d][b])
        out += w * data[tuple(sel)]
    return out


def depth_interval_to_vrms_table(z: np.ndarray, v: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Two-way time and v_rms of a depth interval-velocity profile by trapezoid
    integration (processors/migration/velocity_model.py:234-261)."""
    dz = np.diff(z)
    vav = 0.5 * (v[:-1] + v[1:])
    dt_owt = dz / vav
    t_owt = np.concatenate([[0.0], np.cumsum(dt_owt)])
    v2t = np.concatenate([[0.0], np.cumsum(vav ** 2 * dt_owt)])
    with np.errstate(divide="ignore", invalid="ignore"):
        vr2 = v2t / np.maximum(t_owt, 1e-10)
    vr2[0] = v[0] ** 2
    return 2.0 * t_owt, np.sqrt(vr2)


# ---------------------------------------------------------------------------
# factories and Dix conversions (reference models/velocity_model.py:370-513)

def create_constant_velocity(velocity: float, is_time: bool = True) -> VelocityModel:
    return VelocityModel(data=float(velocity), is_time=is_time)


def create_linear_gradient_velocity(v0: float, gradient: float, z_max: float, dz: float = 0.004,
                                    is_time: bool = True) -> VelocityModel:
    z = np.arange(0, z_max + dz, dz, dtype=np.float64)
    return VelocityModel(data=(v0 + gradient * z).astype(np.float32), z_axis=z, is_time=is_time,
                         gradient=gradient, v0=v0, metadata={"model_type": "linear_gradient"})


def _checked_field(data, axes, names):
    data = np.asarray(data, dtype=np.float32)
    if data.ndim != len(axes):
        raise ValueError(f"Data must be {len(axes)}D, got shape {data.shape}")
    for k, (ax, nm) in enumerate(zip(axes, names)):
        if data.shape[k] != len(ax):
            raise ValueError(f"Data {nm}-dimension ({data.shape[k]}) must match {nm}_axis length ({len(ax)})")
    return data


def create_2d_velocity(data, z_axis, x_axis, is_time: bool = True,
                       metadata: Optional[Dict[str, Any]] = None) -> VelocityModel:
    """v(x, z) from data (n_z, n_x) on its axes (z two-way time in s when is_time)."""
    data = _checked_field(data, (z_axis, x_axis), ("z", "x"))
    return VelocityModel(data=data, z_axis=np.asarray(z_axis, dtype=np.float64),
                         x_axis=np.asarray(x_axis, dtype=np.float64), is_time=is_time,
                         metadata=metadata or {"model_type": "2d_field"})


def create_3d_velocity(data, z_axis, x_axis, y_axis, is_time: bool = True,
                       metadata: Optional[Dict[str, Any]] = None) -> VelocityModel:
    """v(x, y, z) from data (n_z, n_x, n_y) on its axes."""
    data = _checked_field(data, (z_axis, x_axis, y_axis), ("z", "x", "y"))
    return VelocityModel(data=data, z_axis=np.asarray(z_axis, dtype=np.float64),
                         x_axis=np.asarray(x_axis, dtype=np.float64), y_axis=np.asarray(y_axis, dtype=np.float64),
                         is_time=is_time, metadata=metadata or {"model_type": "3d_cube"})


def create_2d_gradient_velocity(v0: float, z_gradient: float, x_gradient: float, z_max: float, x_max: float,
                                dz: float = 0.004, dx: float = 25.0, is_time: bool = True) -> VelocityModel:
    """v(x, z) = v0 + z_gradient z + x_gradient x sampled every dz, dx from the origin."""
    z = np.arange(0, z_max + dz, dz, dtype=np.float64)
    x = np.arange(0, x_max + dx, dx, dtype=np.float64)
    return VelocityModel(data=(v0 + z_gradient * z[:, None] + x_gradient * x[None, :]).astype(np.float32),
                         z_axis=z, x_axis=x, is_time=is_time,
                         metadata={"model_type": "2d_linear_gradient", "v0": v0, "z_gradient": z_gradient,
                                   "x_gradient": x_gradient})


def create_layered_velocity(layer_depths, layer_velocities, z_max: float, dz: float = 0.004,
                            is_time: bool = True) -> VelocityModel:
    """Piecewise-constant v(z): layer i holds from layer_depths[i] (its top)
    to the next top, the last layer to z_max; samples above the first top
    are 0 (the first top is meant to be 0)."""
    tops = np.asarray(layer_depths, dtype=np.float64)
    vel = np.asarray(layer_velocities, dtype=np.float64)
    if len(tops) != len(vel):
        raise ValueError("layer_depths and layer_velocities must have same length")
    z = np.arange(0, z_max + dz, dz, dtype=np.float64)
    k = np.searchsorted(tops, z, side="right") - 1  # the layer whose top is at or above z
    data = np.where(k >= 0, vel[np.clip(k, 0, None)], 0.0).astype(np.float32)
    return VelocityModel(data=data, z_axis=z, is_time=is_time,
                         metadata={"model_type": "layered", "n_layers": len(vel)})


def create_from_rms_velocity(t_axis, v_rms, is_time: bool = True) -> VelocityModel:
    return VelocityModel(data=np.asarray(v_rms, dtype=np.float32), z_axis=np.asarray(t_axis, dtype=np.float64),
                         is_time=is_time, metadata={"velocity_type": "rms"})


def rms_to_interval_velocity(t_axis, v_rms):
    """Dix: v_int(n) = sqrt((v_rms(n)^2 t(n) - v_rms(n-1)^2 t(n-1)) / (t(n) - t(n-1)))."""
    t = np.asarray(t_axis, dtype=np.float64)
    vr = np.asarray(v_rms, dtype=np.float64)
    vi = np.zeros(len(t), dtype=np.float32)
    vi[0] = vr[0]
    for i in range(1, len(t)):
        d = t[i] - t[i - 1]
        num = vr[i] ** 2 * t[i] - vr[i - 1] ** 2 * t[i - 1]
        vi[i] = math.sqrt(num / d) if d > 0 and num > 0 else vr[i]
    return t_axis, vi


def interval_to_rms_velocity(t_axis, v_int):
    """v_rms(n) = sqrt(sum v_int(i)^2 dt(i) / t(n))."""
    t = np.asarray(t_axis, dtype=np.float64)
    vi = np.asarray(v_int, dtype=np.float64)
    vr = np.zeros(len(t), dtype=np.float32)
    vr[0] = vi[0]
    acc = vi[0] ** 2 * t[0] if t[0] > 0 else 0.0
    for i in range(1, len(t)):
        acc += vi[i] ** 2 * (t[i] - t[i - 1])
        vr[i] = math.sqrt(acc / t[i]) if t[i] > 0 else vi[i]
    return t_axis, vr


# ---------------------------------------------------------------------------
# cubic B-spline velocity lattice

def _bspline_matrix(n_out: int, n_coef: int, dtype=torch.float64) -> torch.Tensor:
    """(n_out, n_coef) uniform cubic B-spline basis sampling the coefficient
    lattice at n_out evenly spaced points spanning it."""
    if n_coef < 4:
        raise ValueError("a cubic B-spline needs >= 4 coefficients per axis")
    B = torch.zeros(n_out, n_coef, dtype=dtype)
    span = n_coef - 3
    for i in range(n_out):
        u = (i / max(n_out - 1, 1)) * span
        k = min(int(math.floor(u)), span - 1)
        t = u - k
        w = [(1 - t) ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6, (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6,
             t ** 3 / 6]
        for j in range(4):
            B[i, k + j] = w[j]
    return B


class BSplineVelocity:
    """v_rms on an image grid from a (cx, cy, ct) cubic B-spline lattice of
    coefficients (m/s).  ``evaluate(coef)`` is differentiable in coef."""

    def __init__(self, n_il: int, n_xl: int, n_tau: int, shape=(32, 32, 64), device="cuda"):
        self.shape = tuple(shape)
        self.Bx = _bspline_matrix(n_il, shape[0]).float().to(device)
        self.By = _bspline_matrix(n_xl, shape[1]).float().to(device)
        self.Bt = _bspline_matrix(n_tau, shape[2]).float().to(device)

    def evaluate(self, coef: torch.Tensor) -> torch.Tensor:
        """coef (cx, cy, ct) -> v (n_il, n_xl, n_tau), contiguous."""
        a = torch.matmul(coef, self.Bt.t())                          # (cx, cy, nt)
        a = torch.matmul(self.By, a)                                  # (cx, n_xl, nt) via broadcast
        a = torch.matmul(self.Bx, a.reshape(self.shape[0], -1))       # (n_il, n_xl*nt)
        return a.reshape(self.Bx.shape[0], self.By.shape[0], self.Bt.shape[0]).contiguous()

    def fit(self, v: torch.Tensor) -> torch.Tensor:
        """Least-squares coefficients reproducing a grid velocity."""
        px, py, pt = (torch.linalg.pinv(B.double().cpu()).float().to(B.device) for B in (self.Bx, self.By, self.Bt))
        a = torch.matmul(v, pt.t())
        a = torch.matmul(py, a)
        a = torch.matmul(px, a.reshape(v.shape[0], -1))
        return a.reshape(self.shape).contiguous()
