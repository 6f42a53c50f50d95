"""VTI anisotropy model (reference models/anisotropy_model.py:23-596).

Thomsen epsilon / delta with eta = (epsilon - delta) / (1 + 2 delta)
(:133-155).  ``eta_on_grid`` evaluates eta at every output sample of an
image grid in the kernels' layout, for the per-leg anelliptic traveltime
(processors/migration/traveltime_vti.py:164-222 with v0 = v_rms, delta = 0).
"""
from __future__ import annotations

import logging
from dataclasses import dataclass, field
from enum import Enum
from typing import Any, Dict, Optional, Tuple, Union

import numpy as np
import torch

from .velocity_model import _multilinear

logger = logging.getLogger(__name__)


class AnisotropyType(Enum):
    CONSTANT = "constant"
    V_OF_Z = "v_of_z"
    V_2D = "v_2d"
    V_3D = "v_3d"


@dataclass
class AnisotropyModel:
    epsilon: Union[float, np.ndarray] = 0.0
    delta: Union[float, np.ndarray] = 0.0
    eta: Optional[Union[float, np.ndarray]] = None
    z_axis: Optional[np.ndarray] = None
    x_axis: Optional[np.ndarray] = None
    y_axis: Optional[np.ndarray] = None
    anisotropy_type: AnisotropyType = AnisotropyType.CONSTANT
    metadata: Dict[str, Any] = field(default_factory=dict)

    def __post_init__(self):
        if isinstance(self.epsilon, np.ndarray):
            self.anisotropy_type = {1: AnisotropyType.V_OF_Z, 2: AnisotropyType.V_2D,
                                    3: AnisotropyType.V_3D}[self.epsilon.ndim]
        else:
            self.anisotropy_type = AnisotropyType.CONSTANT
        if np.asarray(self.epsilon).shape != np.asarray(self.delta).shape:
            raise ValueError(f"epsilon shape {np.asarray(self.epsilon).shape} must match delta shape "
                             f"{np.asarray(self.delta).shape}")
        if self.eta is None:
            self.eta = self.compute_eta(self.epsilon, self.delta)
        self._check_bounds()

    def _check_bounds(self):
        eps, dl = np.asarray(self.epsilon), np.asarray(self.delta)
        if np.any(eps < -0.3) or np.any(eps > 0.6):
            logger.warning("Epsilon values outside typical range [-0.2, 0.5]")
        if np.any(dl < -0.4) or np.any(dl > 0.4):
            logger.warning("Delta values outside typical range [-0.3, 0.3]")
        if np.any(1 + 2 * dl <= 0):
            raise ValueError("Unstable anisotropy: (1 + 2*delta) must be positive")
        # 1 + 2 eta = (1 + 2 epsilon) / (1 + 2 delta): the anelliptic time
        # (and the migration kernels' window bounds) need it positive
        if np.any(1 + 2 * np.asarray(self.eta, dtype=np.float64) <= 0):
            raise ValueError("Unstable anisotropy: (1 + 2*eta) must be positive (epsilon > -0.5)")

    @staticmethod
    def compute_eta(epsilon, delta):
        den = 1 + 2 * np.asarray(delta, dtype=np.float64)
        den = np.where(np.abs(den) < 1e-10, 1e-10, den)
        out = (np.asarray(epsilon, dtype=np.float64) - np.asarray(delta, dtype=np.float64)) / den
        return float(out) if np.ndim(out) == 0 else out

    @property
    def is_isotropic(self) -> bool:
        return bool(np.allclose(self.epsilon, 0, atol=1e-6) and np.allclose(self.delta, 0, atol=1e-6))

    @property
    def is_elliptic(self) -> bool:
        return bool(np.allclose(self.eta, 0, atol=1e-6))

    def get_parameters_at(self, z, x=None, y=None) -> Tuple:
        t = self.anisotropy_type
        if t is AnisotropyType.CONSTANT:
            return self.epsilon, self.delta, self.eta
        if t is AnisotropyType.V_OF_Z:
            if self.z_axis is None:
                raise ValueError("z_axis required for V_OF_Z model")
            return tuple(np.interp(z, self.z_axis, np.asarray(a, dtype=np.float64))
                         for a in (self.epsilon, self.delta, self.eta))
        z = np.atleast_1d(np.asarray(z, dtype=np.float64))
        pts = [(self.x_axis, np.broadcast_to(np.atleast_1d(x), z.shape))]
        if t is AnisotropyType.V_3D:
            pts.append((self.y_axis, np.broadcast_to(np.atleast_1d(y), z.shape)))
        pts.append((self.z_axis, z))
        return tuple(_multilinear(np.asarray(a, dtype=np.float64), pts)
                     for a in (self.epsilon, self.delta, self.eta))

    def eta_on_grid(self, p, device="cuda", full: bool = False) -> torch.Tensor:
        """eta at every output sample of the PstmParams grid: (n_tau,) for
        constant / 1D models (unless full=True), else (n_il, n_xl, n_tau).
        The z axis of a 1D/2D/3D model is two-way time in seconds."""
        tau = p.tau0 + p.dtau * np.arange(p.n_tau)
        t = self.anisotropy_type
        if t in (AnisotropyType.CONSTANT, AnisotropyType.V_OF_Z):
            e = np.full(p.n_tau, float(self.eta)) if t is AnisotropyType.CONSTANT else \
                np.interp(tau, self.z_axis, np.asarray(self.eta, dtype=np.float64))
            e = torch.as_tensor(e, dtype=torch.float32, device=device)
            return e[None, None, :].expand(p.n_il, p.n_xl, p.n_tau).contiguous() if full else e
        il = np.arange(p.n_il, dtype=np.float64)
        xl = np.arange(p.n_xl, dtype=np.float64)
        X = p.origin_x + il[:, None] * p.il_dx + xl[None, :] * p.xl_dx
        Y = p.origin_y + il[:, None] * p.il_dy + xl[None, :] * p.xl_dy
        out = np.empty((p.n_il, p.n_xl, p.n_tau), dtype=np.float32)
        for i in range(p.n_il):
            zz = np.broadcast_to(tau[None, :], (p.n_xl, p.n_tau)).ravel()
            xx = np.broadcast_to(X[i][:, None], (p.n_xl, p.n_tau)).ravel()
            yy = np.broadcast_to(Y[i][:, None], (p.n_xl, p.n_tau)).ravel()
            out[i] = self.get_parameters_at(zz, xx, yy)[2].reshape(p.n_xl, p.n_tau)
        return torch.as_tensor(out, device=device)

    def to_dict(self) -> Dict[str, Any]:
        conv = lambda a: a.tolist() if isinstance(a, np.ndarray) else a
        d = {"epsilon": conv(self.epsilon), "delta": conv(self.delta), "eta": conv(self.eta),
             "anisotropy_type": self.anisotropy_type.value, "metadata": dict(self.metadata)}
        for ax in ("z_axis", "x_axis", "y_axis"):
            if getattr(self, ax) is not None:
                d[ax] = conv(np.asarray(getattr(self, ax)))
        return d

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "AnisotropyModel":
        arr = lambda v: np.asarray(v, dtype=np.float64) if isinstance(v, list) else v
        return cls(epsilon=arr(d.get("epsilon", 0.0)), delta=arr(d.get("delta", 0.0)), eta=arr(d.get("eta")),
                   z_axis=arr(d.get("z_axis")), x_axis=arr(d.get("x_axis")), y_axis=arr(d.get("y_axis")),
                   metadata=d.get("metadata", {}))

    def get_summary(self) -> str:
        """The model's kind and its Thomsen parameters (value, or min..max)."""
        def span(a):
            a = np.asarray(a, dtype=np.float64)
            return f"{float(a.min()):g}" if a.min() == a.max() else f"{float(a.min()):g}..{float(a.max()):g}"
        return (f"{self.anisotropy_type.value} VTI model: epsilon {span(self.epsilon)}, delta {span(self.delta)}, "
                f"eta {span(self.eta)}")


def create_isotropic() -> AnisotropyModel:
    return AnisotropyModel(0.0, 0.0)


def create_constant_anisotropy(epsilon: float, delta: float) -> AnisotropyModel:
    return AnisotropyModel(epsilon=epsilon, delta=delta)


def create_eta_field(eta: np.ndarray, x_axis, y_axis, t_axis) -> AnisotropyModel:
    """Anelliptic-only 3D model (epsilon = eta, delta = 0) on (x, y, t) axes."""
    eta = np.asarray(eta, dtype=np.float64)
    return AnisotropyModel(epsilon=eta, delta=np.zeros_like(eta), eta=eta, x_axis=np.asarray(x_axis),
                           y_axis=np.asarray(y_axis), z_axis=np.asarray(t_axis))


# Typical shale (epsilon, delta) by intensity (reference models/anisotropy_model.py:395-421)
SHALE_PARAMETERS = {"weak": (0.1, 0.05), "moderate": (0.2, 0.1), "strong": (0.35, 0.15)}


def create_shale_anisotropy(intensity: str = "moderate") -> AnisotropyModel:
    if intensity not in SHALE_PARAMETERS:
        raise ValueError(f"intensity must be one of {list(SHALE_PARAMETERS)}")
    eps, dl = SHALE_PARAMETERS[intensity]
    return AnisotropyModel(epsilon=eps, delta=dl, metadata={"formation": "shale", "intensity": intensity})


def create_1d_anisotropy(z_axis, epsilon, delta) -> AnisotropyModel:
    """epsilon and delta tabulated along z_axis (two-way time for the kernels)."""
    return AnisotropyModel(epsilon=np.asarray(epsilon, dtype=np.float64), delta=np.asarray(delta, dtype=np.float64),
                           z_axis=np.asarray(z_axis, dtype=np.float64))


def create_gradient_anisotropy(epsilon_surface: float, epsilon_gradient: float, delta_surface: float,
                               delta_gradient: float, z_max: float, dz: float = 0.01) -> AnisotropyModel:
    """epsilon and delta growing linearly from their surface values, sampled
    every dz on [0, z_max]."""
    z = np.arange(0.0, z_max + dz, dz)
    m = create_1d_anisotropy(z, epsilon_surface + epsilon_gradient * z, delta_surface + delta_gradient * z)
    m.metadata.update(epsilon_surface=epsilon_surface, epsilon_gradient=epsilon_gradient,
                      delta_surface=delta_surface, delta_gradient=delta_gradient)
    return m


def _v_factor(v0, f):
    # v0 sqrt(f) with f floored at 0.01, as the reference guards unphysical input
    return v0 * np.sqrt(np.maximum(f, 0.01))


def compute_vti_phase_velocity(v0: float, theta, epsilon, delta):
    """Weak-anisotropy phase velocity at angle theta (radians) from vertical:
    v0 sqrt(1 + 2 delta sin^2 cos^2 + 2 epsilon sin^4)."""
    s2 = np.sin(theta) ** 2
    return _v_factor(v0, 1.0 + 2.0 * np.asarray(delta) * s2 * (1.0 - s2) + 2.0 * np.asarray(epsilon) * s2 * s2)


def compute_nmo_velocity(v0: float, delta):
    """Short-spread moveout velocity v0 sqrt(1 + 2 delta)."""
    return _v_factor(v0, 1.0 + 2.0 * np.asarray(delta))


def compute_horizontal_velocity(v0: float, epsilon):
    """Horizontal P velocity v0 sqrt(1 + 2 epsilon)."""
    return _v_factor(v0, 1.0 + 2.0 * np.asarray(epsilon))


def compute_effective_eta(epsilon, delta):
    """eta = (epsilon - delta) / (1 + 2 delta)."""
    return AnisotropyModel.compute_eta(epsilon, delta)
