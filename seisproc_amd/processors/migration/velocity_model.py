"""Engine-side velocity model (reference processors/migration/velocity_model.py:30-345).

``VelocityModel(velocity_type, v0, gradient, velocity_array, z_axis)`` with
``interval_velocity(z)`` and ``rms_velocity_at_time(t_ms)`` for 'constant',
'gradient' (closed form v_rms^2 = v0^2 + v0 k z + (k z)^2 / 3 at
z = v0 t + k v0 t^2 / 2, t one-way, :124-210) and 'file' (trapezoid v_rms(t)
table, :212-261).  ``vrms_table(t_ms)`` gives the FP32 v_rms the HIP kernels
take (one value per output time).
"""
from __future__ import annotations

import logging
from dataclasses import dataclass
from typing import Optional, Union

import numpy as np
import torch

from ...models.velocity_model import depth_interval_to_vrms_table

logger = logging.getLogger(__name__)


@dataclass
class VelocityModel:
    velocity_type: str
    v0: float
    gradient: float = 0.0
    velocity_array: Optional[np.ndarray] = None
    z_axis: Optional[np.ndarray] = None

    def __post_init__(self):
        if self.velocity_type not in ("constant", "gradient", "file"):
            raise ValueError(f"Unknown velocity type: {self.velocity_type}")
        if self.velocity_type == "file" and self.velocity_array is None:
            raise ValueError("velocity_array required for 'file' type")

    def interval_velocity(self, z):
        if self.velocity_type == "constant":
            if isinstance(z, torch.Tensor):
                return torch.full_like(z, self.v0)
            if isinstance(z, np.ndarray):
                return np.full_like(z, self.v0)
            return self.v0
        if self.velocity_type == "gradient":
            return self.v0 + self.gradient * z
        if isinstance(z, torch.Tensor):
            return torch.from_numpy(np.interp(z.cpu().numpy(), self.z_axis, self.velocity_array)).to(z.device)
        return np.interp(z, self.z_axis, self.velocity_array)

    def _rms_np(self, t_ms: np.ndarray) -> np.ndarray:
        t_ms = np.asarray(t_ms, dtype=np.float64)
        if self.velocity_type == "constant":
            return np.full(t_ms.shape, self.v0, dtype=np.float64)
        if self.velocity_type == "gradient":
            v0, k = self.v0, self.gradient
            if abs(k) < 1e-10:
                return np.full(t_ms.shape, v0, dtype=np.float64)
            t = t_ms / 2000.0
            z = np.maximum(v0 * t + 0.5 * k * v0 * t * t, 0.0)
            vr = np.sqrt(np.maximum(v0 * v0 + v0 * k * z + (k * z) ** 2 / 3.0, 0.5 * v0 * v0))
            return np.where(t_ms < 1.0, v0, vr)
        tt, vr = depth_interval_to_vrms_table(np.asarray(self.z_axis, dtype=np.float64),
                                              np.asarray(self.velocity_array, dtype=np.float64))
        return np.interp(t_ms, tt * 1000.0, vr)

    def rms_velocity_at_time(self, t_ms: Union[float, np.ndarray, torch.Tensor]):
        """v_rms at two-way time t_ms (milliseconds)."""
        if isinstance(t_ms, torch.Tensor):
            return torch.as_tensor(self._rms_np(t_ms.detach().cpu().numpy()), dtype=torch.float32,
                                   device=t_ms.device)
        out = self._rms_np(t_ms)
        if np.ndim(t_ms) == 0:
            return float(out)
        return out.astype(np.float32)

    def vrms_table(self, t_ms: np.ndarray, device="cuda") -> torch.Tensor:
        return torch.as_tensor(self._rms_np(t_ms), dtype=torch.float32, device=device)

    def get_summary(self) -> str:
        """The model's kind and its defining numbers."""
        if self.velocity_type == "constant":
            return f"velocity {self.v0:g} m/s everywhere"
        if self.velocity_type == "gradient":
            return f"velocity {self.v0:g} m/s at the surface, increasing {self.gradient:g} m/s per m"
        return (f"velocity table of {len(self.velocity_array)} depths, "
                f"{float(self.velocity_array[0]):g} m/s at the top")


def create_velocity_model(velocity_type: str, v0: float, gradient: float = 0.0,
                          velocity_file: Optional[str] = None) -> VelocityModel:
    if velocity_type == "constant":
        return VelocityModel("constant", v0)
    if velocity_type == "gradient":
        return VelocityModel("gradient", v0, gradient=gradient)
    if velocity_type == "file":
        if velocity_file is None:
            raise ValueError("velocity_file required for 'file' type")
        data = np.loadtxt(velocity_file)
        return VelocityModel("file", float(data[0, 1]), velocity_array=data[:, 1].astype(np.float32),
                             z_axis=data[:, 0].astype(np.float32))
    raise ValueError(f"Unknown velocity type: {velocity_type}")


def create_velocity_model_from_config(config: dict) -> VelocityModel:
    return create_velocity_model(config.get("velocity_type", "constant"), config.get("velocity_v0", 2500.0),
                                 config.get("velocity_gradient", 0.0), config.get("velocity_file"))
