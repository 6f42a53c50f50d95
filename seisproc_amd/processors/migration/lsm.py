"""Least-squares Kirchhoff migration (CGLS on the normal equations).

Modelling operator D = L^T (demigration, adjoint kernel: image -> data),
migration L (forward kernel: data -> image).  Solves
    min_m || D m - d ||^2 (+ damp ||m||^2)
with conjugate gradients on L D m = L d; every iteration costs one adjoint
and one forward launch, all on the GPU with FP64 reductions of the inner
products.  The two operators are the torch.autograd.Function objects of
autograd.py (applied without recording a graph), so the CG iterations use
exactly the pair whose gradients autograd returns.  The reference has no
least-squares migration.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import torch

from ...abi_types import FLAG_KAHAN
from .autograd import demigrate, migrate


@dataclass
class LSMResult:
    image: torch.Tensor
    data_residual_norms: List[float] = field(default_factory=list)
    gradient_norms: List[float] = field(default_factory=list)
    iterations: int = 0
    seconds: float = 0.0
    seconds_per_iteration: float = 0.0


def _dot(a: torch.Tensor, b: torch.Tensor) -> float:
    return float(torch.dot(a.reshape(-1).double(), b.reshape(-1).double()))


def cgls(plan, data: torch.Tensor, vel: torch.Tensor, n_iter: int = 30, *, eta: Optional[torch.Tensor] = None,
         flags: Optional[int] = None, m0: Optional[torch.Tensor] = None, damp: float = 0.0,
         callback: Optional[Callable[[int, float, float], None]] = None) -> LSMResult:
    """CGLS least-squares migration of `data` (n_traces, n_samples) FP32."""
    fl = (plan.params.flags if flags is None else flags) & ~FLAG_KAHAN

    def L(d):  # data -> image
        with torch.no_grad():
            return migrate(d, vel, plan, flags=fl, eta=eta)

    def Lt(m):  # image -> data
        with torch.no_grad():
            return demigrate(m, vel, plan, flags=fl, eta=eta)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m = torch.zeros(plan.image_shape, dtype=torch.float32, device=data.device) if m0 is None else m0.clone()
    r = data.clone() if m0 is None else data - Lt(m)
    s = L(r) - damp * m
    p = s.clone()
    gamma = _dot(s, s)
    res = LSMResult(image=m, data_residual_norms=[_dot(r, r) ** 0.5], gradient_norms=[gamma ** 0.5])
    for it in range(n_iter):
        q = Lt(p)
        denom = _dot(q, q) + damp * _dot(p, p)
        if denom <= 0.0 or gamma <= 0.0:
            break
        alpha = gamma / denom
        m.add_(p, alpha=alpha)
        r.add_(q, alpha=-alpha)
        s = L(r) - damp * m
        gamma_new = _dot(s, s)
        beta = gamma_new / gamma
        gamma = gamma_new
        p.mul_(beta).add_(s)
        res.data_residual_norms.append(_dot(r, r) ** 0.5)
        res.gradient_norms.append(gamma ** 0.5)
        res.iterations = it + 1
        if callback:
            callback(it + 1, res.data_residual_norms[-1], res.gradient_norms[-1])
    torch.cuda.synchronize()
    res.seconds = time.perf_counter() - t0
    res.seconds_per_iteration = res.seconds / max(res.iterations, 1)
    res.image = m
    return res
