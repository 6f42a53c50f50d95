"""Image-power velocity updates.

J(c) = sum over image samples of image(v(c))^2, where v = B c is the RMS
velocity evaluated from a cubic B-spline lattice of coefficients c (e.g.
32 x 32 x 64) and image = L(v) d is the Kirchhoff migration of the data.
dJ/dc = B^T [ 2 image * d image / d v ] : the bracket comes from the
velocity-gradient kernel (traveltimes recomputed, masks held fixed), the
B^T contraction from torch autograd through BSplineVelocity.evaluate.
"""
from __future__ import annotations

from typing import Tuple

import torch

from ...models.velocity_model import BSplineVelocity
from .autograd import migrate


def image_power(plan, traces: torch.Tensor, bspline: BSplineVelocity, coef: torch.Tensor,
                flags: int = 0) -> Tuple[torch.Tensor, torch.Tensor]:
    """(J, dJ/dcoef) for FP32 traces; isotropic, no antialias."""
    c = coef.detach().clone().requires_grad_(True)
    v = bspline.evaluate(c)
    img = migrate(traces, v, plan, flags=flags)
    J = (img.double() ** 2).sum()
    (g,) = torch.autograd.grad(J, c)
    return J.detach(), g


def ascent_step(plan, traces: torch.Tensor, bspline: BSplineVelocity, coef: torch.Tensor, step_mps: float = 5.0,
                flags: int = 0) -> Tuple[torch.Tensor, float]:
    """One normalised gradient-ascent step on image power: the coefficient
    with the largest |gradient| moves by step_mps m/s."""
    J, g = image_power(plan, traces, bspline, coef, flags)
    gmax = float(g.abs().max())
    if gmax == 0.0:
        return coef.clone(), float(J)
    return (coef + step_mps * g / gmax).detach(), float(J)
