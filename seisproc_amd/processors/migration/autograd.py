"""torch.autograd wrappers of the HIP migration operators.

  migrate(traces, vel, plan)    image = L(v) d         (forward kernel)
      d image / d traces   -> adjoint kernel (demigration)   L^T
      d image / d v_rms    -> velocity-gradient kernel (masks held fixed,
                              traveltimes recomputed, nothing stored)
  demigrate(image, vel, plan)   data = L^T(v) m        (adjoint kernel)
      d data / d image     -> forward kernel             L

Both operators are linear in their first argument, so the adjoint pair is
exact up to FP32 rounding (tests/test_gpu_adjoint.py checks <L d, m> =
<d, L^T m> to 1e-5).  The reference has no adjoint or gradient; these serve
least-squares migration (lsm.py) and image-power velocity updates
(velocity_update.py).
"""
from __future__ import annotations

from typing import Optional

import torch

from ...abi_types import FLAG_AA, FLAG_KAHAN, FLAG_VTI, FLAG_ZERO_OUTSIDE


def _vel3d(plan, vel: torch.Tensor) -> torch.Tensor:
    if vel.dim() == 3:
        return vel.contiguous()
    return vel[None, None, :].expand(plan.image_shape).contiguous()


class KirchhoffMigrationFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, traces: torch.Tensor, vel: torch.Tensor, plan, flags: int, eta: Optional[torch.Tensor]):
        image, _ = plan.forward(traces.contiguous(), vel.contiguous(), eta=eta, with_fold=False, flags=flags)
        ctx.plan, ctx.flags, ctx.eta = plan, flags, eta
        ctx.vel_was_1d = vel.dim() == 1
        ctx.save_for_backward(traces, vel)
        return image

    @staticmethod
    def backward(ctx, grad_image: torch.Tensor):
        traces, vel = ctx.saved_tensors
        plan, flags = ctx.plan, ctx.flags
        g = grad_image.contiguous()
        g_tr = g_vel = None
        if ctx.needs_input_grad[0]:
            g_tr = plan.adjoint(g, vel.contiguous(), eta=ctx.eta, flags=flags & ~FLAG_KAHAN)
        if ctx.needs_input_grad[1]:
            if flags & (FLAG_AA | FLAG_VTI | FLAG_ZERO_OUTSIDE):
                raise NotImplementedError("the velocity gradient is available for isotropic, non-antialiased, "
                                          "clamped migration")
            g3 = plan.vel_grad(traces.contiguous(), _vel3d(plan, vel), g, flags=flags & ~FLAG_KAHAN)
            g_vel = g3.sum(dim=(0, 1)) if ctx.vel_was_1d else g3
        return g_tr, g_vel, None, None, None


class KirchhoffDemigrationFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image: torch.Tensor, vel: torch.Tensor, plan, flags: int, eta: Optional[torch.Tensor]):
        ctx.plan, ctx.flags, ctx.eta = plan, flags, eta
        ctx.save_for_backward(vel)
        return plan.adjoint(image.contiguous(), vel.contiguous(), eta=eta, flags=flags & ~FLAG_KAHAN)

    @staticmethod
    def backward(ctx, grad_traces: torch.Tensor):
        (vel,) = ctx.saved_tensors
        g_img = None
        if ctx.needs_input_grad[0]:
            g_img, _ = ctx.plan.forward(grad_traces.contiguous(), vel.contiguous(), eta=ctx.eta, with_fold=False,
                                        flags=ctx.flags)
        if ctx.needs_input_grad[1]:
            raise NotImplementedError("velocity gradient of demigration is not provided")
        return g_img, None, None, None, None


def migrate(traces: torch.Tensor, vel: torch.Tensor, plan, flags: Optional[int] = None,
            eta: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Differentiable Kirchhoff migration: image [il][xl][tau] of FP32 traces."""
    return KirchhoffMigrationFn.apply(traces, vel, plan, plan.params.flags if flags is None else flags, eta)


def demigrate(image: torch.Tensor, vel: torch.Tensor, plan, flags: Optional[int] = None,
              eta: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Differentiable Kirchhoff demigration (modelling): traces of an image."""
    return KirchhoffDemigrationFn.apply(image, vel, plan, plan.params.flags if flags is None else flags, eta)
