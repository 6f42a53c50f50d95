"""GPU parity of the adjoint (demigration) and velocity-gradient kernels.

Adjoint: dot-product test <L d, m> = <d, L^T m> to 1e-5 relative (FP32
kernels, FP64 inner products) and agreement with the oracle's exact
transpose.  Velocity gradient: agreement with the oracle's analytic gradient,
which tests/test_oracle_kat.py pins by central finite differences.
"""
import numpy as np
import pytest
import torch

import oracle
from seisproc_amd import synthetic as S
from seisproc_amd.abi_types import FLAG_AA, FLAG_TAPER, FLAG_VTI, WEIGHT_SPREADING

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _setup(flags=0, lateral=True, seed=20, **kw):
    from seisproc_amd.processors.migration.hip_ops import PstmPlan
    s = S.tiny_survey(seed=seed, n_bins=14, n_classes=3, n_samples=301, n_img=12, aperture=400.0,
                      flags=flags & (FLAG_AA | FLAG_VTI), lateral=lateral)
    vel = s.vel if s.vel.dim() == 3 else s.vel[None, None, :].expand(12, 12, -1).contiguous()
    p = s.params.replace(flags=flags, vel_mode=1, **kw)
    return s, PstmPlan(p, s.src_xy, s.rcv_xy), p, vel


@pytest.mark.parametrize("flags", [0, FLAG_AA, FLAG_AA | FLAG_VTI | FLAG_TAPER])
def test_adjoint_dot_product(flags):
    # with the taper, also the offset window's near and far tapers
    win = dict(min_offset_m=100.0, max_offset_m=500.0, taper_width=0.3) if flags & FLAG_TAPER else {}
    s, plan, p, vel = _setup(flags, **win)
    eta = s.eta.to(DEV) if s.eta is not None else None
    g = torch.Generator(device=DEV).manual_seed(3)
    d = torch.randn((p.n_traces, p.n_samples), generator=g, device=DEV)
    m = torch.randn(plan.image_shape, generator=g, device=DEV)
    Ld, _ = plan.forward(d, vel.to(DEV), eta=eta, with_fold=False)
    Ltm = plan.adjoint(m, vel.to(DEV), eta=eta)
    lhs = float((Ld.double() * m.double()).sum())
    rhs = float((d.double() * Ltm.double()).sum())
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"[dot flags={flags}] <Ld,m>={lhs:.9e} <d,L^T m>={rhs:.9e} rel={rel:.2e}")
    assert rel <= 1e-5


@pytest.mark.parametrize("n_samples", [1500, 2009, 2010, 2444, 2445, 3000, 3902, 3903, 5000])
def test_adjoint_dot_product_long_traces(n_samples):
    # plain adjoint tiers by trace length and their capacity edges: 4x2 tiles
    # at 7 per CU (1D velocity, traces up to 2009 samples), at 6 (2444), at 4
    # (3902), 8x4 tiles at 2 (6074)
    from seisproc_amd.processors.migration.hip_ops import PstmPlan
    s = S.tiny_survey(seed=23, n_bins=8, n_classes=2, n_samples=n_samples, n_img=6, aperture=300.0)
    p = s.params.replace(n_tau=min(n_samples, 900))
    vel = s.vel[:p.n_tau].contiguous().to(DEV)
    plan = PstmPlan(p, s.src_xy, s.rcv_xy)
    g = torch.Generator(device=DEV).manual_seed(6)
    d = torch.randn((p.n_traces, p.n_samples), generator=g, device=DEV)
    m = torch.randn(plan.image_shape, generator=g, device=DEV)
    Ld, _ = plan.forward(d, vel, with_fold=False)
    Ltm = plan.adjoint(m, vel)
    lhs = float((Ld.double() * m.double()).sum())
    rhs = float((d.double() * Ltm.double()).sum())
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"[dot n_samples={n_samples}] rel={rel:.2e}")
    assert rel <= 1e-5


@pytest.mark.parametrize("n_samples,dtau,flags", [(12001, 0.024, 0), (8001, 0.016, FLAG_AA)])
def test_adjoint_segmented_long_windows(n_samples, dtau, flags):
    # windows longer than the LDS share (one 256-time block spans the record):
    # the adjoint stages them in segments too; dot-product test and the
    # oracle's transpose
    from seisproc_amd.processors.migration.hip_ops import PstmPlan
    s = S.tiny_survey(seed=24, n_bins=6, n_classes=2, n_samples=n_samples, n_img=4, aperture=2500.0)
    n_tau = int(round((n_samples - 1) * 0.0005 / dtau))
    p = s.params.replace(dt=0.0005, n_tau=n_tau, dtau=dtau, flags=flags, max_angle_deg=85.0, vel_mode=0)
    vel = torch.full((n_tau,), 1500.0, device=DEV)
    plan = PstmPlan(p, s.src_xy, s.rcv_xy)
    g = torch.Generator(device=DEV).manual_seed(8)
    d = torch.randn((p.n_traces, p.n_samples), generator=g, device=DEV)
    m = torch.randn(plan.image_shape, generator=g, device=DEV)
    Ld, _ = plan.forward(d, vel, with_fold=False)
    Ltm = plan.adjoint(m, vel)
    lhs = float((Ld.double() * m.double()).sum())
    rhs = float((d.double() * Ltm.double()).sum())
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    sx, sy, rx, ry = s.numpy_geometry()
    ref = oracle.adjoint(p, m.double().cpu().numpy(), sx, sy, rx, ry, vel.double().cpu().numpy())
    err = np.linalg.norm(Ltm.double().cpu().numpy() - ref) / np.linalg.norm(ref)
    print(f"[segmented adjoint n={n_samples} flags={flags}] dot rel={rel:.2e} oracle rel_rms={err:.3e}")
    assert rel <= 1e-5 and err <= 1e-4


@pytest.mark.parametrize("flags", [0, FLAG_AA, FLAG_TAPER])
def test_adjoint_matches_oracle(flags):
    win = dict(min_offset_m=100.0, max_offset_m=500.0, taper_width=0.3) if flags & FLAG_TAPER else {}
    s, plan, p, vel = _setup(flags, **win)
    g = torch.Generator(device=DEV).manual_seed(4)
    m = torch.randn(plan.image_shape, generator=g, device=DEV)
    out = plan.adjoint(m, vel.to(DEV))
    sx, sy, rx, ry = s.numpy_geometry()
    ref = oracle.adjoint(p, m.double().cpu().numpy(), sx, sy, rx, ry, vel.double().numpy())
    err = np.linalg.norm(out.double().cpu().numpy() - ref) / np.linalg.norm(ref)
    print(f"[adjoint flags={flags}] rel_rms={err:.3e}")
    assert err <= 1e-4


@pytest.mark.parametrize("wm", [0, WEIGHT_SPREADING, 3])
def test_vel_grad_matches_oracle(wm):
    s, plan, p, vel = _setup(0, weight_mode=wm)
    g = torch.Generator(device=DEV).manual_seed(5)
    up = torch.randn(plan.image_shape, generator=g, device=DEV)
    grad = plan.vel_grad(s.traces.to(DEV), vel.to(DEV), up, flags=0)
    sx, sy, rx, ry = s.numpy_geometry()
    ref = oracle.vel_grad(p.replace(weight_mode=wm), s.traces.double().numpy(), sx, sy, rx, ry,
                          vel.double().numpy(), up.double().cpu().numpy())
    err = np.linalg.norm(grad.double().cpu().numpy() - ref) / np.linalg.norm(ref)
    print(f"[vel_grad wm={wm}] rel_rms={err:.3e}")
    # The derivative of linear interpolation jumps at every sample knot; FP32
    # traveltimes differ from the FP64 oracle's by ~1e-4 samples, so the few
    # contributions that close to a knot take the neighbouring slope.  That
    # bounds agreement at the 1e-3 level rather than the 1e-4 of the image.
    assert err <= 3e-3
