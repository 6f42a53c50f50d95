"""GPU: autograd wrappers, least-squares migration and the image-power
velocity gradient, checked against the FP64 oracle."""
import numpy as np
import pytest
import torch

import oracle
from seisproc_amd import synthetic as S
from seisproc_amd.models.velocity_model import BSplineVelocity

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _survey(seed=30, lateral=True, n_img=10):
    from seisproc_amd.processors.migration.hip_ops import PstmPlan
    s = S.tiny_survey(seed=seed, n_bins=12, n_classes=2, n_samples=201, n_img=n_img, aperture=350.0,
                      lateral=lateral)
    vel = s.vel if s.vel.dim() == 3 else s.vel[None, None, :].expand(n_img, n_img, -1).contiguous()
    p = s.params.replace(vel_mode=1)
    return s, PstmPlan(p, s.src_xy, s.rcv_xy), p, vel.to(DEV)


def test_autograd_gradients_are_adjoint_and_vel_grad():
    from seisproc_amd.processors.migration.autograd import demigrate, migrate
    s, plan, p, vel = _survey()
    d = s.traces.to(DEV).clone().requires_grad_(True)
    v = vel.clone().requires_grad_(True)
    g = torch.randn(plan.image_shape, device=DEV, generator=torch.Generator(device=DEV).manual_seed(1))
    img = migrate(d, v, plan, flags=0)
    J = (img * g).sum()
    gd, gv = torch.autograd.grad(J, (d, v))
    # the adjoint accumulates with atomics (summation order varies run to run)
    ref_d = plan.adjoint(g, vel, flags=0)
    assert float((gd - ref_d).norm() / ref_d.norm()) < 1e-6
    torch.testing.assert_close(gv, plan.vel_grad(s.traces.to(DEV), vel, g, flags=0), rtol=0, atol=0)
    # demigration's gradient is migration
    m = torch.randn(plan.image_shape, device=DEV).requires_grad_(True)
    dd = demigrate(m, vel, plan, flags=0)
    h = torch.randn_like(dd)
    (gm,) = torch.autograd.grad((dd * h).sum(), m)
    ref, _ = plan.forward(h, vel, with_fold=False, flags=0)
    torch.testing.assert_close(gm, ref, rtol=0, atol=0)


def test_cgls_reduces_residual_and_recovers_reflectivity():
    from seisproc_amd.processors.migration.lsm import cgls
    s, plan, p, vel = _survey(seed=31, lateral=False)
    gen = torch.Generator(device=DEV).manual_seed(2)
    m_true = torch.zeros(plan.image_shape, device=DEV)
    spikes = torch.rand(plan.image_shape, device=DEV, generator=gen) < 0.01
    m_true[spikes] = torch.randn(int(spikes.sum()), device=DEV, generator=gen)
    data = plan.adjoint(m_true, vel, flags=0)
    res = cgls(plan, data, vel, n_iter=12, flags=0)
    r = res.data_residual_norms
    print(f"[cgls] residual {r[0]:.3e} -> {r[-1]:.3e} in {res.iterations} its, {res.seconds_per_iteration*1e3:.1f} ms/it")
    assert all(b <= a * (1 + 1e-6) for a, b in zip(r, r[1:]))
    assert r[-1] < 0.5 * r[0]
    # the LSM image fits the data better than the plain migration image (best scaled)
    mig, _ = plan.forward(data, vel, with_fold=False, flags=0)
    dm = plan.adjoint(mig, vel, flags=0)
    a = float((dm * data).sum() / (dm * dm).sum())
    assert r[-1] < float((data - a * dm).norm())


def test_image_power_gradient_matches_oracle_finite_differences():
    # J(c) = sum image(B c)^2; masks frozen at the base model.  Central
    # differences of the FP64 oracle along a random coefficient direction.
    from seisproc_amd.processors.migration.velocity_update import image_power
    s, plan, p, _ = _survey(seed=32, n_img=8)
    bs = BSplineVelocity(p.n_il, p.n_xl, p.n_tau, shape=(4, 4, 6), device=DEV)
    coef = torch.full((4, 4, 6), 2000.0, device=DEV)
    coef += 300.0 * torch.linspace(0, 1, 6, device=DEV)[None, None, :]
    tr = s.traces.to(DEV)
    J, g = image_power(plan, tr, bs, coef)
    rng = np.random.default_rng(3)
    delta = torch.as_tensor(rng.standard_normal((4, 4, 6)), dtype=torch.float32, device=DEV)
    an = float((g.double() * delta.double()).sum())
    sx, sy, rx, ry = s.numpy_geometry()
    d = s.traces.double().numpy()
    v0 = bs.evaluate(coef).double().cpu().numpy()

    dv = bs.evaluate(delta).double().cpu().numpy()  # v is linear in the coefficients

    def J_oracle(h):
        v = v0 + h * dv  # FP64: an FP32 v would round away most of a 0.01 m/s step
        img, _ = oracle.forward(p, d, sx, sy, rx, ry, v, mask_vel=v0)
        return float((img ** 2).sum())

    for h, tol in ((1e-2, 2e-3), (1.0, 3e-2)):
        fd = (J_oracle(h) - J_oracle(-h)) / (2 * h)
        rel = abs(fd - an) / abs(fd)
        print(f"[image power] h={h} m/s fd={fd:.6e} analytic={an:.6e} rel={rel:.2e}")
        assert rel < tol
