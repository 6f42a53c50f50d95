"""The torch-CPU restatement of the reference tile loop (the timed CPU
baseline, oracle/ref_torch.py) agrees with the FP64 oracle."""
import numpy as np
import torch

import oracle
from oracle.ref_torch import demigrate_kirchhoff_full_cpu, migrate_kirchhoff_full_cpu
from seisproc_amd import synthetic as S


def test_ref_torch_matches_oracle(oracle_lib):
    s = S.tiny_survey(seed=21, n_bins=8, n_classes=2, n_samples=201, n_img=6, aperture=300.0)
    p = s.params
    v = 2000.0
    vel = np.full(p.n_tau, v)
    sx, sy, rx, ry = (torch.tensor(a, dtype=torch.float32) for a in s.numpy_geometry())
    il = torch.arange(p.n_il, dtype=torch.float32) * p.il_dx + p.origin_x
    xl = torch.arange(p.n_xl, dtype=torch.float32) * p.xl_dy + p.origin_y
    ox = il[:, None].expand(-1, p.n_xl).reshape(-1)
    oy = xl[None, :].expand(p.n_il, -1).reshape(-1)
    tau_ms = torch.arange(p.n_tau, dtype=torch.float32) * p.dtau * 1000.0
    depth = v * (tau_ms / 1000.0) / 2.0
    img, fold = migrate_kirchhoff_full_cpu(s.traces.t().contiguous(), sx, sy, rx, ry, ox, oy, depth, v,
                                           p.dt * 1000.0, 0.0, p.aperture_m, p.max_angle_deg, p.n_il, p.n_xl)
    ref, rfold = oracle.forward(p, s.traces.double().numpy(), *s.numpy_geometry(), vel)
    got = img.permute(1, 2, 0).double().numpy()
    gf = fold.permute(1, 2, 0).double().numpy()
    assert rfold.sum() > 0
    assert np.abs(gf - rfold).sum() <= 1e-4 * rfold.sum() + 1
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-4, err


def test_ref_torch_adjoint_is_the_transpose_and_matches_oracle(oracle_lib):
    # CPU side of the least-squares iteration baseline: <L d, m> = <d, L^T m>
    # against the forward restatement, and L^T m against the FP64 oracle adjoint
    s = S.tiny_survey(seed=22, n_bins=8, n_classes=2, n_samples=201, n_img=6, aperture=300.0)
    p = s.params
    v = 2000.0
    sx, sy, rx, ry = (torch.tensor(a, dtype=torch.float32) for a in s.numpy_geometry())
    il = torch.arange(p.n_il, dtype=torch.float32) * p.il_dx + p.origin_x
    xl = torch.arange(p.n_xl, dtype=torch.float32) * p.xl_dy + p.origin_y
    ox = il[:, None].expand(-1, p.n_xl).reshape(-1)
    oy = xl[None, :].expand(p.n_il, -1).reshape(-1)
    depth = v * (torch.arange(p.n_tau, dtype=torch.float32) * p.dtau) / 2.0
    args = (sx, sy, rx, ry, ox, oy, depth, v, p.dt * 1000.0, 0.0, p.aperture_m, p.max_angle_deg, p.n_il, p.n_xl)
    d = s.traces.t().contiguous()
    m = torch.randn(p.n_tau, p.n_il, p.n_xl, generator=torch.Generator().manual_seed(5))
    Ld, _ = migrate_kirchhoff_full_cpu(d, *args)
    Ltm = demigrate_kirchhoff_full_cpu(m, p.n_samples, *args)
    lhs = float((Ld.double() * m.double()).sum())
    rhs = float((d.double() * Ltm.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs)), (lhs, rhs)
    ref = oracle.adjoint(p, m.permute(1, 2, 0).double().numpy(), *s.numpy_geometry(), np.full(p.n_tau, v))
    got = Ltm.t().double().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-4, err
