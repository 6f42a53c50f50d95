"""GPU: the drop-in surface (MigrationEngine.migrate_bin_full,
migrate_kirchhoff_full, PSTMProcessor.process, KirchhoffMigrator) against the
FP64 oracle and against each other."""
import numpy as np
import pytest
import torch

import oracle
from seisproc_amd import synthetic as S
from seisproc_amd.models import (AnisotropyModel, MigrationConfig, MigrationGeometry, OutputGrid, SeismicData,
                                 VelocityModel, WeightMode)

pytestmark = pytest.mark.gpu


def _zero_offset_survey():
    s = S.tiny_survey(seed=40, n_bins=12, n_classes=2, n_samples=251, n_img=10, aperture=400.0)
    return s


def test_engine_migrate_bin_full_matches_oracle():
    from seisproc_amd.processors.migration import MigrationEngine
    s = _zero_offset_survey()
    sx, sy, rx, ry = s.numpy_geometry()
    p = s.params
    eng = MigrationEngine()
    img, fold = eng.migrate_bin_full(s.traces.t().numpy(), sx, sy, rx, ry, p.origin_x, p.origin_y, p.il_dx,
                                     p.xl_dy, 0.0, p.n_il, p.n_xl, p.dt * 1000.0, 0.0, velocity_mps=2000.0,
                                     max_aperture_m=p.aperture_m, max_angle_deg=p.max_angle_deg, normalize=False)
    assert img.shape == (p.n_tau, p.n_il, p.n_xl)
    ref, rfold = oracle.forward(p, s.traces.double().numpy(), sx, sy, rx, ry, np.full(p.n_tau, 2000.0))
    ref = ref.transpose(2, 0, 1)
    rfold = rfold.transpose(2, 0, 1)
    assert np.abs(fold - rfold).sum() <= 1e-4 * rfold.sum() + 1
    assert np.linalg.norm(img - ref) / np.linalg.norm(ref) < 1e-4
    norm, _ = eng.migrate_bin_full(s.traces.t().numpy(), sx, sy, rx, ry, p.origin_x, p.origin_y, p.il_dx, p.xl_dy,
                                   0.0, p.n_il, p.n_xl, p.dt * 1000.0, 0.0, velocity_mps=2000.0,
                                   max_aperture_m=p.aperture_m, normalize=True, min_fold=1)
    mask = rfold >= 1
    np.testing.assert_allclose(norm[mask], (img / np.maximum(fold, 1))[mask], rtol=1e-6)
    np.testing.assert_array_equal(norm[~mask], img[~mask])
    # trace batches accumulate to the same image
    bimg, bfold = eng.migrate_bin_batched(s.traces.t().numpy(), sx, sy, rx, ry, p.origin_x, p.origin_y, p.il_dx,
                                          p.xl_dy, 0.0, p.n_il, p.n_xl, p.dt * 1000.0, 0.0, 2000.0,
                                          max_aperture_m=p.aperture_m, normalize=False, batch_size=100)
    np.testing.assert_array_equal(bfold, fold)
    assert np.linalg.norm(bimg - img) / np.linalg.norm(img) < 1e-5
    with pytest.raises(NotImplementedError):
        eng.migrate_bin_time_domain()


def test_engine_rms_velocity_and_bf16():
    from seisproc_amd.processors.migration import MigrationEngine
    s = S.tiny_survey(seed=41, n_bins=12, n_classes=2, n_samples=251, n_img=10, aperture=400.0, lateral=True)
    sx, sy, rx, ry = s.numpy_geometry()
    p = s.params
    eng = MigrationEngine()
    kw = dict(traces=s.traces.t().numpy(), source_x=sx, source_y=sy, receiver_x=rx, receiver_y=ry,
              origin_x=p.origin_x, origin_y=p.origin_y, il_spacing=p.il_dx, xl_spacing=p.xl_dy, azimuth_deg=0.0,
              n_il=p.n_il, n_xl=p.n_xl, dt_ms=p.dt * 1000.0, t_min_ms=0.0, max_aperture_m=p.aperture_m,
              normalize=False)
    img, fold = eng.migrate_bin_full(**kw, velocity_model=s.vel)
    ref, _ = oracle.forward(p.replace(vel_mode=1), s.traces.double().numpy(), sx, sy, rx, ry, s.vel.double().numpy())
    ref = ref.transpose(2, 0, 1)
    assert np.linalg.norm(img - ref) / np.linalg.norm(ref) < 1e-4
    b16, _ = eng.migrate_bin_full(**kw, velocity_model=s.vel, trace_dtype="bf16")
    e16 = np.linalg.norm(b16 - ref) / np.linalg.norm(ref)
    print(f"[engine bf16] rel_rms vs FP64 oracle: {e16:.3e}")
    assert e16 < 1e-3
    # 1D RMS model object
    vm = VelocityModel(data=np.array([1800.0, 2300.0], np.float32), z_axis=np.array([0.0, 1.0]))
    img1, _ = eng.migrate_bin_full(**kw, velocity_model=vm)
    v1 = vm.rms_velocity_at_time(np.arange(p.n_tau) * p.dt)
    ref1, _ = oracle.forward(p.replace(vel_mode=0), s.traces.double().numpy(), sx, sy, rx, ry, v1)
    assert np.linalg.norm(img1 - ref1.transpose(2, 0, 1)) / np.linalg.norm(ref1) < 1e-4


def test_functional_migrate_kirchhoff_full_reference_signature():
    from seisproc_amd.processors.migration import migrate_kirchhoff_full
    s = _zero_offset_survey()
    p = s.params
    sx, sy, rx, ry = (torch.tensor(a, dtype=torch.float32) for a in s.numpy_geometry())
    il = torch.arange(p.n_il, dtype=torch.float32) * p.il_dx + p.origin_x
    xl = torch.arange(p.n_xl, dtype=torch.float32) * p.xl_dy + p.origin_y
    xg, ig = torch.meshgrid(xl, il, indexing="xy")
    depth = 2000.0 * (torch.arange(p.n_tau, dtype=torch.float32) * p.dt) / 2.0
    img, fold = migrate_kirchhoff_full(s.traces.t().contiguous(), sx, sy, rx, ry, ig.flatten(), xg.flatten(), depth,
                                       2000.0, p.dt * 1000.0, 0.0, p.aperture_m, p.max_angle_deg, p.n_il, p.n_xl)
    assert img.device.type == "cpu" and img.shape == (p.n_tau, p.n_il, p.n_xl)
    ref, _ = oracle.forward(p, s.traces.double().numpy(), *s.numpy_geometry(), np.full(p.n_tau, 2000.0))
    got = img.permute(1, 2, 0).double().numpy()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-4
    with pytest.raises(ValueError):
        migrate_kirchhoff_full(s.traces.t().contiguous(), sx, sy, rx, ry, torch.rand(p.n_il * p.n_xl),
                               xg.flatten(), depth, 2000.0, p.dt * 1000.0, 0.0, p.aperture_m, 60.0, p.n_il, p.n_xl)


def _seismic_data(s):
    sx, sy, rx, ry = s.numpy_geometry()
    return SeismicData(traces=s.traces.t().numpy().copy(), sample_rate=s.params.dt * 1000.0,
                       headers={"SourceX": sx, "SourceY": sy, "GroupX": rx, "GroupY": ry})


def test_pstm_processor_process_and_registry():
    from seisproc_amd.processors import BaseProcessor, create_processor
    s = _zero_offset_survey()
    p = s.params
    g = OutputGrid(n_time=p.n_tau, n_inline=p.n_il, n_xline=p.n_xl, dt=p.dtau, d_inline=p.il_dx, d_xline=p.xl_dy,
                   x_origin=p.origin_x, y_origin=p.origin_y)
    cfg = MigrationConfig(output_grid=g, max_aperture_m=p.aperture_m, weight_mode=WeightMode.FULL,
                          antialias_enabled=False, normalize_by_fold=False)
    proc = create_processor("PSTMProcessor", config=cfg, velocity=2000.0)
    data = _seismic_data(s)
    out = proc.process(data)
    assert out.traces.shape == (p.n_tau, p.n_il * p.n_xl) and out.sample_rate == p.dtau * 1000.0
    assert out.headers["inline"][p.n_xl] == 2 and out.headers["crossline"][1] == 2
    ref, rfold = oracle.forward(p, s.traces.double().numpy(), *s.numpy_geometry(), np.full(p.n_tau, 2000.0))
    ref = ref.transpose(2, 0, 1).reshape(p.n_tau, -1)
    assert np.linalg.norm(out.traces - ref) / np.linalg.norm(ref) < 1e-4
    assert out.metadata["contributions"] == pytest.approx(rfold.sum(), rel=1e-4)
    assert data.traces is not out.traces  # input untouched
    again = BaseProcessor.from_dict(proc.to_dict()).process(data)
    np.testing.assert_array_equal(again.traces, out.traces)


def test_kirchhoff_migrator_dataset_equals_single_gather():
    from seisproc_amd.processors import create_processor
    s = _zero_offset_survey()
    p = s.params
    g = OutputGrid(n_time=p.n_tau, n_inline=p.n_il, n_xline=p.n_xl, dt=p.dtau, d_inline=p.il_dx, d_xline=p.xl_dy,
                   x_origin=p.origin_x, y_origin=p.origin_y)
    cfg = MigrationConfig(output_grid=g, max_aperture_m=p.aperture_m, antialias_enabled=False)
    mig = create_processor("KirchhoffMigrator", velocity=VelocityModel(2000.0), config=cfg)
    sx, sy, rx, ry = s.numpy_geometry()
    n0 = s.class_sizes[0]
    tr = s.traces.t().numpy()
    gathers = [SeismicData(tr[:, :n0].copy(), p.dt * 1000.0), SeismicData(tr[:, n0:].copy(), p.dt * 1000.0)]
    geos = [MigrationGeometry(sx[:n0], sy[:n0], rx[:n0], ry[:n0]), MigrationGeometry(sx[n0:], sy[n0:], rx[n0:], ry[n0:])]
    res = mig.migrate_dataset(gathers, geos)
    one = mig.migrate_dataset([SeismicData(tr.copy(), p.dt * 1000.0)], [MigrationGeometry(sx, sy, rx, ry)])
    np.testing.assert_array_equal(res.fold, one.fold)
    assert np.linalg.norm(res.image - one.image) / np.linalg.norm(one.image) < 1e-5
    # SPREADING weight (reference migrator) against the oracle
    ref, rfold = oracle.forward(p.replace(weight_mode=1), s.traces.double().numpy(), sx, sy, rx, ry,
                                np.full(p.n_tau, 2000.0))
    refn = np.where(rfold >= 1, ref / np.maximum(rfold, 1), 0.0).transpose(2, 0, 1)
    assert np.linalg.norm(one.image - refn) / np.linalg.norm(refn) < 1e-4


def test_optimized_migrator_factory_equals_class_api():
    # the interactive path's factory (main_window.py:4642-4653 calls it with
    # prefer_gpu and optimization_level='high') runs the same HIP sum as
    # KirchhoffMigrator: bit-identical image and fold at every level
    from seisproc_amd.processors.migration import KirchhoffMigrator, create_optimized_kirchhoff_migrator
    s = _zero_offset_survey()
    p = s.params
    g = OutputGrid(n_time=p.n_tau, n_inline=p.n_il, n_xline=p.n_xl, dt=p.dtau, d_inline=p.il_dx, d_xline=p.xl_dy,
                   x_origin=p.origin_x, y_origin=p.origin_y)
    cfg = MigrationConfig(output_grid=g, max_aperture_m=p.aperture_m, antialias_enabled=False)
    sx, sy, rx, ry = s.numpy_geometry()
    gather = SeismicData(s.traces.t().numpy().copy(), p.dt * 1000.0)
    geo = MigrationGeometry(sx, sy, rx, ry)
    ref_img, ref_fold = KirchhoffMigrator(VelocityModel(2000.0), cfg).migrate_gather(gather, geo)
    for level in ("low", "medium", "high"):
        mig = create_optimized_kirchhoff_migrator(velocity=VelocityModel(2000.0), config=cfg, prefer_gpu=True,
                                                  optimization_level=level)
        img, fold = mig.migrate_gather(gather=gather, geometry=geo)
        assert torch.equal(img, ref_img) and torch.equal(fold, ref_fold)
        st = mig.get_optimization_stats()
        assert st.total_traces == s.n_traces and st.traces_processed == s.n_traces
        assert "GPUTiling=" + str(level == "high") in mig.get_description()
        # bin lifecycle and serialisation of the interface
        mig.initialize_for_bin(geo)
        mig.reset_for_new_bin()
        assert mig.get_optimization_stats().total_traces == 0
        d = mig.to_dict()
        assert d["class_name"] == "OptimizedKirchhoffMigrator" and d["config"] == cfg.to_dict()
    # SPREADING weight against the oracle (raw image and fold, [tau][il][xl])
    ref, rfold = oracle.forward(p.replace(weight_mode=1), s.traces.double().numpy(), sx, sy, rx, ry,
                                np.full(p.n_tau, 2000.0))
    np.testing.assert_array_equal(ref_fold.cpu().numpy(), rfold.transpose(2, 0, 1))
    r = ref.transpose(2, 0, 1)
    assert np.linalg.norm(ref_img.cpu().numpy() - r) / np.linalg.norm(r) < 1e-4


def test_processor_vti_with_anisotropy_model():
    from seisproc_amd.processors import create_processor
    s = _zero_offset_survey()
    p = s.params
    g = OutputGrid(n_time=p.n_tau, n_inline=p.n_il, n_xline=p.n_xl, dt=p.dtau, d_inline=p.il_dx, d_xline=p.xl_dy,
                   x_origin=p.origin_x, y_origin=p.origin_y)
    cfg = MigrationConfig(output_grid=g, max_aperture_m=p.aperture_m, weight_mode=WeightMode.FULL,
                          antialias_enabled=True, use_anisotropy=True, normalize_by_fold=False)
    proc = create_processor("PSTMProcessor", config=cfg, velocity=2000.0, anisotropy=AnisotropyModel(0.1, 0.0),
                            aa_max=8.0)
    out = proc.process(_seismic_data(s))
    pp = p.replace(flags=1 | 2, vel_mode=0)
    ref, _ = oracle.forward(pp, s.traces.double().numpy(), *s.numpy_geometry(), np.full(p.n_tau, 2000.0),
                            eta=np.full(p.n_tau, 0.1))
    ref = ref.transpose(2, 0, 1).reshape(p.n_tau, -1)
    assert np.linalg.norm(out.traces - ref) / np.linalg.norm(ref) < 1e-4
