"""Model / configuration layer (CPU).  Known answers restated from the
reference's tests:
  tests/test_config_adapter.py      wizard mapping, counts, validation
  tests/test_velocity_model.py      constant / gradient v_rms
  tests/test_anisotropy_model.py    eta formula, 1D interpolation
  tests/test_kirchhoff_kernel.py:386-406  fold normalisation
"""
import json
import math

import numpy as np
import pytest
import torch

from seisproc_amd.abi_types import FLAG_AA, FLAG_KAHAN, FLAG_VTI, WEIGHT_FULL, WEIGHT_SPREADING
from seisproc_amd.models import (AnisotropyModel, MigrationConfig, MigrationJobConfig, OutputGrid, SeismicData,
                                 VelocityModel, WeightMode)
from seisproc_amd.models.anisotropy_model import create_eta_field
from seisproc_amd.models.velocity_model import (BSplineVelocity, create_linear_gradient_velocity,
                                                interval_to_rms_velocity, rms_to_interval_velocity)
from seisproc_amd.processors.migration.base_migrator import MigrationResult
from seisproc_amd.processors.migration.config_adapter import ConfigAdapter
from seisproc_amd.processors.migration.kirchhoff_kernel import normalize_by_fold
from seisproc_amd.processors.migration import velocity_model as engine_vm

WIZARD = {"name": "Test Migration", "x_origin": 1000.0, "y_origin": 2000.0, "output_bin_il": 25.0,
          "output_bin_xl": 25.0, "grid_azimuth_deg": 45.0, "inline_min": 1, "inline_max": 100, "inline_step": 1,
          "xline_min": 1, "xline_max": 100, "xline_step": 1, "time_min_ms": 0, "time_max_ms": 4000,
          "dt_ms": 4.0, "velocity_type": "constant", "velocity_v0": 3000.0, "max_aperture_m": 5000.0,
          "max_angle_deg": 60.0}


# ---------------------------------------------------------------- ConfigAdapter
def test_adapter_extracts_wizard_values():
    p = ConfigAdapter(WIZARD).params
    assert (p.origin_x, p.origin_y, p.il_spacing, p.xl_spacing, p.azimuth_deg) == (1000.0, 2000.0, 25.0, 25.0, 45.0)
    assert (p.n_il, p.n_xl, p.dt_ms, p.t_min_ms, p.n_samples) == (100, 100, 4.0, 0.0, 1001)
    assert (p.velocity_mps, p.max_aperture_m, p.max_angle_deg) == (3000.0, 5000.0, 60.0)


def test_adapter_counts_with_steps_and_bin_fallbacks():
    p = ConfigAdapter({"inline_min": 1, "inline_max": 199, "inline_step": 2, "xline_min": 1, "xline_max": 301,
                       "xline_step": 3}).params
    assert (p.n_il, p.n_xl) == (100, 101)
    assert ConfigAdapter({"input_bin_il": 50.0, "input_bin_xl": 50.0}).params.il_spacing == 50.0
    assert ConfigAdapter({}).params.xl_spacing == 25.0


def test_adapter_engine_params_and_validation():
    ad = ConfigAdapter(dict(WIZARD, output_bin_il=50.0, output_bin_xl=50.0, max_aperture_m=3000.0))
    tr = np.zeros((10, 3), np.float32)
    sx = np.arange(3.0)
    kw = ad.get_engine_params(tr, sx, sx, sx, sx, normalize=False, min_fold=3)
    assert kw["il_spacing"] == 50.0 and kw["n_times"] == 1001 and kw["normalize"] is False and kw["min_fold"] == 3
    assert np.array_equal(kw["traces"], tr)
    ok, msg = ConfigAdapter({k: v for k, v in WIZARD.items() if k != "grid_azimuth_deg"}).validate()
    assert ok, msg
    ok, msg = ConfigAdapter({"velocity_v0": 0.0}).validate()
    assert not ok and "velocity" in msg.lower()
    ok, msg = ConfigAdapter({"velocity_v0": 500.0}).validate()
    assert not ok and "unusual" in msg.lower()
    ok, msg = ConfigAdapter({"inline_min": 100, "inline_max": 10}).validate()
    assert not ok and "inline" in msg.lower()
    ok, msg = ConfigAdapter({"inline_max": 100000, "xline_max": 100000}).validate()
    assert not ok and "large" in msg.lower()



def test_adapter_summary_names_each_parameter():
    # the reference's summary checks (tests/test_config_adapter.py:366-410): the
    # grid size, velocity, aperture and bin size appear, and each section is named
    s = ConfigAdapter(WIZARD).get_summary()
    assert isinstance(s, str)
    for value in ("100", "3000", "5000", "25"):
        assert value in s
    s = ConfigAdapter({"velocity_v0": 2500.0}).get_summary().lower()
    for word in ("grid", "origin", "velocity", "aperture"):
        assert word in s


def test_model_summaries():
    # the reference's anisotropy summary checks (tests/test_anisotropy_model.py:
    # 315-331): the kind is named and the constant's epsilon is shown
    from seisproc_amd.models.anisotropy_model import AnisotropyModel, create_constant_anisotropy
    s = create_constant_anisotropy(0.2, 0.1).get_summary()
    assert "constant" in s and "0.2" in s
    z = np.array([0.0, 1.0, 2.0])
    assert "v_of_z" in AnisotropyModel(epsilon=0.1 + 0.05 * z, delta=0.05 + 0.02 * z, z_axis=z).get_summary()
    from seisproc_amd.models.migration_config import MigrationConfig
    c = MigrationConfig().get_summary().lower()
    for word in ("aperture", "offsets", "traveltime", "weights", "antialias"):
        assert word in c
    assert "2500" in engine_vm.VelocityModel("constant", 2500.0).get_summary()

# ---------------------------------------------------------------- MigrationConfig
def test_migration_config_validation_and_roundtrip():
    with pytest.raises(ValueError, match="max_angle_deg"):
        MigrationConfig(max_angle_deg=95.0)
    with pytest.raises(ValueError, match="min_offset_m"):
        MigrationConfig(min_offset_m=200.0, max_offset_m=100.0)
    with pytest.raises(ValueError, match="n_time"):
        OutputGrid(n_time=0)
    c = MigrationConfig(weight_mode=WeightMode.FULL, max_aperture_m=1234.0)
    c2 = MigrationConfig.from_dict(json.loads(json.dumps(c.to_dict())))
    assert c2.to_dict() == c.to_dict()


def test_config_to_kernel_params():
    g = OutputGrid(n_time=501, n_inline=10, n_xline=20, dt=0.002, d_inline=12.5, d_xline=25.0, t0=0.1,
                   x_origin=5e5, y_origin=6.7e6, inline_azimuth=30.0)
    c = MigrationConfig(output_grid=g, weight_mode=WeightMode.FULL, antialias_enabled=True, use_anisotropy=True)
    p = c.to_pstm_params(100, 1001, 0.004)
    assert (p.n_il, p.n_xl, p.n_tau, p.tau0, p.dtau, p.dt) == (10, 20, 501, 0.1, 0.002, 0.004)
    assert p.weight_mode == WEIGHT_FULL and p.flags & FLAG_AA and p.flags & FLAG_VTI and p.flags & FLAG_KAHAN
    # the reference kernels ignore the azimuth: x = x0 + il d_il, y = y0 + xl d_xl
    assert (p.il_dx, p.il_dy, p.xl_dx, p.xl_dy) == (12.5, 0.0, 0.0, 25.0)
    # honoring the azimuth reproduces OutputGrid.get_coordinates
    q = c.to_pstm_params(100, 1001, 0.004, honor_azimuth=True)
    for i, j in ((3, 7), (9, 0), (0, 19)):
        x, y = g.get_coordinates(i, j)
        assert abs(q.origin_x + i * q.il_dx + j * q.xl_dx - x) < 1e-6
        assert abs(q.origin_y + i * q.il_dy + j * q.xl_dy - y) < 1e-6
    assert MigrationConfig().to_pstm_params(1, 10, 0.002).weight_mode == WEIGHT_SPREADING


# ---------------------------------------------------------------- velocity
def test_engine_velocity_model_known_answers():
    m = engine_vm.VelocityModel("constant", 2500.0)
    assert m.rms_velocity_at_time(1000.0) == 2500.0
    assert np.allclose(m.rms_velocity_at_time(np.array([0.0, 500.0])), 2500.0)
    g = engine_vm.VelocityModel("gradient", 2000.0, gradient=0.5)
    assert g.interval_velocity(0.0) == 2000.0 and g.interval_velocity(1000.0) == 2500.0
    assert g.rms_velocity_at_time(0.0) == 2000.0
    v = [g.rms_velocity_at_time(t) for t in (500.0, 1000.0, 2000.0)]
    assert v[0] < v[1] < v[2]
    assert 1500.0 < engine_vm.VelocityModel("gradient", 1500.0, gradient=0.5).rms_velocity_at_time(2000.0) < 2500.0
    t = torch.tensor([0.0, 500.0, 1000.0, 2000.0])
    vt = g.rms_velocity_at_time(t)
    assert isinstance(vt, torch.Tensor) and torch.isclose(vt[0], torch.tensor(2000.0)) and bool((vt[1:] > vt[:-1]).all())
    # closed form at t = 1 s two-way: z = v0 t + k v0 t^2 / 2 with t = 0.5 s one-way
    z = 2000.0 * 0.5 + 0.5 * 0.5 * 2000.0 * 0.25
    assert abs(g.rms_velocity_at_time(1000.0) - math.sqrt(2000.0**2 + 2000.0 * 0.5 * z + (0.5 * z) ** 2 / 3)) < 1e-6



def test_depth_interval_model_to_vrms_closed_forms():
    # depth-domain interval velocities -> v_rms at two-way times (the engine
    # model's table, reference processors/migration/velocity_model.py:234-261)
    two = VelocityModel(data=np.array([2000.0, 2000.0, 3000.0, 3000.0], np.float32),
                        z_axis=np.array([0.0, 1000.0, 1000.001, 2000.0]), is_time=False)
    t1, t2 = 0.5, 1000.0 / 3000.0  # one-way times of the two layers
    want = math.sqrt((2000.0 ** 2 * t1 + 3000.0 ** 2 * t2) / (t1 + t2))
    got = two.rms_velocity_at_time(np.array([0.0, 2 * t1, 2 * (t1 + t2)]))
    np.testing.assert_allclose(got, [2000.0, 2000.0, want], rtol=2e-6)
    # v(z) = v0 + k z: one-way time t = ln(1 + k z / v0) / k, v(t) = v0 e^(k t),
    # v_rms^2 = v0^2 (e^(2 k t) - 1) / (2 k t)
    v0, k = 1800.0, 0.6
    z = np.arange(0.0, 4000.0 + 0.5, 0.5)
    lin = VelocityModel(data=(v0 + k * z).astype(np.float32), z_axis=z, is_time=False)
    t = np.array([0.3, 1.0, 1.4])  # the profile ends at t = ln(1 + k 4000 / v0) / k = 1.41 s
    want = v0 * np.sqrt(np.expm1(2 * k * t) / (2 * k * t))
    np.testing.assert_allclose(lin.rms_velocity_at_time(2 * t), want, rtol=1e-6)
    # v_rms, not the interval velocity v0 e^(k t) at that time (3280 m/s at 1 s)
    assert v0 * math.exp(k * 1.0) - lin.rms_velocity_at_time(np.array([2.0]))[0] > 700.0

def test_velocity_model_types_and_dix_roundtrip():
    assert VelocityModel(2000.0).is_constant
    m = create_linear_gradient_velocity(1500.0, 600.0, 3.0, 0.002)
    assert abs(m.rms_velocity_at_time(np.array([1.0]))[0] - 2100.0) < 1e-3
    t = np.linspace(0.0, 3.0, 31)
    vi = 1500.0 + 400.0 * t
    _, vr = interval_to_rms_velocity(t, vi)
    _, vi2 = rms_to_interval_velocity(t, vr)
    np.testing.assert_allclose(vi2[1:], vi[1:], rtol=2e-3)
    v3 = VelocityModel(data=np.full((3, 4, 5), 2000.0, np.float32), z_axis=np.linspace(0, 4, 3),
                       x_axis=np.arange(4.0) * 100, y_axis=np.arange(5.0) * 100)
    assert np.allclose(v3.get_velocity_at(1.0, 150.0, 220.0), 2000.0)
    d = VelocityModel.from_dict(json.loads(json.dumps(m.to_dict())))
    np.testing.assert_allclose(d.data, m.data)



def test_lateral_and_layered_velocity_factories():
    # 2D/3D fields (z first), the 2D gradient and layered v(z) (reference
    # models/velocity_model.py:516-700)
    from seisproc_amd.models.velocity_model import (VelocityType, create_2d_gradient_velocity, create_2d_velocity,
                                                    create_3d_velocity, create_layered_velocity)
    z, x, y = np.array([0.0, 1.0, 2.0]), np.array([0.0, 100.0]), np.array([0.0, 50.0, 100.0, 150.0])
    m2 = create_2d_velocity(np.array([[1500, 1600], [2000, 2100], [2500, 2600]]), z, x)
    assert m2.velocity_type is VelocityType.V_OF_XZ and m2.metadata["model_type"] == "2d_field"
    assert float(m2.get_velocity_at(0.5, 50.0)[0]) == pytest.approx(1800.0)
    with pytest.raises(ValueError, match="x-dimension"):
        create_2d_velocity(np.ones((3, 3)) * 2000, z, x)
    m3 = create_3d_velocity(np.full((3, 2, 4), 2000.0), z, x, y)
    assert m3.velocity_type is VelocityType.V_OF_XYZ and m3.data.shape == (3, 2, 4)
    with pytest.raises(ValueError, match="3D"):
        create_3d_velocity(np.ones((3, 2)), z, x, y)
    g = create_2d_gradient_velocity(1500.0, 500.0, 0.1, z_max=2.0, x_max=1000.0, dz=0.5, dx=250.0)
    assert g.data.shape == (5, 5) and g.data[4, 4] == pytest.approx(1500.0 + 500.0 * 2.0 + 0.1 * 1000.0)
    lay = create_layered_velocity([0.0, 0.5, 1.2], [1500.0, 2200.0, 3000.0], z_max=2.0, dz=0.1)
    assert lay.metadata["n_layers"] == 3
    np.testing.assert_array_equal(lay.data[[0, 4, 5, 11, 12, 20]], [1500, 1500, 2200, 2200, 3000, 3000])
    with pytest.raises(ValueError):
        create_layered_velocity([0.0, 1.0], [1500.0], z_max=2.0)

def test_bspline_velocity_partition_of_unity_and_fit():
    b = BSplineVelocity(20, 18, 60, shape=(6, 5, 8), device="cpu")
    coef = torch.full((6, 5, 8), 2500.0)
    assert torch.allclose(b.evaluate(coef), torch.full((20, 18, 60), 2500.0), atol=1e-3)
    ii = torch.linspace(0, 1, 20)[:, None, None]
    kk = torch.linspace(0, 1, 60)[None, None, :]
    v = (1800.0 + 300.0 * ii + 600.0 * kk).expand(20, 18, 60).contiguous()
    fit = b.evaluate(b.fit(v))
    assert float((fit - v).abs().max()) < 1.0  # cubic B-splines reproduce linear fields


# ---------------------------------------------------------------- anisotropy
def test_eta_formula_and_interpolation():
    assert abs(AnisotropyModel.compute_eta(0.25, 0.1) - (0.25 - 0.1) / 1.2) < 1e-10
    assert AnisotropyModel.compute_eta(0.0, 0.0) == 0.0
    assert abs(AnisotropyModel.compute_eta(0.2, 0.2)) < 1e-10
    m = AnisotropyModel(epsilon=np.array([0.1, 0.2, 0.3]), delta=np.array([0.05, 0.1, 0.15]),
                        z_axis=np.array([0.0, 1.0, 2.0]))
    eps, dl, eta = m.get_parameters_at(0.5)
    assert abs(eps - 0.15) < 1e-6 and abs(dl - 0.075) < 1e-6
    with pytest.raises(ValueError):
        AnisotropyModel(0.1, -0.6)



def test_anisotropy_factories_and_vti_velocities():
    # the reference's anisotropy tests restated (tests/test_anisotropy_model.py:
    # 110-280) plus identities that pin the relations independently
    from seisproc_amd.models import anisotropy_model as am
    m = am.create_1d_anisotropy(np.array([0.0, 1.0, 2.0]), np.array([0.1, 0.2, 0.3]), np.array([0.05, 0.1, 0.15]))
    assert m.anisotropy_type is am.AnisotropyType.V_OF_Z
    e, d, _ = m.get_parameters_at(0.5)
    assert abs(e - 0.15) < 1e-12 and abs(d - 0.075) < 1e-12
    g = am.create_gradient_anisotropy(0.1, 0.1, 0.05, 0.05, z_max=2.0, dz=1.0)
    assert g.anisotropy_type is am.AnisotropyType.V_OF_Z and g.epsilon[0] == 0.1
    assert abs(g.get_parameters_at(1.0)[0] - 0.2) < 1e-12
    assert [(s.epsilon, s.delta) for s in map(am.create_shale_anisotropy, ("weak", "moderate", "strong"))] == \
        [(0.1, 0.05), (0.2, 0.1), (0.35, 0.15)]
    with pytest.raises(ValueError):
        am.create_shale_anisotropy("invalid")
    with pytest.raises(ValueError, match="must match"):
        am.AnisotropyModel(epsilon=np.array([0.1, 0.2]), delta=np.array([0.05, 0.1, 0.15]))
    with pytest.raises(ValueError, match="Unstable"):
        am.AnisotropyModel(epsilon=0.0, delta=-0.6)
    v0, eps, dl = 2500.0, 0.2, 0.1
    assert abs(am.compute_vti_phase_velocity(v0, 0.0, eps, dl) - v0) < 1e-9
    vh = am.compute_horizontal_velocity(v0, eps)
    assert abs(am.compute_vti_phase_velocity(v0, np.pi / 2, eps, dl) - vh) < 1e-9
    vn = am.compute_nmo_velocity(v0, dl)
    # 1 + 2 eta = (v_horizontal / v_nmo)^2
    assert abs(1 + 2 * am.compute_effective_eta(eps, dl) - (vh / vn) ** 2) < 1e-12
    # weak anisotropy: the phase velocity is Thomsen's linear form to second order
    th = np.linspace(0.0, np.pi / 2, 7)
    lin = v0 * (1 + 1e-3 * np.sin(th) ** 2 * np.cos(th) ** 2 + 2e-3 * np.sin(th) ** 4)
    np.testing.assert_allclose(am.compute_vti_phase_velocity(v0, th, 2e-3, 1e-3), lin, rtol=1e-5)

def test_eta_field_on_grid():
    from seisproc_amd.abi_types import PstmParams
    p = PstmParams.from_dict(dict(n_il=3, n_xl=4, n_tau=5, dtau=0.5, il_dx=10.0, xl_dy=10.0))
    eta = np.random.default_rng(0).uniform(0, 0.15, (3, 4, 5))
    m = create_eta_field(eta, np.arange(3) * 10.0, np.arange(4) * 10.0, np.arange(5) * 0.5)
    np.testing.assert_allclose(m.eta_on_grid(p, device="cpu").numpy(), eta, atol=1e-6)
    c = AnisotropyModel(0.1, 0.0)
    assert torch.allclose(c.eta_on_grid(p, device="cpu"), torch.full((5,), 0.1))


# ---------------------------------------------------------------- normalisation / results
def test_normalize_by_fold_known_answers():
    # tests/test_kirchhoff_kernel.py:386-406
    img = torch.tensor([[[10.0, 20.0], [30.0, 0.0]]])
    fold = torch.tensor([[[2.0, 4.0], [3.0, 0.0]]])
    out = normalize_by_fold(img, fold, 1)
    assert out.tolist() == [[[5.0, 5.0], [10.0, 0.0]]]
    out = normalize_by_fold(torch.tensor([[[10.0, 20.0]]]), torch.tensor([[[1.0, 0.5]]]), 1)
    assert out.tolist() == [[[10.0, 20.0]]]


def test_migration_result_fold_normalization():
    r = MigrationResult(np.array([[[10.0, 20.0, 7.0]]]), np.array([[[2, 4, 1]]]))
    n = r.apply_fold_normalization(min_fold=2)
    assert n.image.tolist() == [[[5.0, 5.0, 0.0]]] and n.metadata["fold_normalized"]
    assert r.shape == (1, 1, 3) and r.get_statistics()["fold_max"] == 4


# ---------------------------------------------------------------- job config / data / processor params
def test_job_config_roundtrip(tmp_path):
    j = MigrationJobConfig(name="j", input_file=str(tmp_path / "in.zarr"), output_directory=str(tmp_path / "out"),
                           inline_max=199, inline_step=2, time_max_ms=2000, dt_ms=2.0, binning_preset="land_3d")
    g = j.get_output_grid()
    assert (g.n_inline, g.n_xline, g.n_time, g.dt) == (100, 100, 1001, 0.002)
    f = tmp_path / "job.json"
    j.save(str(f))
    j2 = MigrationJobConfig.load(str(f))
    assert j2.to_dict() == j.to_dict()
    errs = j.validate()
    assert any("does not exist" in e for e in errs)
    assert j.get_migration_config().weight_mode is WeightMode.SPREADING
    vf = tmp_path / "v.txt"
    np.savetxt(vf, np.array([[0.0, 1500.0], [2.0, 2700.0]]))
    j3 = MigrationJobConfig(name="v", input_file="x", output_directory=str(tmp_path), velocity_file=str(vf))
    assert j3.get_velocity_model().v0 == 1500.0



def test_job_config_binning_validation_and_templates(tmp_path):
    # the reference's job-config tests restated (tests/test_migration_job.py:
    # 95-410): validation messages, the gradient velocity, preset and explicit
    # binning tables, serialisation of a table, and the templates
    from seisproc_amd.models.binning import create_uniform_offset_binning
    from seisproc_amd.models import migration_job as mj
    inp = tmp_path / "in.sgy"
    inp.write_bytes(b"dummy")
    out = str(tmp_path / "out")
    ok = MigrationJobConfig(name="v", input_file=str(inp), output_directory=out, binning_preset="land_3d")
    assert ok.validate() == [] and ok.is_valid() and ok.velocity_v0 == 2500.0 and ok.created_at
    bad = lambda **kw: MigrationJobConfig(name="b", input_file=str(inp), output_directory=out, **kw).validate()
    assert any("binning" in e.lower() for e in bad())
    assert any("velocity" in e.lower() for e in bad(binning_preset="land_3d", velocity_v0=-100.0))
    assert any("time" in e.lower() for e in bad(binning_preset="land_3d", time_min_ms=5000.0, time_max_ms=2000.0))
    m = MigrationJobConfig(name="m", input_file=str(inp), output_directory=out, binning_preset="land_3d",
                           max_aperture_m=2500.0, max_angle_deg=55.0, velocity_v0=3000.0, velocity_gradient=0.5)
    assert (m.get_migration_config().max_aperture_m, m.get_migration_config().max_angle_deg) == (2500.0, 55.0)
    vm = m.get_velocity_model()
    assert vm.v0 == 3000.0 and vm.gradient == 0.5
    table = m.get_binning_table()
    assert len(table.bins) == 10 and table.bins[0].name == "land3d_0000_0200" and table.offset_range == (0, 5000)
    custom = create_uniform_offset_binning(0, 3000, 3)
    custom.name = "Custom 3 Bins"
    c = MigrationJobConfig(name="c", input_file=str(inp), output_directory=out, binning_table=custom)
    assert c.get_binning_table() is custom
    back = MigrationJobConfig.from_dict(json.loads(json.dumps(c.to_dict())))
    assert back.binning_table.name == "Custom 3 Bins" and len(back.binning_table.bins) == 3
    assert set(mj.list_templates()) == {"land_3d", "marine", "wide_azimuth_ovt", "full_stack"}
    land = mj.create_land_3d_template(str(inp), out)
    assert land.binning_preset == "land_3d" and land.max_aperture_m == 3000.0 and "Land 3D" in land.name
    marine = mj.create_from_template("marine", str(inp), out)
    assert marine.binning_preset == "marine" and marine.velocity_v0 == 1500.0 and marine.velocity_gradient == 0.3
    assert mj.create_from_template("land_3d", str(inp), out, velocity_v0=3500.0).velocity_v0 == 3500.0
    assert mj.create_full_stack_template(str(inp), out).create_stack_volume is True
    assert mj.create_wide_azimuth_template(str(inp), out).get_binning_table().n_bins == 16
    with pytest.raises(ValueError, match="Unknown template"):
        mj.create_from_template("nonexistent", str(inp), out)



def test_binning_table_checks():
    # the reference's checks restated (tests/test_binning.py:321-367), plus
    # an azimuth overlap that only part of the sectors share and exact gaps
    from seisproc_amd.models.binning import BinningTable, OffsetAzimuthBin as B
    t = BinningTable(name="t", bins=[B("a", 0.0, 500.0), B("b", 500.0, 1000.0)])
    c = t.check_coverage(np.array([100.0, 600.0, 1500.0, 300.0]), np.zeros(4))
    assert (c["n_traces"], c["n_unassigned"], c["n_multi_bin"], c["coverage_percent"]) == (4, 1, 0, 75.0)
    assert c["bin_counts"] == {"a": 2, "b": 1}
    assert t.check_overlaps() == []
    assert BinningTable(bins=[B("a", 0.0, 600.0), B("b", 400.0, 1000.0)]).check_overlaps() == [("a", "b")]
    sectors = BinningTable(bins=[B("n", 0.0, 100.0, 350.0, 20.0), B("e", 0.0, 100.0, 15.0, 90.0),
                                 B("s", 0.0, 100.0, 100.0, 200.0)])
    assert sectors.check_overlaps() == [("n", "e")]  # 15..20 degrees, across the wrap of "n"
    g = BinningTable(bins=[B("a", 0.0, 300.0), B("b", 500.0, 800.0), B("c", 700.0, 900.0), B("d", 950.0, 1000.0)])
    assert g.check_gaps(offset_step=50.0) == [(300.0, 500.0), (900.0, 950.0)]

def test_binning_factories_and_presets():
    # the reference's binning tests restated (tests/test_binning.py:440-540)
    from seisproc_amd.models.binning import (BinningPreset, create_full_stack_binning, create_narrow_azimuth_binning,
                                             create_ovt_binning)
    from seisproc_amd.utils.binning_presets import get_preset, get_preset_description, list_presets
    ovt = create_ovt_binning([(0, 500), (500, 1000)], n_azimuth_sectors=4)
    assert ovt.n_bins == 8 and ovt.preset is BinningPreset.OVT
    assert all(abs(b.azimuth_width - 90.0) < 0.1 for b in ovt)
    assert ovt.assign_trace(700.0, 95.0) == ["ovt_off0500_az090"]
    nar = create_narrow_azimuth_binning(0.0, 5000.0, inline_azimuth=0.0, azimuth_width=20.0)
    assert nar.n_bins == 4 and nar.preset is BinningPreset.NARROW_AZIMUTH
    assert {"inline_pos", "xline_pos"} <= {b.name for b in nar}
    # the inline sector wraps through north; a sector ending at north ends at 360
    assert nar.assign_trace(100.0, 355.0) == ["inline_pos"] and nar.assign_trace(100.0, 5.0) == ["inline_pos"]
    assert nar.assign_trace(100.0, 45.0) == []
    tilted = create_narrow_azimuth_binning(inline_azimuth=350.0, azimuth_width=20.0)
    assert tilted.get_bin("inline_pos").azimuth_max == 360.0 and tilted.assign_trace(10.0, 345.0) == ["inline_pos"]
    fs = create_full_stack_binning(offset_max=5000.0)
    assert fs.n_bins == 1 and fs.bins[0].offset_max == 5000.0 and fs.bins[0].is_full_azimuth
    assert {"land_3d", "marine", "wide_azimuth_ovt"} <= set(list_presets())
    assert get_preset("land_3d").preset is BinningPreset.COMMON_OFFSET and get_preset("land_3d").n_bins == 10
    assert get_preset("marine").n_bins == 6 and get_preset("marine").offset_range == (0, 8000)
    assert get_preset("wide_azimuth_ovt").preset is BinningPreset.OVT and get_preset("wide_azimuth_ovt").n_bins == 16
    assert get_preset("narrow_azimuth").n_bins == 4 and get_preset("full_stack").n_bins == 1
    with pytest.raises(ValueError):
        get_preset("nonexistent")
    assert "land" in get_preset_description("land_3d").lower()
    assert get_preset_description("nonexistent") == "No description available"


def test_geometry_queries_and_synthetic_layouts():
    # the reference's geometry tests restated (tests/test_migration_models.py:
    # 330-355) plus the selections' bounds and azimuth wrap
    from seisproc_amd.models.migration_geometry import (MigrationGeometry, create_land_3d_geometry,
                                                        create_synthetic_geometry)
    g = create_synthetic_geometry(n_shots=5, n_receivers_per_shot=24, receiver_spacing=25.0, near_offset=100.0)
    assert g.n_traces == 120 and g.offset[0] == pytest.approx(100.0)
    assert g.get_offset_range() == pytest.approx((100.0, 100.0 + 23 * 25.0))
    assert np.allclose(g.azimuth, 180.0)  # receivers trail south of shots on a north-going line
    st = g.get_statistics()
    assert (st["n_unique_sources"], st["n_unique_receivers"]) == (5, 5 * 24 - 4 * 20)
    assert {"x_range", "y_range", "offset_mean", "offset_std", "azimuth_range"} <= set(st)
    land = create_land_3d_geometry(n_source_lines=2, n_sources_per_line=3, n_receiver_lines=2,
                                   n_receivers_per_line=5)
    assert land.n_traces == 60 and land.get_source_position(5) == (0.0, 0.0) and land.get_source_position(10) == (50.0, 0.0)
    assert land.get_receiver_position(7) == (200.0, 50.0)
    # azimuths 0, 90, 180, 270, 350 degrees at offsets 100..500 m
    az = np.radians([0.0, 90.0, 180.0, 270.0, 350.0])
    off = np.array([100.0, 200.0, 300.0, 400.0, 500.0])
    z = np.zeros(5)
    q = MigrationGeometry(z, z, off * np.sin(az), off * np.cos(az))
    assert q.get_indices_for_bin(150.0, 400.0, 0.0, 360.0).tolist() == [1, 2, 3]  # inclusive
    assert q.filter_by_azimuth(340.0, 10.0).trace_indices.tolist() == [0, 4]      # wraps through north
    assert q.filter_by_offset_azimuth(0.0, 450.0, 80.0, 200.0).trace_indices.tolist() == [1, 2]
    assert q.get_cdp_position(1) == pytest.approx((100.0, 0.0))
    assert q.get_survey_bounds()["offset"] == pytest.approx((100.0, 500.0))


def test_base_migrator_params_and_serialisation():
    # a BaseMigrator keeps its keyword parameters (reference base_migrator.py:
    # 24-66, 230-243) without replacing params a processor base has set
    from seisproc_amd.processors.migration.base_migrator import BaseMigrator

    class Probe(BaseMigrator):
        def _validate_inputs(self): pass
        def migrate_gather(self, gather, geometry): pass
        def migrate_dataset(self, gathers, geometries, batch_size=1): pass
        def estimate_memory_gb(self, n_traces, n_samples): return 1.0
        def get_description(self): return "probe"

    cfg = MigrationConfig()
    m = Probe(VelocityModel(2000.0), cfg, chunk=7)
    d = m.to_dict()
    assert d["class_name"] == "Probe" and d["params"] == {"chunk": 7} and d["config"] == cfg.to_dict()
    assert d["velocity_model"]["data"] == 2000.0

    class WithParams(Probe):
        def __init__(self):
            self.params = {"kept": 1}
            super().__init__(VelocityModel(2000.0), cfg, extra=2)

    assert WithParams().params == {"kept": 1, "extra": 2}

def test_seismic_data_contract():
    d = SeismicData(traces=np.zeros((11, 3), np.int16), sample_rate=2.0)
    assert d.traces.dtype == np.float32 and d.n_samples == 11 and d.duration == 20.0 and d.nyquist_freq == 250.0
    with pytest.raises(ValueError):
        SeismicData(traces=np.zeros(5), sample_rate=2.0)
    assert "3 traces" in repr(d) and "11 samples" in repr(d)


def test_pstm_processor_params_roundtrip_without_gpu():
    from seisproc_amd.processors import BaseProcessor, PROCESSOR_REGISTRY
    cfg = MigrationConfig(output_grid=OutputGrid(n_time=11, n_inline=2, n_xline=3))
    proc = PROCESSOR_REGISTRY["PSTMProcessor"](config=cfg, velocity=2000.0, trace_dtype="bf16")
    d = json.loads(json.dumps(proc.to_dict()))
    again = BaseProcessor.from_dict(d)
    assert again.config.to_dict() == cfg.to_dict() and again.params["trace_dtype"] == "bf16"
    assert "HIP" in again.get_description()
    with pytest.raises(ValueError):
        PROCESSOR_REGISTRY["PSTMProcessor"](config=cfg, velocity=2000.0, trace_dtype="int8")
    with pytest.raises(ValueError):
        BaseProcessor.from_dict({"module": "seisproc_amd.models", "class_name": "SeismicData", "params": {}})


def test_optimized_migrator_factory_has_no_cpu_path():
    # create_optimized_kirchhoff_migrator (optimized_kirchhoff_migrator.py:721-775)
    # keeps the reference's keywords; without a ROCm GPU it raises instead of
    # falling back to a CPU migration
    import torch
    from seisproc_amd.models import MigrationConfig, VelocityModel
    from seisproc_amd.processors.migration import create_optimized_kirchhoff_migrator
    if torch.cuda.is_available():
        pytest.skip("checks the no-GPU behaviour")
    with pytest.raises(RuntimeError, match="GPU"):
        create_optimized_kirchhoff_migrator(velocity=VelocityModel(2000.0), config=MigrationConfig(),
                                            prefer_gpu=False, optimization_level="medium")


def test_vti_leg_range_bound_keeps_the_anelliptic_product_in_fp32():
    # The VTI kernels evaluate each leg as t = N rsq(N D), N = p^2 + 2 (1 + eta)
    # tqs p + tqs^2, D = tqs + (1 + 2 eta) p (samples: tqs = (tau / 2 dt)^2,
    # p = (h / (v dt))^2).  N D grows with p, tqs and eta, so its largest value
    # of a call is at the latest output time, a leg at the aperture, the
    # smallest velocity and the largest eta.  Every call the bound accepts
    # keeps that product below the FP32 maximum; calls near the limit that it
    # refuses would exceed it by a large factor once the leg times pass it.
    from seisproc_amd.abi_types import PstmParams
    from seisproc_amd.processors.migration.hip_ops import VTI_MAX_LEG2, vti_leg2_bound
    rng = np.random.default_rng(7)
    f32_max = float(np.finfo(np.float32).max)
    accepted = refused = 0
    for _ in range(20000):
        dt = 10 ** rng.uniform(-4.5, -1.5)
        n_tau = int(rng.integers(1, 4000))
        dtau = dt * 10 ** rng.uniform(-0.5, 2.5)
        tau0 = rng.uniform(0, 2.0)
        ap = 10 ** rng.uniform(1, 5)
        vmin = 10 ** rng.uniform(-1, 4)
        eta = rng.uniform(-0.49, 2.0)
        p = PstmParams.from_dict(dict(dt=dt, dtau=dtau, tau0=tau0, n_tau=n_tau, aperture_m=ap))
        tqs = ((tau0 + (n_tau - 1) * dtau) / (2 * dt)) ** 2
        pp = (ap / (vmin * dt)) ** 2
        nd = (pp * pp + 2 * (1 + eta) * tqs * pp + tqs * tqs) * (tqs + (1 + 2 * eta) * pp)
        if vti_leg2_bound(p, vmin, eta) < VTI_MAX_LEG2:
            accepted += 1
            assert nd < f32_max / 4, (dt, n_tau, dtau, tau0, ap, vmin, eta, nd)
        else:
            refused += 1
    assert accepted > 1000 and refused > 1000
    # the limit sits at 2^21-sample legs without anisotropy
    p = PstmParams.from_dict(dict(dt=1e-3, dtau=1e-3, tau0=0.0, n_tau=1, aperture_m=2 ** 21 * 1e-3 * 1.0))
    assert abs(vti_leg2_bound(p, 1.0, 0.0) - 2.0 ** 42) < 1e-6 * 2.0 ** 42
