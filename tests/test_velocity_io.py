"""CPU: velocity and anisotropy file readers/writers (the reference's text,
CDP-time-velocity, inline-xline-time-velocity, JSON and .npz layouts) and
their sampling onto the image grid."""
import json

import numpy as np
import pytest

from seisproc_amd.abi_types import PstmParams
from seisproc_amd.models.anisotropy_model import AnisotropyModel
from seisproc_amd.models.velocity_model import VelocityModel, VelocityType
from seisproc_amd.utils import anisotropy_io as aio
from seisproc_amd.utils import velocity_io as vio


def _grid(n_il=3, n_xl=2, n_tau=6, dtau=0.2):
    return PstmParams.from_dict(dict(n_traces=1, n_samples=n_tau, n_il=n_il, n_xl=n_xl, n_tau=n_tau, dt=dtau,
                                     dtau=dtau, il_dx=10.0, xl_dy=20.0))


def test_text_tv_units_comments_and_sorting(tmp_path):
    f = tmp_path / "v.txt"
    f.write_text("# header comment\n// another\n1000 6.0\n0 5.0\n\nbad line here\n500 5.5 extra\n")
    m = vio.read_velocity_text(f, time_unit="ms", velocity_unit="ft/ms")
    assert m.velocity_type is VelocityType.V_OF_Z
    np.testing.assert_allclose(m.z_axis, [0.0, 0.5, 1.0])
    np.testing.assert_allclose(m.data, np.float32([5.0, 5.5, 6.0]) * np.float32(304.8), rtol=1e-6)
    assert vio.detect_velocity_format(f) is vio.VelocityFileFormat.ASCII_TV
    with pytest.raises(ValueError):
        vio.read_velocity_text(f, time_unit="hours")
    (tmp_path / "empty.txt").write_text("# nothing\n")
    with pytest.raises(ValueError, match="No valid"):
        vio.read_velocity_text(tmp_path / "empty.txt")


def test_round_trips_text_json_npz(tmp_path):
    m = VelocityModel(data=np.array([1500.0, 2100.0, 2900.0], np.float32), z_axis=np.array([0.0, 1.0, 2.5]))
    for name in ("a.txt", "a.json", "a.npz"):
        vio.write_velocity_auto(m, tmp_path / name)
        back = vio.read_velocity_auto(tmp_path / name)
        np.testing.assert_allclose(back.z_axis, m.z_axis, atol=1e-6)
        np.testing.assert_allclose(back.data, m.data, rtol=1e-6)
    c = VelocityModel(2250.0)
    vio.write_velocity_text(c, tmp_path / "c.txt")
    cb = vio.read_velocity_auto(tmp_path / "c.txt")
    np.testing.assert_allclose(cb.rms_velocity_at_time(np.array([0.0, 3.0])), [2250.0, 2250.0])
    # the .npz metadata is a plain string: readable without pickle
    with np.load(tmp_path / "a.npz", allow_pickle=False) as z:
        assert json.loads(str(z["metadata"])) == {}


def test_cdp_tv_and_il_xl_tv_tables_on_the_image_grid(tmp_path):
    f = tmp_path / "cdp.txt"
    rows = [(c, t, 1500.0 + 2.0 * c + 0.5 * t) for c in (10, 12, 14) for t in (0.0, 500.0, 1000.0)]
    f.write_text("\n".join(f"{c} {t} {v}" for c, t, v in rows))
    assert vio.detect_velocity_format(f) is vio.VelocityFileFormat.ASCII_CDPTV
    m = vio.read_velocity_auto(f)
    assert m.velocity_type is VelocityType.V_OF_XZ and m.metadata["lateral_axes"] == "cdp"
    p = _grid(n_il=5, n_xl=1, n_tau=6, dtau=0.2)
    v = m.vrms_on_grid(p, device="cpu", inline_start=10).numpy()  # image column il <-> CDP 10 + il
    tau = 0.2 * np.arange(6)
    want = 1500.0 + 2.0 * np.clip(10 + np.arange(5), 10, 14)[:, None] + 0.5 * 1000.0 * tau[None, :]
    np.testing.assert_allclose(v[:, 0, :], want, rtol=1e-5)

    g = tmp_path / "ilxl.txt"
    g.write_text("\n".join(f"{il} {xl} {t} {1000 + 10 * il + xl + t}" for il in (1, 2, 3) for xl in (1, 2)
                           for t in (0.0, 1000.0)))
    assert vio.detect_velocity_format(g) is vio.VelocityFileFormat.ASCII_ILXLTV
    m3 = vio.read_velocity_auto(g)
    assert m3.velocity_type is VelocityType.V_OF_XYZ and m3.data.shape == (2, 3, 2)
    v3 = m3.vrms_on_grid(_grid(), device="cpu").numpy()
    il = np.arange(1, 4)[:, None, None]
    xl = np.arange(1, 3)[None, :, None]
    t = (0.2 * np.arange(6))[None, None, :]
    np.testing.assert_allclose(v3, 1000 + 10 * il + xl + 1000 * np.minimum(t, 1.0), rtol=1e-5)
    g.write_text("1 1 0 1500\n1 1 1000 1600\n2 2 0 1500\n")
    with pytest.raises(ValueError, match="lattice"):
        vio.read_velocity_ilxltv(g)


def test_anisotropy_readers(tmp_path):
    f = tmp_path / "a.txt"
    f.write_text("# z eps delta\n1000 0.2 0.05\n0 0.1 0.0\n")
    m = aio.read_anisotropy_text(f, z_unit="ms")
    np.testing.assert_allclose(m.z_axis, [0.0, 1.0])
    np.testing.assert_allclose(m.eta, [0.1, (0.2 - 0.05) / 1.1], rtol=1e-6)
    for name in ("b.txt", "b.json", "b.npz"):
        aio.write_anisotropy_auto(m, tmp_path / name)
        back = aio.read_anisotropy_auto(tmp_path / name)
        np.testing.assert_allclose(back.eta, m.eta, rtol=1e-5)
    c = AnisotropyModel(epsilon=0.12, delta=0.02)
    aio.write_anisotropy_npy(c, tmp_path / "c.npz")
    assert aio.read_anisotropy_npy(tmp_path / "c.npz").eta == pytest.approx(0.1 / 1.04)
    e = m.eta_on_grid(_grid(n_tau=6, dtau=0.2), device="cpu").numpy()
    np.testing.assert_allclose(e, np.interp(0.2 * np.arange(6), [0.0, 1.0], m.eta), rtol=1e-6)
