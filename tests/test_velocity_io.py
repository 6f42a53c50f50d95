"""CPU: velocity and anisotropy file readers/writers (the reference's text,
CDP-time-velocity, inline-xline-time-velocity, JSON and .npz layouts) and
their sampling onto the image grid."""
import json
import math

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


# ---------------------------------------------------------------- SEG-Y
def test_ibm_float_known_answers():
    # IBM System/360 hexadecimal floats: 0x42640000 = 100.0, 0xC276A000 =
    # -118.625, 0x41100000 = 1.0, 0x3F800000 = 0.03125 (= 0.5 x 16^-1)
    from seisproc_amd.utils.segy import float_to_ibm, ibm_to_float
    words = np.array([0x42640000, 0xC276A000, 0x41100000, 0x3F800000, 0], dtype=np.uint32)
    vals = np.array([100.0, -118.625, 1.0, 0.03125, 0.0])
    np.testing.assert_array_equal(ibm_to_float(words), vals)
    np.testing.assert_array_equal(float_to_ibm(vals), words)
    # round trip of velocities and of arbitrary floats to the 24-bit fraction
    v = np.arange(1400.0, 6000.0, 0.5)
    np.testing.assert_array_equal(ibm_to_float(float_to_ibm(v)), v)
    x = np.random.default_rng(0).normal(size=1000) * 1e3
    assert np.max(np.abs(ibm_to_float(float_to_ibm(x)) - x) / np.abs(x)) < 2 ** -20


def _cube(n_il=3, n_xl=4, nt=50):
    t = 0.004 * np.arange(nt)
    il, xl = np.meshgrid(np.arange(101, 101 + n_il), np.arange(7, 7 + n_xl), indexing="ij")
    v = 1500.0 + 800.0 * t[None, None, :] + 10.0 * il[..., None] + xl[..., None]
    return t, il.ravel(), xl.ravel(), v.reshape(n_il * n_xl, nt)


@pytest.mark.parametrize("fmt,little,n_ext", [(1, False, 0), (5, False, 1), (5, True, 0), (2, False, 0)])
def test_read_velocity_segy_cube(tmp_path, fmt, little, n_ext):
    from seisproc_amd.utils.segy import write_segy
    from seisproc_amd.utils.velocity_io import read_velocity_auto, read_velocity_segy
    t, il, xl, v = _cube()
    f = tmp_path / "vel.sgy"
    write_segy(f, v, 0.004, format_code=fmt, inline=il, crossline=xl, cdp=np.arange(1, il.size + 1),
               little_endian=little, n_extended_text=n_ext)
    m = read_velocity_segy(f)
    assert m.metadata["lateral_axes"] == "inline_xline" and m.data.shape == (50, 3, 4)
    np.testing.assert_allclose(m.z_axis, t, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(m.x_axis, [101, 102, 103])
    np.testing.assert_array_equal(m.y_axis, [7, 8, 9, 10])
    want = v.reshape(3, 4, 50).transpose(2, 0, 1)
    np.testing.assert_allclose(m.data, np.round(want) if fmt == 2 else want, rtol=0, atol=1e-3)
    assert read_velocity_auto(f).data.shape == (50, 3, 4)


def test_read_velocity_segy_cdp_line_and_single_trace(tmp_path):
    from seisproc_amd.utils.segy import write_segy
    from seisproc_amd.utils.velocity_io import read_velocity_segy
    t = 0.002 * np.arange(30)
    v = 1800.0 + 500.0 * t[None, :] + 20.0 * np.arange(5)[:, None]
    f = tmp_path / "line.segy"
    write_segy(f, v[::-1], 0.002, cdp=np.array([50, 40, 30, 20, 10]), delay_ms=0)
    m = read_velocity_segy(f)
    assert m.metadata["lateral_axes"] == "cdp" and m.metadata["cdp_range"] == (10, 50)
    np.testing.assert_array_equal(m.x_axis, [10, 20, 30, 40, 50])
    np.testing.assert_allclose(m.data, v.T, rtol=0, atol=1e-3)  # sorted by CDP
    one = tmp_path / "one.sgy"
    write_segy(one, v[:1], 0.002, delay_ms=8)
    m1 = read_velocity_segy(one)
    assert m1.data.ndim == 1 and abs(m1.z_axis[0] - 0.008) < 1e-12
    np.testing.assert_allclose(m1.data, v[0], rtol=0, atol=1e-3)



def test_segy_header_counts_unsigned_and_trace_header_fallback(tmp_path):
    # samples per trace and the interval are unsigned 16-bit fields: a
    # 40,000-sample record reads back whole; a binary header holding zero in
    # either field defers to the first trace header (bytes 115-118)
    from seisproc_amd.utils.segy import BIN_BYTES, TEXT_BYTES, read_segy, write_segy
    v = 1500.0 + np.arange(40000, dtype=np.float64)[None, :] * 0.25 + np.array([[0.0], [7.0]])
    f = tmp_path / "long.sgy"
    write_segy(f, v, 0.0005, format_code=5)
    vol = read_segy(f)
    assert vol.samples.shape == (2, 40000) and vol.interval_s == 0.0005
    np.testing.assert_array_equal(vol.samples, v.astype(np.float32))
    raw = bytearray(f.read_bytes())
    raw[TEXT_BYTES + 16:TEXT_BYTES + 18] = b"\0\0"  # interval
    raw[TEXT_BYTES + 20:TEXT_BYTES + 22] = b"\0\0"  # samples per trace
    g = tmp_path / "zero_bin.sgy"
    g.write_bytes(bytes(raw))
    vol0 = read_segy(g)
    assert vol0.samples.shape == (2, 40000) and vol0.interval_s == 0.0005
    np.testing.assert_array_equal(vol0.samples, vol.samples)
    raw[TEXT_BYTES + BIN_BYTES + 114:TEXT_BYTES + BIN_BYTES + 116] = b"\0\0"  # and trace 0's count
    g.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="sample count"):
        read_segy(g)

def test_preview_and_info_to_model(tmp_path):
    from seisproc_amd.utils.segy import write_segy
    from seisproc_amd.utils.velocity_io import preview_velocity_file, velocity_info_to_model
    t, il, xl, v = _cube()
    f = tmp_path / "vel.sgy"
    write_segy(f, v, 0.004, inline=il, crossline=xl)
    info = preview_velocity_file(f)
    assert info.is_valid and info.n_locations == 12 and info.n_time_samples == 50
    assert info.inline_range == (101, 103) and info.xline_range == (7, 10) and info.sample_interval == 0.004
    assert info.velocity_range[0] == pytest.approx(v.min(), abs=1e-3)
    assert velocity_info_to_model(info).data.shape == (50, 3, 4)
    txt = tmp_path / "tv.txt"
    txt.write_text("# t v\n0 1500\n1000 2500\n2000 3100\n")
    i2 = preview_velocity_file(txt)
    assert i2.is_valid and i2.raw_data["format"] == "tv" and i2.sample_interval == 1000.0
    m = velocity_info_to_model(i2, time_unit="ms", velocity_unit="km/s")
    np.testing.assert_allclose(m.z_axis, [0.0, 1.0, 2.0])
    np.testing.assert_allclose(m.data, [1.5e6, 2.5e6, 3.1e6])
    bad = preview_velocity_file(tmp_path / "missing.txt")
    assert not bad.is_valid and "not found" in bad.error_message
    with pytest.raises(ValueError, match="Invalid"):
        velocity_info_to_model(bad)



def test_preview_and_info_to_model_tables_json_npz(tmp_path):
    # the preview of each remaining format and the model built from it (reference
    # utils/velocity_io.py:1175-1290): CDP tables through their reader, inline-
    # crossline tables as v(t, inline) on the first inline's times, JSON and .npz
    # through theirs; a file of one numeric column is not a velocity table
    from seisproc_amd.utils.velocity_io import (VelocityFileFormat, preview_velocity_file, velocity_info_to_model,
                                                write_velocity_json, write_velocity_npy)
    cdp = tmp_path / "cdp.txt"
    cdp.write_text("".join(f"{c} {t} {1500 + c + t}\n" for c in (30, 10, 20) for t in (0, 500, 1000)))
    i1 = preview_velocity_file(cdp)
    assert i1.is_valid and i1.format is VelocityFileFormat.ASCII_CDPTV
    assert (i1.n_locations, i1.n_time_samples, i1.cdp_range) == (3, 3, (10, 30)) and i1.sample_interval == 500.0
    m1 = velocity_info_to_model(i1)
    np.testing.assert_allclose(m1.z_axis, [0.0, 0.5, 1.0])
    np.testing.assert_allclose(m1.data[:, 0], [1510, 2010, 2510])  # sorted by CDP

    ilxl = tmp_path / "ilxl.txt"
    rows = [(il, xl, t, 1500 + 10 * il + xl + t) for il in (5, 6) for xl in (1, 2) for t in (0, 1000, 2000)]
    ilxl.write_text("".join("%d %d %d %d\n" % r for r in rows))
    i2 = preview_velocity_file(ilxl)
    assert i2.is_valid and i2.format is VelocityFileFormat.ASCII_ILXLTV
    assert (i2.n_locations, i2.n_time_samples, i2.inline_range, i2.xline_range) == (4, 3, (5, 6), (1, 2))
    m2 = velocity_info_to_model(i2, velocity_unit="km/s")
    # the first inline's times, every crossline's included (the reference's axis)
    np.testing.assert_allclose(m2.z_axis, [0.0, 0.0, 1.0, 1.0, 2.0, 2.0])
    np.testing.assert_array_equal(m2.x_axis, [5, 6])
    # (np.interp over the repeated times picks one crossline's value)
    assert m2.data.shape == (6, 2) and 1000.0 * 1561 <= m2.data[0, 1] <= 1000.0 * 1562

    vm = VelocityModel(data=np.array([1500.0, 2000.0, 2600.0], np.float32), z_axis=np.array([0.0, 1.0, 2.0]))
    js, nz = tmp_path / "v.json", tmp_path / "v.npz"
    write_velocity_json(vm, js)
    write_velocity_npy(vm, nz)
    for f, fmt in ((js, "json"), (nz, "npz")):
        info = preview_velocity_file(f)
        assert info.is_valid and info.raw_data["format"] == fmt
        assert (info.n_locations, info.n_time_samples, info.time_range) == (1, 3, (0.0, 2.0))
        assert info.velocity_range == (1500.0, 2600.0) and info.sample_interval == 1.0
        np.testing.assert_allclose(velocity_info_to_model(info).data, vm.data)

    one = tmp_path / "one.txt"
    one.write_text("1500\n1600\n")
    i3 = preview_velocity_file(one)
    assert not i3.is_valid and i3.format is VelocityFileFormat.UNKNOWN


def test_convert_velocity_file_rms_interval_round_trip(tmp_path):
    # the reference's conversion test restated (tests/test_velocity_io.py:
    # 300-330), then the reverse conversion recovers the RMS function
    from seisproc_amd.utils.velocity_io import convert_velocity_file, read_velocity_auto
    rms = tmp_path / "rms.txt"
    rms.write_text("0\t2000\n500\t2200\n1000\t2500\n1500\t2800\n")
    itv = tmp_path / "interval.txt"
    m = convert_velocity_file(str(rms), str(itv), input_type="rms", output_type="interval", time_unit="ms")
    assert itv.exists() and m.metadata.get("velocity_type") == "interval" and abs(m.data[0] - 2000.0) < 1.0
    # Dix: the second interval carries sqrt((2200^2 0.5 - 2000^2 0) / 0.5) = 2200, the third
    # sqrt((2500^2 1.0 - 2200^2 0.5) / 0.5)
    assert abs(m.data[2] - math.sqrt((2500.0 ** 2 - 2200.0 ** 2 * 0.5) / 0.5)) < 0.05
    back = tmp_path / "back.json"
    r = convert_velocity_file(str(itv), str(back), input_type="interval", output_type="rms", time_unit="s")
    np.testing.assert_allclose(r.data, [2000.0, 2200.0, 2500.0, 2800.0], rtol=1e-5)
    assert read_velocity_auto(back).metadata["velocity_type"] == "rms"
    with pytest.raises(ValueError):
        convert_velocity_file(str(rms), str(itv), input_type="rms", output_type="average")


def test_custom_and_suggested_offset_binning():
    # the reference's checks (tests/test_binning.py:538-560) and the edges
    from seisproc_amd.utils.binning_presets import create_custom_offset_binning, suggest_binning
    t = create_custom_offset_binning(0.0, 2000.0, 4)
    assert t.n_bins == 4 and [b.offset_max for b in t] == [500.0, 1000.0, 1500.0, 2000.0]
    lg = create_custom_offset_binning(0.0, 1000.0, 3, logarithmic=True)
    np.testing.assert_allclose([b.offset_max for b in lg], [10.0, 100.0, 1000.0])
    assert lg.bins[0].offset_min == 0.0
    off = np.random.default_rng(3).uniform(0, 3000, 1000)
    s = suggest_binning(off, np.zeros(1000), target_traces_per_bin=200)
    assert s.n_bins == 5 and s.name == "Suggested - 5 Bins"
    counts = [int(b.contains_batch(off, np.zeros(1000)).sum()) for b in s]
    assert sum(counts) == 999 and max(counts) - min(counts) <= 2  # the largest offset sits on the last upper edge


def test_velocity_at_cdp_summary_and_well_anisotropy():
    # velocity functions by CDP and the summary (reference utils/velocity_io.py:
    # 1291-1368); anisotropy from a well (utils/anisotropy_io.py:438-486)
    from seisproc_amd.models.velocity_model import create_2d_velocity
    from seisproc_amd.utils.anisotropy_io import estimate_anisotropy_from_wells
    t, v = vio.get_velocity_at_cdp(VelocityModel(2500.0), 7)
    assert t.tolist() == [0.0, 6.0] and v.tolist() == [2500.0, 2500.0]
    one = VelocityModel(data=np.array([1500.0, 2500.0], np.float32), z_axis=np.array([0.0, 2.0]))
    assert vio.get_velocity_at_cdp(one, 3, times=np.array([1.0]))[1].tolist() == [2000.0]
    line = create_2d_velocity(np.array([[1500.0, 1700.0], [2500.0, 2900.0]]), [0.0, 2.0], [10.0, 20.0])
    t2, v2 = vio.get_velocity_at_cdp(line, 15)
    np.testing.assert_allclose(v2, [1600.0, 2700.0])
    s = vio.get_velocity_summary(line)
    assert "v_of_xz" in s and "1500..2900 m/s" in s and "10..20" in s
    assert "2500 m/s" in vio.get_velocity_summary(VelocityModel(2500.0))
    z = np.array([0.0, 1.0, 2.0])
    w = estimate_anisotropy_from_wells({"z": z, "epsilon": np.array([0.1, 0.2, 0.3]), "delta": np.array([0.0, 0.05, 0.1])})
    assert w.metadata["source"] == "well_estimate" and abs(w.get_parameters_at(0.5)[0] - 0.15) < 1e-12
    e = estimate_anisotropy_from_wells({"z": z, "vp": np.full(3, 3000.0)}, method="empirical")
    assert np.allclose(e.epsilon, 0.15) and np.allclose(e.delta, 0.08) and e.metadata["method"] == "empirical"
    with pytest.raises(ValueError):
        estimate_anisotropy_from_wells({"z": z})

def test_models_from_picks():
    from seisproc_amd.utils.anisotropy_io import create_anisotropy_from_picks
    from seisproc_amd.utils.velocity_io import create_velocity_from_picks
    m = create_velocity_from_picks([(1.0, 2500.0), (0.0, 1500.0), (2.0, 3000.0)], dt=0.5)
    np.testing.assert_allclose(m.z_axis, [0.0, 0.5, 1.0, 1.5, 2.0])
    np.testing.assert_allclose(m.data, [1500, 2000, 2500, 2750, 3000])
    assert m.metadata["n_picks"] == 3
    m2 = create_velocity_from_picks([(0.5, 2000.0)], dt=0.25, t_max=1.0)
    np.testing.assert_allclose(m2.data, [2000.0] * 5)
    with pytest.raises(ValueError):
        create_velocity_from_picks([])
    a = create_anisotropy_from_picks([(0.0, 0.1, 0.05), (1.0, 0.2, 0.05)], dz=0.5)
    np.testing.assert_allclose(a.epsilon, [0.1, 0.15, 0.2], rtol=1e-6)
    np.testing.assert_allclose(a.eta, (np.array([0.1, 0.15, 0.2]) - 0.05) / 1.1, rtol=1e-6)
