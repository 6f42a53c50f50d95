"""CPU: the multi-volume output manager (reference seisio/migration_output.py),
with the reference's tests (tests/test_migration_output.py) restated and the
three file formats read back."""
import numpy as np
import pytest
import torch

from seisproc_amd.models.binning import create_uniform_offset_binning
from seisproc_amd.models.migration_config import OutputGrid
from seisproc_amd.seisio import MigrationOutputManager, OutputVolumeInfo, create_output_manager


def _grid():
    return OutputGrid(n_time=100, n_inline=20, n_xline=20, dt=0.004, d_inline=25.0, d_xline=25.0,
                      inline_start=100, xline_start=200)


def _manager(tmp_path, **kw):
    m = MigrationOutputManager(output_directory=str(tmp_path), output_grid=_grid(),
                               binning_table=create_uniform_offset_binning(0, 2000, 4), **kw)
    return m


def test_volumes_accumulation_and_normalisation(tmp_path):
    m = _manager(tmp_path)
    with pytest.raises(RuntimeError, match="not initialized"):
        m.add_migrated_data("offset_0000_0500", np.ones((100, 20, 20), np.float32))
    m.initialize()
    assert len(m._volumes) == 4 and m._stack_volume is not None and m._initialized
    assert all(v.shape == (100, 20, 20) for v in m._volumes.values())
    name = m.binning_table.bins[0].name
    m.add_migrated_data(name, np.full((100, 20, 20), 10.0, np.float32))
    assert np.allclose(m._volumes[name], 10.0) and np.allclose(m._stack_volume, 10.0)
    with pytest.raises(ValueError, match="Unknown bin"):
        m.add_migrated_data("nonexistent_bin", np.ones((100, 20, 20), np.float32))
    assert m.get_volume(name).shape == (100, 20, 20) and m.get_volume("stack") is not None
    s = m.get_summary()
    assert s["n_bins"] == 4 and s["create_stack"] is True and s["volumes"][name]["n_traces_migrated"] == 1
    assert not m._volume_info[name].is_complete
    m.finalize_bin(name)
    assert m._volume_info[name].is_complete

    f = _manager(tmp_path / "fold", create_fold=True)
    f.initialize()
    for _ in range(5):
        f.accumulate_to_bin(name, 50, 10, 10, 2.0)
    f.accumulate_to_bin(name, 50, 11, 11, 4.0)
    assert f._volumes[name][50, 10, 10] == 10.0 and f._fold_volumes[name][50, 10, 10] == 5
    f.normalize_by_fold(min_fold=2)
    assert f._volumes[name][50, 10, 10] == 2.0 and f._volumes[name][50, 11, 11] == 0.0
    assert f._stack_volume[50, 10, 10] == 2.0

    n = _manager(tmp_path / "nostack", create_stack=False)
    n.initialize()
    assert n._stack_volume is None and "stack" not in n._volume_info
    info = OutputVolumeInfo(name="b", filepath="/o/b.npy", bin_name="b", offset_center=500.0, azimuth_center=45.0)
    assert info.n_traces_migrated == 0 and not info.is_complete


def test_headers_and_device_volumes(tmp_path):
    m = _manager(tmp_path, create_fold=True)
    m.initialize()
    name = m.binning_table.bins[1].name
    h = m.get_output_headers(name, 5, 10)
    assert (h["INLINE_3D"], h["CROSSLINE_3D"]) == (105, 210) and h["OFFSET"] == 750.0
    assert (h["CDP_X"], h["CDP_Y"]) == pytest.approx(_grid().get_coordinates(5, 10))
    img = torch.arange(20 * 20 * 100, dtype=torch.float32).reshape(20, 20, 100)  # [il][xl][tau]
    m.add_device_volume(name, img, torch.ones_like(img))
    np.testing.assert_array_equal(m.get_volume(name), img.permute(2, 0, 1).numpy())
    assert m.get_fold(name).dtype == np.int32 and int(m.get_fold("stack").sum()) == 20 * 20 * 100


@pytest.mark.parametrize("fmt", ["npy", "zarr", "segy"])
def test_save_all_formats_read_back(tmp_path, fmt):
    from seisproc_amd.utils.segy import read_segy
    from seisproc_amd.utils.zarr_v2 import ZarrV2Array
    m = _manager(tmp_path, output_format=fmt, create_fold=True)
    m.initialize()
    name = m.binning_table.bins[0].name
    rng = np.random.default_rng(1)
    data = rng.normal(size=(100, 20, 20)).astype(np.float32)
    m.add_migrated_data(name, data, fold=np.full((100, 20, 20), 3.0))
    m.save_all()
    for info in m._volume_info.values():
        assert (tmp_path / info.filepath).exists()
    path = m._volume_info[name].filepath
    fold_path = m._fold_path(path)
    if fmt == "npy":
        back, fold = np.load(path), np.load(fold_path)
    elif fmt == "zarr":
        back, fold = ZarrV2Array.open(path).read(), ZarrV2Array.open(fold_path).read()
    else:
        vol = read_segy(path)
        back = vol.samples.T.reshape(100, 20, 20)
        assert vol.headers["inline"][20] == 101 and vol.headers["crossline"][1] == 201
        fold = read_segy(fold_path).samples.T.reshape(100, 20, 20)
    np.testing.assert_array_equal(back, data)
    np.testing.assert_array_equal(fold, 3)


def test_from_job_config(tmp_path):
    from seisproc_amd.models.migration_job import MigrationJobConfig
    j = MigrationJobConfig(name="My Job", input_file="x", output_directory=str(tmp_path), binning_preset="marine",
                           output_format="zarr", create_fold_volume=True, time_max_ms=400.0, dt_ms=4.0,
                           inline_max=10, xline_max=8)
    m = create_output_manager(j)
    m.initialize()
    assert m.base_name == "my_job" and len(m._volumes) == 6 and m.create_fold
    assert m._volume_info["stack"].filepath.endswith("my_job_stack.zarr")
    with pytest.raises(ValueError, match="Unknown output format"):
        MigrationOutputManager(str(tmp_path), _grid(), create_uniform_offset_binning(0, 2000, 4), output_format="tiff")
