"""CPU: the migration output writer (reference format: traces.zarr + fold.npy
+ Parquet headers + metadata.json, il-major traces) and its Zarr v2 store."""
import json

import numpy as np
import pandas as pd
import pytest
import torch

from seisproc_amd.utils.migration_output_storage import MigrationOutputStorage, create_migration_output_storage
from seisproc_amd.utils.zarr_v2 import ZarrV2Array


def test_zarr_v2_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((7, 23)).astype(np.float32)
    z = ZarrV2Array.create(tmp_path / "a.zarr", a.shape, (7, 10))
    z.write_columns(0, a)
    meta = json.loads((tmp_path / "a.zarr" / ".zarray").read_text())
    assert meta == {"zarr_format": 2, "shape": [7, 23], "chunks": [7, 10], "dtype": "<f4", "compressor": None,
                    "fill_value": 0.0, "order": "C", "filters": None}
    # chunk files named by chunk coordinates, stored at full chunk shape (edge padded with the fill value)
    names = sorted(f.name for f in (tmp_path / "a.zarr").iterdir() if not f.name.startswith("."))
    assert names == ["0.0", "0.1", "0.2"]
    last = np.frombuffer((tmp_path / "a.zarr" / "0.2").read_bytes(), dtype="<f4").reshape(7, 10)
    np.testing.assert_array_equal(last[:, :3], a[:, 20:])
    np.testing.assert_array_equal(last[:, 3:], 0.0)
    np.testing.assert_array_equal(ZarrV2Array.open(tmp_path / "a.zarr").read(), a)
    with pytest.raises(ValueError):
        z.write_columns(5, a[:, :10])


def _cube(n_time=11, n_il=5, n_xl=7, seed=1):
    rng = np.random.default_rng(seed)
    img = rng.standard_normal((n_time, n_il, n_xl)).astype(np.float32)
    fold = rng.integers(0, 4, (n_time, n_il, n_xl)).astype(np.float32)
    return img, fold


def test_full_cube_files_match_reference_layout(tmp_path):
    img, fold = _cube()
    n_time, n_il, n_xl = img.shape
    st = create_migration_output_storage(str(tmp_path), n_time, n_il, n_xl, dt_ms=2.0, d_inline=12.5, d_xline=25.0,
                                         inline_start=100, xline_start=7, x_origin=500000.0, y_origin=6.7e6,
                                         chunk_size=4)
    st.write_full_cube(img, fold)
    md = st.finalize(normalize=False, extra_metadata={"aperture_m": 1500.0})
    traces = ZarrV2Array.open(tmp_path / "traces.zarr").read()
    assert traces.shape == (n_time, n_il * n_xl)
    # il-major: trace = il * n_xl + xl
    np.testing.assert_array_equal(traces[:, 3 * n_xl + 2], img[:, 3, 2])
    np.testing.assert_array_equal(np.load(tmp_path / "fold.npy"), fold.reshape(n_time, -1))
    h = pd.read_parquet(tmp_path / "headers.parquet")
    assert list(h.columns) == ["trace_index", "inline", "xline", "cdp_x", "cdp_y", "ensemble_number",
                               "trace_in_ensemble"]
    r = h.iloc[3 * n_xl + 2]
    assert (r.inline, r.xline, r.trace_in_ensemble, r.ensemble_number) == (103, 9, 2, 103)
    # azimuth 0: inline axis points north (y), crossline east (x)
    assert r.cdp_x == pytest.approx(500000.0 + 2 * 25.0) and r.cdp_y == pytest.approx(6.7e6 + 3 * 12.5)
    e = pd.read_parquet(tmp_path / "ensemble_index.parquet")
    assert list(e.inline) == list(range(100, 100 + n_il)) and list(e.start_trace) == [i * n_xl for i in range(n_il)]
    assert list(e.end_trace) == [i * n_xl + n_xl - 1 for i in range(n_il)]
    ti = pd.read_parquet(tmp_path / "trace_index.parquet")
    assert list(ti.columns) == ["trace_index", "global_trace_id"] and len(ti) == n_il * n_xl
    on_disk = json.loads((tmp_path / "metadata.json").read_text())
    assert on_disk["shape"] == [n_time, n_il * n_xl] and on_disk["sample_rate"] == 2.0
    assert on_disk["migration_info"]["inline_end"] == 100 + n_il - 1
    assert on_disk["migration_params"] == {"aperture_m": 1500.0}
    assert on_disk["processing_stats"]["max_fold"] == float(fold.max()) == md["processing_stats"]["max_fold"]


def test_rotated_grid_coordinates(tmp_path):
    st = MigrationOutputStorage(str(tmp_path))
    st.initialize_output(3, 2, 2, d_inline=10.0, d_xline=20.0, grid_azimuth_deg=90.0)
    h = st._build_headers()
    # azimuth 90: inline axis points east, crossline south
    assert h.cdp_x.iloc[2] == pytest.approx(10.0) and h.cdp_y.iloc[2] == pytest.approx(0.0, abs=1e-9)
    assert h.cdp_x.iloc[1] == pytest.approx(0.0, abs=1e-9) and h.cdp_y.iloc[1] == pytest.approx(-20.0)


def test_normalize_inline_writes_and_device_layout(tmp_path):
    img, fold = _cube(seed=2)
    n_time, n_il, n_xl = img.shape
    a = create_migration_output_storage(str(tmp_path / "a"), n_time, n_il, n_xl, chunk_size=3)
    for i in range(n_il):
        a.write_inline(i, img[:, i, :], fold[:, i, :])
    a.finalize(normalize=True)
    got = ZarrV2Array.open(tmp_path / "a" / "traces.zarr").read()
    f2 = fold.reshape(n_time, -1)
    with np.errstate(divide="ignore", invalid="ignore"):
        want = np.where(f2 > 0, img.reshape(n_time, -1) / f2, 0.0).astype(np.float32)
    np.testing.assert_array_equal(got, want)
    # device layout [il][xl][tau] (a CPU tensor here) writes the same files
    b = create_migration_output_storage(str(tmp_path / "b"), n_time, n_il, n_xl, chunk_size=3)
    b.write_device_cube(torch.from_numpy(img).permute(1, 2, 0).contiguous(),
                        torch.from_numpy(fold).permute(1, 2, 0).contiguous(), traces_per_copy=4)
    b.finalize(normalize=True)
    np.testing.assert_array_equal(ZarrV2Array.open(tmp_path / "b" / "traces.zarr").read(), want)
    np.testing.assert_array_equal(np.load(tmp_path / "b" / "fold.npy"), f2)
    with pytest.raises(ValueError):
        b.write_full_cube(img[:, :2], fold[:, :2])
    with pytest.raises(RuntimeError):
        MigrationOutputStorage(str(tmp_path / "c")).write_inline(0, img[:, 0], fold[:, 0])
