"""CPU: the bin-by-bin job driver (binning, per-bin outputs, stack,
checkpoint/resume after an injected failure).  The migration of one bin is
replaced by a deterministic CPU operator so that only the driver logic runs
here; tests/test_gpu_job.py runs the same job on the HIP kernels."""
import json

import numpy as np
import pytest
import torch

from seisproc_amd.abi_types import PstmParams
from seisproc_amd.models.binning import (BinningTable, OffsetAzimuthBin, create_common_offset_binning,
                                         create_full_stack_binning, create_uniform_offset_binning)
from seisproc_amd.processors.migration.job_driver import ArrayBinSource, MigrationJobRunner
from seisproc_amd.utils.zarr_v2 import ZarrV2Array


def _job(seed=0, n=90, ns=40, n_il=4, n_xl=3):
    rng = np.random.default_rng(seed)
    mx, my = rng.uniform(0, 100, n), rng.uniform(0, 100, n)
    off = rng.uniform(0, 600, n)
    az = rng.uniform(0, 2 * np.pi, n)
    sx, sy = mx - 0.5 * off * np.sin(az), my - 0.5 * off * np.cos(az)
    rx, ry = mx + 0.5 * off * np.sin(az), my + 0.5 * off * np.cos(az)
    traces = rng.standard_normal((n, ns)).astype(np.float32)
    p = PstmParams.from_dict(dict(n_traces=n, n_samples=ns, n_il=n_il, n_xl=n_xl, n_tau=ns, dt=0.004, dtau=0.004,
                                  il_dx=25.0, xl_dy=25.0, aperture_m=500.0))
    return p, traces, (sx, sy, rx, ry)


def _fake_migrate(p, traces, geo):
    """Deterministic stand-in for one bin's migration: image columns weight
    the traces by inverse midpoint distance."""
    il = torch.arange(p.n_il, dtype=torch.float64)[:, None].expand(p.n_il, p.n_xl).reshape(-1) * p.il_dx
    xl = torch.arange(p.n_xl, dtype=torch.float64)[None, :].expand(p.n_il, p.n_xl).reshape(-1) * p.xl_dy
    mx, my = 0.5 * (geo[:, 0] + geo[:, 2]), 0.5 * (geo[:, 1] + geo[:, 3])
    d2 = (il[:, None] - mx[None, :]) ** 2 + (xl[:, None] - my[None, :]) ** 2
    w = (1.0 / (1.0 + d2)).float() * (d2 < p.aperture_m ** 2)
    img = (w @ traces).reshape(p.n_il, p.n_xl, p.n_tau)
    fold = (w > 0).float().sum(1)[:, None].expand(-1, p.n_tau).reshape(p.n_il, p.n_xl, p.n_tau).contiguous()
    return img, fold


def test_binning_membership_and_presets():
    b = OffsetAzimuthBin("wrap", 100.0, 200.0, azimuth_min=350.0, azimuth_max=10.0)
    assert b.contains(150.0, 355.0) and b.contains(150.0, 5.0) and not b.contains(150.0, 20.0)
    assert not b.contains(200.0, 0.0) and b.contains(100.0, 0.0)  # [min, max)
    assert b.azimuth_width == pytest.approx(20.0) and b.azimuth_center == pytest.approx(0.0)
    with pytest.raises(ValueError):
        OffsetAzimuthBin("bad", 200.0, 100.0)
    t = create_uniform_offset_binning(0.0, 600.0, 3)
    assert [x.name for x in t] == ["offset_0000_0200", "offset_0200_0400", "offset_0400_0600"]
    assert BinningTable.from_json(t.to_json()).to_dict() == t.to_dict()
    assert create_full_stack_binning().n_enabled_bins == 1
    with pytest.raises(ValueError):
        BinningTable(bins=[OffsetAzimuthBin("a", 0, 1), OffsetAzimuthBin("a", 1, 2)])


def test_bin_source_assigns_by_offset_and_azimuth():
    p, tr, geo = _job()
    t = create_uniform_offset_binning(0.0, 600.0, 3)
    t.bins[1].enabled = False
    src = ArrayBinSource(tr, *geo, t)
    names = [n for n, _ in src.iterate_bins()]
    assert names == ["offset_0000_0200", "offset_0400_0600"]
    off = np.hypot(geo[2] - geo[0], geo[3] - geo[1])
    for name, idx in src.iterate_bins():
        b = t.get_bin(name)
        assert np.all((off[idx] >= b.offset_min) & (off[idx] < b.offset_max))
    np.testing.assert_array_equal(src.read(idx), tr[idx])
    # the BinnedDataset queries (reference utils/dataset_indexer.py:677-736)
    assert src.get_bin_trace_count("offset_0200_0400") == 0  # disabled: not assigned
    assert src.get_bin_summary() == {n: int(i.size) for n, i in src.iterate_bins()}
    rep = src.get_coverage_report()
    in_enabled = int(np.count_nonzero(((off >= 0) & (off < 200)) | ((off >= 400) & (off < 600))))
    assert rep["n_traces"] == len(off) and rep["total_assigned"] == in_enabled and rep["n_multi_bin"] == 0
    assert rep["n_unassigned"] == len(off) - in_enabled
    np.testing.assert_array_equal(src.get_bin_trace_numbers("offset_0400_0600"), idx)


def _read_bin(d, name):
    return ZarrV2Array.open(d / name / "traces.zarr").read(), np.load(d / name / "fold.npy")


def test_resume_after_injected_failure_is_bitwise_identical(tmp_path):
    p, tr, geo = _job(seed=3)
    t = create_uniform_offset_binning(0.0, 600.0, 4)
    vel = torch.full((p.n_tau,), 2000.0)
    mk = lambda d: MigrationJobRunner(p, vel, d, device=torch.device("cpu"), migrate_fn=_fake_migrate)
    full = mk(tmp_path / "full").run(ArrayBinSource(tr, *geo, t))
    assert len(full.completed) == 4

    calls = []

    def fail_after_two(cur, total, msg):
        calls.append(cur)
        if cur == 2:
            raise KeyboardInterrupt("injected stop after bin 2")

    with pytest.raises(KeyboardInterrupt):
        mk(tmp_path / "cut").run(ArrayBinSource(tr, *geo, t), progress_callback=fail_after_two)
    st = json.loads((tmp_path / "cut" / "checkpoint" / "state.json").read_text())
    assert st["completed"] == full.completed[:2]
    again = mk(tmp_path / "cut").run(ArrayBinSource(tr, *geo, t))
    assert again.skipped == full.completed[:2] and again.completed == full.completed[2:]
    assert torch.equal(again.stack_image, full.stack_image) and torch.equal(again.stack_fold, full.stack_fold)
    for name in full.completed:
        a, fa = _read_bin(tmp_path / "full", name)
        b, fb = _read_bin(tmp_path / "cut", name)
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(fa, fb)
    # the stack is the sum of the bins' raw images
    raw = sum(_fake_migrate(p.replace(n_traces=len(i)), torch.from_numpy(tr[i]),
                            torch.from_numpy(np.stack([g[i] for g in geo], 1)))[0]
              for _, i in ArrayBinSource(tr, *geo, t).iterate_bins())
    torch.testing.assert_close(full.stack_image, raw, rtol=1e-6, atol=1e-6)


def test_checkpoint_of_another_job_is_refused(tmp_path):
    p, tr, geo = _job(seed=4)
    t = create_common_offset_binning([(0.0, 300.0), (300.0, 600.0)])
    vel = torch.full((p.n_tau,), 2000.0)
    MigrationJobRunner(p, vel, tmp_path, device=torch.device("cpu"), migrate_fn=_fake_migrate) \
        .run(ArrayBinSource(tr, *geo, t))
    other = MigrationJobRunner(p.replace(aperture_m=400.0), vel, tmp_path, device=torch.device("cpu"),
                               migrate_fn=_fake_migrate)
    with pytest.raises(RuntimeError, match="different job"):
        other.run(ArrayBinSource(tr, *geo, t))
    # an explicit restart ignores the old state
    res = other.run(ArrayBinSource(tr, *geo, t), resume=False)
    assert len(res.completed) == 2 and not res.skipped
    # a checkpoint kept without a running stack cannot resume a job that keeps one
    nostack = tmp_path / "nostack"
    MigrationJobRunner(p, vel, nostack, device=torch.device("cpu"), migrate_fn=_fake_migrate, stack=False) \
        .run(ArrayBinSource(tr, *geo, t))
    with pytest.raises(RuntimeError, match="different job"):
        MigrationJobRunner(p, vel, nostack, device=torch.device("cpu"), migrate_fn=_fake_migrate) \
            .run(ArrayBinSource(tr, *geo, t))


def test_runner_needs_the_gpu_without_an_injected_migration(tmp_path):
    p, tr, geo = _job()
    with pytest.raises(RuntimeError, match="ROCm"):
        MigrationJobRunner(p, torch.full((p.n_tau,), 2000.0), tmp_path, device=torch.device("cpu"))


def test_stop_inside_checkpoint_write_keeps_stack_and_state_consistent(tmp_path, monkeypatch):
    # a stop between writing a generation's stack file and replacing state.json
    # must leave the previous generation intact: resume adds each bin once
    import os
    from seisproc_amd.processors.migration import job_driver as jd
    p, tr, geo = _job(seed=5)
    t = create_uniform_offset_binning(0.0, 600.0, 4)
    vel = torch.full((p.n_tau,), 2000.0)
    mk = lambda d: MigrationJobRunner(p, vel, d, device=torch.device("cpu"), migrate_fn=_fake_migrate)
    full = mk(tmp_path / "full").run(ArrayBinSource(tr, *geo, t))

    real_replace = os.replace
    n = {"calls": 0}

    def flaky_replace(a, b):
        # stop at the 3rd checkpoint: its stack file is written, state.json not yet replaced
        if str(b).endswith("state.json"):
            n["calls"] += 1
            if n["calls"] == 3:
                assert (tmp_path / "cut" / "checkpoint" / "stack.3.npz").exists()
                raise OSError("injected stop before state.json is replaced")
        return real_replace(a, b)

    monkeypatch.setattr(jd.os, "replace", flaky_replace)
    with pytest.raises(OSError, match="injected"):
        mk(tmp_path / "cut").run(ArrayBinSource(tr, *geo, t))
    monkeypatch.setattr(jd.os, "replace", real_replace)
    st = json.loads((tmp_path / "cut" / "checkpoint" / "state.json").read_text())
    assert st["completed"] == full.completed[:2] and st["stack_file"] == "stack.2.npz"
    again = mk(tmp_path / "cut").run(ArrayBinSource(tr, *geo, t))
    assert again.skipped == full.completed[:2]
    assert torch.equal(again.stack_image, full.stack_image) and torch.equal(again.stack_fold, full.stack_fold)
    assert sorted(f.name for f in (tmp_path / "cut" / "checkpoint").glob("stack.*.npz")) == ["stack.4.npz"]


def test_checkpoint_of_other_input_data_is_refused(tmp_path):
    # same parameters, binning and per-bin counts, different geometry or traces
    p, tr, geo = _job(seed=6)
    t = create_uniform_offset_binning(0.0, 600.0, 2)
    vel = torch.full((p.n_tau,), 2000.0)
    mk = lambda: MigrationJobRunner(p, vel, tmp_path, device=torch.device("cpu"), migrate_fn=_fake_migrate)
    mk().run(ArrayBinSource(tr, *geo, t))
    moved = (geo[0] + 1.0, geo[1], geo[2] + 1.0, geo[3])
    with pytest.raises(RuntimeError, match="different job"):
        mk().run(ArrayBinSource(tr, *moved, t))
    other = tr.copy()
    other[0, 0] += 1.0
    with pytest.raises(RuntimeError, match="different job"):
        mk().run(ArrayBinSource(other, *geo, t))
    res = mk().run(ArrayBinSource(tr, *geo, t))
    assert not res.completed and len(res.skipped) == 2
