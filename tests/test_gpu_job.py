"""GPU: the bin-by-bin job driver on the HIP kernels -- pinned-host staging
with copies on a second stream, per-bin outputs, the stack against the FP64
oracle, and a resume after an injected stop that reproduces the
uninterrupted job bit for bit."""
import numpy as np
import pytest
import torch

import oracle
from seisproc_amd import synthetic as S
from seisproc_amd.models.binning import create_common_offset_binning
from seisproc_amd.processors.migration.job_driver import ArrayBinSource, MigrationJobRunner
from seisproc_amd.utils.zarr_v2 import ZarrV2Array

pytestmark = pytest.mark.gpu


def _job():
    s = S.tiny_survey(seed=50, n_bins=12, n_classes=3, n_samples=251, n_img=10, aperture=400.0)
    sx, sy, rx, ry = s.numpy_geometry()
    binning = create_common_offset_binning([(100.0, 200.0), (250.0, 350.0), (400.0, 500.0)])
    src = ArrayBinSource(s.traces.numpy(), sx, sy, rx, ry, binning)
    return s, src


def test_job_stack_matches_oracle_and_resume_is_bitwise(tmp_path):
    s, src = _job()
    p = s.params
    vel = s.vel.to("cuda")
    assert list(src.bin_summary().values()) == s.class_sizes
    full = MigrationJobRunner(p, vel, tmp_path / "full").run(src)
    assert full.completed == [b.name for b in src.binning.enabled_bins]
    ref, rfold = oracle.forward(p, s.traces.double().numpy(), *s.numpy_geometry(), s.vel.double().numpy())
    got = full.stack_image.double().cpu().numpy()
    assert np.abs(full.stack_fold.double().cpu().numpy() - rfold).sum() <= 1e-4 * rfold.sum() + 1
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-4
    for name, t in full.timings.items():
        print(f"[job] {name}: {t}")

    def stop_after_one(cur, total, msg):
        if cur == 1:
            raise KeyboardInterrupt("injected stop")

    with pytest.raises(KeyboardInterrupt):
        MigrationJobRunner(p, vel, tmp_path / "cut").run(src, progress_callback=stop_after_one)
    again = MigrationJobRunner(p, vel, tmp_path / "cut").run(src)
    assert again.skipped == full.completed[:1] and again.completed == full.completed[1:]
    assert torch.equal(again.stack_image, full.stack_image) and torch.equal(again.stack_fold, full.stack_fold)
    for name in full.completed:
        a = ZarrV2Array.open(tmp_path / "full" / name / "traces.zarr").read()
        b = ZarrV2Array.open(tmp_path / "cut" / name / "traces.zarr").read()
        np.testing.assert_array_equal(a, b)


def test_job_bf16_traces(tmp_path):
    s, src = _job()
    p = s.params
    res = MigrationJobRunner(p, s.vel.to("cuda"), tmp_path, trace_dtype="bf16", write_bins=False,
                             checkpoint=False).run(src)
    ref, _ = oracle.forward(p, s.traces.double().numpy(), *s.numpy_geometry(), s.vel.double().numpy())
    got = res.stack_image.double().cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"[job bf16] rel_rms vs FP64 oracle: {err:.3e}")
    assert err < 1e-3
