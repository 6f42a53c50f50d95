"""Parity against golden vectors of the reference's own function.

``tests/golden/ref_migrate_kirchhoff_full.npz`` holds inputs and outputs of
``migrate_kirchhoff_full`` (processors/migration/kirchhoff_kernel.py:459-854)
run on the CPU in FP32 by ``tests/golden/make_reference_vectors.py``: a 3D grid,
a wider grid spanning three of the reference's 100-point tiles with the
time-dependent aperture, a 2D line whose record ends before the deep/far
traveltimes (clamp to n - 1.001), and a record starting at t_min > 0 with
output times off the input sample grid.

The FP64 oracle (CPU) and the HIP library (GPU) are compared with the
reference's image and fold: fold equal except aperture/angle boundary pairs
where FP32 and FP64 predicates disagree (counted, at most 1e-4 of the
contributions), image within 1e-4 relative RMS.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from seisproc_amd.abi_types import FLAG_TIME_APERTURE, PstmParams

GOLDEN = Path(__file__).resolve().parent / "golden" / "ref_migrate_kirchhoff_full.npz"


def _cases():
    z = np.load(GOLDEN, allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    out = []
    for m in meta:
        n = m["name"]
        arr = {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(n + "/")}
        out.append((m, arr))
    return out


CASES = _cases()
IDS = [m["name"] for m, _ in CASES]


def _params(m, arr) -> PstmParams:
    n_samples, n_traces = arr["traces"].shape
    return PstmParams.from_dict(dict(
        n_traces=n_traces, n_samples=n_samples, n_il=m["n_il"], n_xl=m["n_xl"], n_tau=m["n_tau"],
        dt=m["dt_ms"] / 1000.0, t0=m["t_min_ms"] / 1000.0, dtau=m["dtau_ms"] / 1000.0, tau0=m["tau0_ms"] / 1000.0,
        origin_x=m["x0"], origin_y=m["y0"], il_dx=m["d_il"], il_dy=0.0, xl_dx=0.0, xl_dy=m["d_xl"],
        aperture_m=m["aperture"], max_angle_deg=m["angle"],
        flags=FLAG_TIME_APERTURE if m["time_aperture"] else 0))


def _compare(img, fold, arr, img_tol=1e-4):
    """img, fold: [il][xl][tau]; the reference arrays are [tau][il][xl]."""
    ref = arr["image"].astype(np.float64).transpose(1, 2, 0)
    rfold = arr["fold"].astype(np.float64).transpose(1, 2, 0)
    flips = float(np.abs(fold - rfold).sum())
    assert rfold.sum() > 0
    assert flips <= 1e-4 * rfold.sum(), (flips, rfold.sum())
    # image samples whose fold agrees hold exactly the same set of contributions
    same = fold == rfold
    err = float(np.linalg.norm((img - ref)[same]) / np.linalg.norm(ref[same]))
    assert err < img_tol, err
    err_all = float(np.linalg.norm(img - ref) / np.linalg.norm(ref))
    assert err_all < 1e-4, err_all
    return flips, err


def test_golden_file_is_consistent():
    assert len(CASES) == 4
    for m, arr in CASES:
        n_samples, n_traces = arr["traces"].shape
        assert arr["image"].shape == (m["n_tau"], m["n_il"], m["n_xl"])
        assert arr["fold"].shape == arr["image"].shape
        assert arr["sx"].shape == (n_traces,)
        assert arr["output_x"].shape == (m["n_il"] * m["n_xl"],)
        # the grid the reference saw is the one the params describe
        il = np.arange(m["n_il"])[:, None]
        xl = np.arange(m["n_xl"])[None, :]
        np.testing.assert_allclose(arr["output_x"], (m["x0"] + il * m["d_il"] + 0 * xl).ravel(), atol=1e-4)
        np.testing.assert_allclose(arr["output_y"], (m["y0"] + 0 * il + xl * m["d_xl"]).ravel(), atol=1e-4)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_oracle_matches_reference_function(case, oracle_lib):
    import oracle
    m, arr = case
    p = _params(m, arr)
    geo = [arr[k].astype(np.float64) for k in ("sx", "sy", "rx", "ry")]
    vel = np.full(p.n_tau, float(m["v"]))
    img, fold = oracle.forward(p, arr["traces"].T.astype(np.float64), *geo, vel)
    # FP64 restatement vs the reference's FP32 arithmetic
    _compare(img, fold, arr, img_tol=5e-6)


def test_clamp_case_reads_past_the_record():
    # the line case is only a clamp test if some traveltimes pass the record end
    m, arr = next(c for c in CASES if c[0]["name"] == "line2d_clamp")
    n_samples = arr["traces"].shape[0]
    t_end = (n_samples - 1) * m["dt_ms"] / 1000.0
    tau_max = (m["tau0_ms"] + (m["n_tau"] - 1) * m["dtau_ms"]) / 1000.0
    assert tau_max > t_end
    rfold = arr["fold"]
    k_end = int(np.ceil((t_end * 1000.0 - m["tau0_ms"]) / m["dtau_ms"]))
    assert rfold[k_end + 1:].sum() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_hip_forward_matches_reference_function(case):
    import torch
    from seisproc_amd.processors.migration.hip_ops import PstmPlan
    m, arr = case
    p = _params(m, arr)
    dev = torch.device("cuda:0")
    src = torch.tensor(np.stack([arr["sx"], arr["sy"]], 1), dtype=torch.float64, device=dev)
    rcv = torch.tensor(np.stack([arr["rx"], arr["ry"]], 1), dtype=torch.float64, device=dev)
    traces = torch.tensor(np.ascontiguousarray(arr["traces"].T), dtype=torch.float32, device=dev)
    vel = torch.full((p.n_tau,), float(m["v"]), dtype=torch.float32, device=dev)
    print(f"golden case {m['name']}: {p.n_traces} traces, image {p.n_il}x{p.n_xl}x{p.n_tau}", flush=True)
    plan = PstmPlan(p, src, rcv, device=dev)
    img, fold = plan.forward(traces, vel)
    torch.cuda.synchronize()
    _compare(img.double().cpu().numpy(), fold.double().cpu().numpy(), arr, img_tol=1e-4)


@pytest.mark.parametrize("mistake", ["weight_without_obliquity", "index_off_by_one", "velocity_1pct",
                                     "aperture_10pct"])
def test_golden_vectors_catch_planted_mistakes(mistake, oracle_lib):
    # each plausible slip in the oracle (a dropped weight term, an index shift,
    # a wrong velocity, a wrong aperture) must fail the comparison
    import oracle
    from seisproc_amd.abi_types import WEIGHT_SPREADING
    m, arr = CASES[0]
    p = _params(m, arr)
    v = float(m["v"])
    if mistake == "weight_without_obliquity":
        p = p.replace(weight_mode=WEIGHT_SPREADING)
    elif mistake == "index_off_by_one":
        p = p.replace(t0=p.t0 - p.dt)
    elif mistake == "velocity_1pct":
        v *= 1.01
    else:
        p = p.replace(aperture_m=p.aperture_m * 1.1)
    geo = [arr[k].astype(np.float64) for k in ("sx", "sy", "rx", "ry")]
    img, fold = oracle.forward(p, arr["traces"].T.astype(np.float64), *geo, np.full(p.n_tau, v))
    with pytest.raises(AssertionError):
        _compare(img, fold, arr, img_tol=5e-6)
