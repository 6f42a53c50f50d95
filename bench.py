"""PSTM throughput benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload pstm3d-precision|pstm3d-single-offset|pstm3d-vti-largest|lsm3d]

Default workload (the metric's precision form): 3D PSTM of 25 offset classes
x 40,000 traces (1,000,000 traces x 1001 samples at 2 ms) into a 400 x 400 x
1001 image at 12.5 m, aperture 1500 m / 60 deg, compensated accumulation,
traces stored BF16 with per-trace FP32 scales.  A step is one forward
migration of every trace into the full image (fold included) with the inputs
resident in HBM.  ``value`` is the BF16-trace rate in Gcontrib/s (one
contribution = one (trace, image sample) pair that passes the aperture and
angle tests = one unit of fold); the FP32 and FP16 forms are timed beside it
and every form's error against the FP64 oracle is measured on sampled image
columns.

The CPU baseline is the torch-CPU restatement of the reference's CPU method
(oracle/ref_torch.py), timed on a bounded sample of the same workload: one
100-point reference tile (10 x 10 image columns) at every 16th output time,
with every trace that reaches it.

``roofline``: the tile kernel reads its taps from trace windows staged in LDS
and is bound by VALU issue, so the primary figure is VALU lane operations per
second (counter totals of one call x this run's kernel time) against the FP32
vector peak; the HBM rate of the algorithmic bytes, the measured HBM traffic
and the LDS traffic are reported beside it.  The counter totals come from
the profiles/*.md summary whose source digest matches the kernel sources
and flags of the library being run (``source_sha()``); without one the
counter figures are null.
``end_to_end`` times one engine-style call from host memory (copy, plan,
quantisation, forward), outside the timed steps.  ``--vt`` runs the
precision form with v_rms = 1500 + 600 t.

``--workload lsm3d`` measures least-squares migration instead: a step is one
CG iteration (one adjoint + one forward launch over 1,048,576 traces x 751
samples and a 256 x 256 x 751 image); the rate counts both launches'
contributions, the CPU baseline times the restated forward and its transpose
on the same kind of sample, and the image-power velocity gradient on a 32 x
32 x 64 B-spline lattice is timed and checked by central differences.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Published MI355X FP32 vector peak, 157.3 TFLOP/s, in VALU lane-operations
# per second (one FMA = 2 FLOP): 78.65 T lane-op/s, i.e. one wave64 FP32
# instruction per 2 cycles per SIMD at 2.4 GHz on 1024 SIMDs.
VALU_PEAK_TLANE = 157.3 / 2.0
# That figure counts packed FP32 (two values per lane per instruction), which
# the kernels here never emit (packed FP32 is switched off at compile time).
# Unpacked FP32 VALU instructions issue at 16 lanes per clock per SIMD -- the
# lane rate of the published FP64 vector peak, 78.6 TFLOP/s -- so the VALU
# ceiling of these kernels is half the published FP32 figure: 39.32 T
# lane-op/s (1024 SIMDs x 16 lanes x 2.4 GHz).  The issue-rate
# probe (seisproc_amd/csrc/valu_rates.hip) measures v_fma_f32 streams there.
VALU_UNPACKED_TLANE = 1024 * 16 * 2.4e9 / 1e12
# LDS: 256 CUs x 128 bytes per clock x 2.4 GHz.
LDS_PEAK_TBS = 256 * 128 * 2.4e9 / 1e12
N_SIMD = 1024
N_SE = 32  # shader engines summed by the SQ counters (8 XCDs x 4)
PROFILES = ROOT / "profiles"


def source_sha() -> str:
    """Digest of the kernel sources and compiler flags the library is built
    from: a counter summary in profiles/ is used only for the build it was
    measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in sorted((ROOT / "seisproc_amd" / "csrc").iterdir()):
        if f.suffix in (".h", ".hip"):
            h.update(f.name.encode())
            h.update(f.read_bytes())
    from seisproc_amd.build import HIPCC_FLAGS
    h.update(" ".join(HIPCC_FLAGS).encode())
    return h.hexdigest()[:16]


def load_counters(workload: str, form: str, classes, any_classes: bool = False):
    """Per-call counter totals of this workload and trace form from the JSON
    block of a profiles/*.md summary stamped with the current source digest.
    any_classes: also accept a summary of the same workload on another number
    of offset classes (the caller scales its totals by contributions; the
    record keeps its own "classes").  Returns (record, path) or (None, reason)."""
    sha = source_sha()
    found = []
    for f in sorted(PROFILES.glob("*.md")):
        text = f.read_text()
        for block in text.split("```json")[1:]:
            try:
                rec = json.loads(block.split("```")[0])
            except ValueError:
                continue
            if rec.get("kind") != "pstm-counters":
                continue
            if rec.get("workload") == workload and rec.get("form") == form and (
                    any_classes or rec.get("classes") == classes):
                found.append((rec, f))
    found.sort(key=lambda rf: rf[0].get("classes") != classes)  # the same class count first
    for rec, f in found:
        if rec.get("source_sha") == sha:
            return rec, str(f.relative_to(ROOT))
    if found:
        return None, f"counter summaries for this workload are for other builds (source {sha} not among " \
                     f"{[r.get('source_sha') for r, _ in found]})"
    return None, "no counter summary for this workload in profiles/"


def roofline(contrib: float, kernel_s: float, bytes_per_contribution: int, cnt):
    """Bound of the dominant kernel.  The tile kernel stages trace windows in
    LDS and is limited by VALU issue, so the primary figure is VALU lane
    operations per second against the FP32 vector peak; the HBM rate of the
    algorithmic bytes (every tap from HBM) and the measured HBM and LDS
    traffic are reported beside it.  Counter figures are per call, from the
    profile summary of the current sources; rates use this run's kernel time."""
    algo = contrib * bytes_per_contribution / kernel_s / 1e9
    r = {"bound": "valu", "achieved": None, "peak": round(VALU_PEAK_TLANE, 2),
         "unit": "T lane-op/s (FP32 VALU; 157.3 TFLOP/s = 78.65 T lane-FMA/s)", "frac": None, "traffic": None,
         "hbm_algorithmic": {"achieved": round(algo, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(algo / HBM_PEAK_GBS, 4), "bytes_per_contribution": bytes_per_contribution}}
    rec, src = cnt
    if rec is None:
        r["counters"] = {"source": None, "note": src}
        return r
    c = rec["counters"]
    scale = contrib / rec["contributions"]  # 1.0 for the same workload
    valu = c["SQ_INSTS_VALU"] * scale
    ach = valu * 64 / kernel_s / 1e12
    fetch = c["FETCH_SIZE_KiB"] * 1024 * rec["calibration"]["fetch_factor"] * scale
    write = c["WRITE_SIZE_KiB"] * 1024 * rec["calibration"]["write_factor"] * scale
    lds_bytes = (c["SQ_INSTS_LDS_LOAD_BANDWIDTH"] + c["SQ_INSTS_LDS_STORE_BANDWIDTH"]) * 64 * scale
    r.update(achieved=round(ach, 2), frac=round(ach / VALU_PEAK_TLANE, 4), traffic=round(fetch + write))
    r["unpacked_fp32"] = {"peak": round(VALU_UNPACKED_TLANE, 2), "frac": round(ach / VALU_UNPACKED_TLANE, 4),
                          "note": "the ceiling without packed FP32: 16 lanes per clock per SIMD"}
    r["valu_per_contribution"] = round(c["SQ_INSTS_VALU"] * 64 / rec["contributions"], 2)
    r["trans_per_contribution"] = round(c["SQ_INSTS_VALU_TRANS_F32"] * 64 / rec["contributions"], 2)
    r["issue_cycles_per_valu"] = round((c["SQ_BUSY_CYCLES"] / N_SE) / (c["SQ_INSTS_VALU"] / N_SIMD), 2)
    r["hbm_measured"] = {"achieved": round((fetch + write) / kernel_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round((fetch + write) / kernel_s / 1e9 / HBM_PEAK_GBS, 4),
                         "fetch_bytes": round(fetch), "write_bytes": round(write)}
    r["lds"] = {"achieved": round(lds_bytes / kernel_s / 1e12, 2), "peak": round(LDS_PEAK_TBS, 2), "unit": "TB/s",
                "frac": round(lds_bytes / kernel_s / 1e12 / LDS_PEAK_TBS, 4), "bytes_per_call": round(lds_bytes),
                "insts_per_contribution": round(c["SQ_INSTS_LDS"] * 64 / rec["contributions"], 2),
                "bank_conflict_cycle_frac": round(c["SQ_LDS_BANK_CONFLICT"] / max(c["SQ_LDS_IDX_ACTIVE"], 1), 4)}
    r["counters"] = {"source": src, "source_sha": rec["source_sha"]}
    if abs(scale - 1.0) > 1e-9:
        r["counters"]["scaled_by_contributions"] = round(scale, 4)
        r["counters"]["profiled_classes"] = rec.get("classes")
    return r


def host_cpu(threads: int) -> dict:
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cores": threads, "host_cpus": os.cpu_count(), "cpu_model": model}


def cpu_threads() -> int:
    """Threads for the CPU baseline: the process's CPU share (OMP_NUM_THREADS
    where set, as on the GPU hosts), else every CPU."""
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="pstm3d-precision",
                    choices=["pstm3d-precision", "pstm3d-single-offset", "pstm3d-vti-largest", "lsm3d", "job"])
    ap.add_argument("--classes", type=int, default=None,
                    help="offset classes (pstm3d-precision: 25, pstm3d-vti-largest: 60)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--cpu-tau-stride", type=int, default=16)
    ap.add_argument("--forms", default=None, help="comma list of trace forms to time (default: all)")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="first run a plain device copy of this many GiB (counter calibration)")
    ap.add_argument("--vt", action="store_true",
                    help="pstm3d-precision with v_rms = 1500 + 600 t instead of 2000 m/s")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip the end-to-end timing (host traces -> plan -> H2D -> forward)")
    ap.add_argument("--error-columns", type=int, default=6, help="image columns checked against the FP64 oracle")
    ap.add_argument("--no-grad", action="store_true", help="lsm3d: skip the image-power velocity gradient")
    return ap.parse_args()


def _workload(args, torch):
    from seisproc_amd import synthetic as S
    dev = "cuda"
    if args.workload == "pstm3d-precision":
        ncls = args.classes or 25
        s = S.precision_3d(n_classes=ncls, device=dev, vt=args.vt)
        cfg = {"workload": "pstm3d-precision", "offset_classes": ncls, "traces": s.n_traces,
               "samples": s.params.n_samples, "dt_ms": 2.0,
               "image": [s.params.n_il, s.params.n_xl, s.params.n_tau], "bin_m": 12.5,
               "aperture_m": 1500.0, "max_angle_deg": 60.0, "antialias": False, "kahan": True,
               "velocity": "v_rms = 1500 + 600 t m/s" if args.vt else "v_rms = 2000 m/s"}
        forms = ["bf16", "fp32", "fp16"]
    elif args.workload == "pstm3d-single-offset":
        s = S.single_offset_3d(device=dev)
        cfg = {"workload": "pstm3d-single-offset", "traces": s.n_traces, "samples": s.params.n_samples,
               "image": [s.params.n_il, s.params.n_xl, s.params.n_tau], "bin_m": 25.0,
               "aperture_m": 3000.0, "antialias": True,
               "velocity": "v_rms = (1500 + 600 t)(1 +- 5% lateral)"}
        forms = ["fp32", "fp16", "bf16"]
    else:
        # --classes n: n classes over the same 0-4000 m offset range (a launch must end within minutes)
        s = S.vti_largest_3d(device=dev, n_classes=args.classes or 60)
        cfg = {"workload": "pstm3d-vti-largest", "traces": s.n_traces, "classes": len(s.class_sizes),
               "samples": s.params.n_samples, "image": [s.params.n_il, s.params.n_xl, s.params.n_tau],
               "bin_m": 12.5, "aperture_m": 4000.0, "antialias": True, "vti": True}
        forms = ["fp16", "fp32"]
    return s, cfg, forms


def _error_columns(s, images, folds, n_cols, torch):
    """FP64 oracle on sampled image columns (every trace in their aperture)
    against every form's image and fold."""
    import numpy as np
    import oracle
    p = s.params
    rng = np.random.default_rng(12)
    picks = [(p.n_il // 2, p.n_xl // 2), (p.n_il // 4, (3 * p.n_xl) // 4), ((3 * p.n_il) // 4, p.n_xl // 3)]
    while len(picks) < n_cols:
        picks.append((int(rng.integers(p.n_il // 8, p.n_il - p.n_il // 8)),
                      int(rng.integers(p.n_xl // 8, p.n_xl - p.n_xl // 8))))
    picks = picks[:n_cols]
    sx, sy, rx, ry = s.numpy_geometry()
    tr = s.traces
    vel_full = s.vel
    refs, rfolds = [], []
    gots = {k: [] for k in images}
    gfolds = {k: [] for k in images}
    for il, xl in picks:
        x = p.origin_x + il * p.il_dx + xl * p.xl_dx
        y = p.origin_y + il * p.il_dy + xl * p.xl_dy
        mx, my = 0.5 * (sx + rx), 0.5 * (sy + ry)
        sel = np.nonzero((mx - x) ** 2 + (my - y) ** 2 < p.aperture_m ** 2)[0]
        q = p.replace(n_traces=len(sel), n_il=1, n_xl=1, origin_x=x, origin_y=y, vel_mode=0)
        vcol = vel_full if vel_full.dim() == 1 else vel_full[il, xl]
        eta = None
        if s.eta is not None:
            eta = s.eta[il, xl].double().cpu().numpy()
        d = tr[torch.as_tensor(sel, device=tr.device)].double().cpu().numpy()
        ref, rfold = oracle.forward(q, d, sx[sel], sy[sel], rx[sel], ry[sel], vcol.double().cpu().numpy(), eta=eta)
        refs.append(ref.ravel())
        rfolds.append(rfold.ravel())
        for k in images:
            gots[k].append(images[k][il, xl].double().cpu().numpy())
            gfolds[k].append(folds[k][il, xl].double().cpu().numpy())
    ref = np.concatenate(refs)
    rfold = np.concatenate(rfolds)
    out = {}
    for k in images:
        g = np.concatenate(gots[k])
        bar = 1e-4 if k == "fp32" else 1e-3
        err = float(np.linalg.norm(g - ref) / np.linalg.norm(ref))
        gf = np.concatenate(gfolds[k])
        flips = float(np.abs(gf - rfold).sum())
        # an aperture- or angle-edge flip (FP32 and FP64 masks deciding a
        # boundary tie differently) moves its sample by a whole contribution:
        # flips have their own bar, the image bar applies to the samples whose
        # fold agrees with the oracle's
        agree = gf == rfold
        err_agree = float(np.linalg.norm((g - ref)[agree]) / np.linalg.norm(ref[agree]))
        out[k] = {"rel_rms": err, "rel_rms_fold_agree": err_agree,
                  "rel_max": float(np.abs(g - ref).max() / np.abs(ref).max()),
                  "fold_flips": flips, "contributions_checked": float(rfold.sum()), "error_bar": bar,
                  "error_ok": bool(err_agree <= bar and flips <= 1e-4 * rfold.sum() + 1)}
    return out, f"FP64 oracle on {len(picks)} image columns {picks} (all traces in aperture)"


def _cpu_baseline(s, stride, torch):
    """Reference CPU method on one 100-point tile (10 x 10 columns) at every
    `stride`-th output time, with every trace its pre-filter keeps."""
    from oracle.ref_torch import migrate_kirchhoff_full_cpu
    p = s.params
    threads = cpu_threads()
    torch.set_num_threads(threads)
    if s.vel.dim() != 1:
        return None
    # the reference CPU method takes one constant velocity; for a v(t) field
    # its rate is taken at the median v_rms (the loop's cost does not depend
    # on the value)
    constant = float(s.vel.max()) == float(s.vel.min())
    v = float(s.vel.median())
    il0, xl0 = p.n_il // 2 - 5, p.n_xl // 2 - 5
    il = torch.arange(il0, il0 + 10, dtype=torch.float64)
    xl = torch.arange(xl0, xl0 + 10, dtype=torch.float64)
    ox = (p.origin_x + il[:, None] * p.il_dx + xl[None, :] * p.xl_dx).reshape(-1).float()
    oy = (p.origin_y + il[:, None] * p.il_dy + xl[None, :] * p.xl_dy).reshape(-1).float()
    src, rcv = s.src_xy, s.rcv_xy
    mx, my = 0.5 * (src[:, 0] + rcv[:, 0]), 0.5 * (src[:, 1] + rcv[:, 1])
    cx, cy = float(ox.double().mean()), float(oy.double().mean())
    reach = p.aperture_m + 10 * max(p.il_dx, p.xl_dy) + 100.0
    sel = torch.nonzero((mx - cx) ** 2 + (my - cy) ** 2 < reach ** 2).flatten()
    tr = s.traces[sel].t().contiguous().cpu()
    g = lambda t: t[sel].float().cpu()
    sx, sy, rx, ry = g(src[:, 0]), g(src[:, 1]), g(rcv[:, 0]), g(rcv[:, 1])
    tau_ms = (torch.arange(0, p.n_tau, stride, dtype=torch.float32) * p.dtau * 1000.0)
    depth = v * (tau_ms / 1000.0) / 2.0
    t0 = time.perf_counter()
    img, fold = migrate_kirchhoff_full_cpu(tr, sx, sy, rx, ry, ox, oy, depth, v, p.dt * 1000.0, 0.0,
                                           p.aperture_m, p.max_angle_deg, 10, 10)
    sec = time.perf_counter() - t0
    contrib = float(fold.double().sum())
    return {"value": contrib / sec / 1e9, "unit": "Gcontrib/s", **host_cpu(threads), "kind": "port",
            "sample": f"reference tile loop (torch CPU FP32, oracle/ref_torch.py) on 10x10 image columns "
                      f"x every {stride}th of {p.n_tau} output times, {len(sel)} traces in reach"
                      f"{'' if constant else f', constant v = {v:.0f} m/s (median v_rms)'}; "
                      f"{contrib:.3e} contributions in {sec:.1f} s, {threads} threads"}


def _cpu_sample_grid(s, stride, torch):
    """Shared set-up of the CPU samples: a 10 x 10-column tile at the image
    centre, every `stride`-th output time, the traces that reach it."""
    p = s.params
    il0, xl0 = p.n_il // 2 - 5, p.n_xl // 2 - 5
    il = torch.arange(il0, il0 + 10, dtype=torch.float64)
    xl = torch.arange(xl0, xl0 + 10, dtype=torch.float64)
    ox = (p.origin_x + il[:, None] * p.il_dx + xl[None, :] * p.xl_dx).reshape(-1).float()
    oy = (p.origin_y + il[:, None] * p.il_dy + xl[None, :] * p.xl_dy).reshape(-1).float()
    src, rcv = s.src_xy, s.rcv_xy
    mx, my = 0.5 * (src[:, 0] + rcv[:, 0]), 0.5 * (src[:, 1] + rcv[:, 1])
    cx, cy = float(ox.double().mean()), float(oy.double().mean())
    reach = p.aperture_m + 10 * max(p.il_dx, p.xl_dy) + 100.0
    sel = torch.nonzero((mx - cx) ** 2 + (my - cy) ** 2 < reach ** 2).flatten()
    g = lambda t: t[sel].float().cpu()
    geo = (g(src[:, 0]), g(src[:, 1]), g(rcv[:, 0]), g(rcv[:, 1]))
    tau_s = torch.arange(0, p.n_tau, stride, dtype=torch.float32) * p.dtau
    return sel, geo, ox, oy, tau_s


def _bench_lsm(args, torch):
    """Least-squares migration (CG iterations of adjoint + forward) and the
    image-power velocity gradient on the differentiation workload."""
    from seisproc_amd import synthetic as S
    from seisproc_amd.models.velocity_model import BSplineVelocity
    from seisproc_amd.processors.migration.lsm import cgls
    from seisproc_amd.processors.migration.velocity_update import image_power
    t_start = time.perf_counter()
    s = S.lsm_3d(device="cuda")
    plan, vel, data, p = s.info["plan"], s.vel, s.traces, s.params
    _, fold = plan.forward(data, vel, with_fold=True, flags=0)
    contrib = float(fold.double().sum())
    del fold
    torch.cuda.synchronize()
    print(f"[bench] lsm3d: {s.n_traces} traces x {p.n_samples}, image {plan.image_shape}, "
          f"{contrib:.4e} contributions per launch, setup {time.perf_counter() - t_start:.1f} s",
          file=sys.stderr, flush=True)
    if args.warmup:
        cgls(plan, data, vel, n_iter=args.warmup, flags=0)
    stamps = [time.perf_counter()]

    def on_iteration(it, rn, gn):
        stamps.append(time.perf_counter())
        print(f"[bench] CG iteration {it}: residual {rn:.4e} ({stamps[-1] - stamps[-2]:.1f} s)", file=sys.stderr,
              flush=True)

    res = cgls(plan, data, vel, n_iter=args.steps, flags=0, callback=on_iteration)
    # the rate averages the iterations after the first: the first one also
    # carries CGLS's initial migration of the data (one more forward launch)
    stamps = stamps[1:]
    per_it = [b - a for a, b in zip(stamps, stamps[1:])]
    sec_it = sum(per_it) / len(per_it) if per_it else res.seconds_per_iteration
    rate = 2.0 * contrib / sec_it / 1e9
    # the dominant kernel of an iteration is the adjoint: time one adjoint
    # and one forward launch with events on the launch stream
    stream = torch.cuda.current_stream()
    k_s = {}
    for name, fn in (("adjoint", lambda: plan.adjoint(res.image, vel, flags=0)),
                     ("forward", lambda: plan.forward(data, vel, with_fold=False, flags=0))):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        k_s[name] = e0.elapsed_time(e1) / 1e3
    print(f"[bench] kernel time per launch: adjoint {k_s['adjoint']:.3f} s, forward {k_s['forward']:.3f} s",
          file=sys.stderr, flush=True)
    r = res.data_residual_norms
    m_true = s.info["reflectivity"]
    model_err = float((res.image - m_true).double().norm() / m_true.double().norm())
    print(f"[bench] CG: {sec_it * 1e3:.1f} ms/iteration, {rate:.1f} Gcontrib/s, residual {r[0]:.4e} -> {r[-1]:.4e}",
          file=sys.stderr, flush=True)

    grad_check = None
    if not args.no_grad:
        # image-power gradient on a 32 x 32 x 64 B-spline lattice, central differences (h = 1 m/s) along a
        # random coefficient direction with the GPU forward kernel
        bs = BSplineVelocity(p.n_il, p.n_xl, p.n_tau, shape=(32, 32, 64), device="cuda")
        coef = bs.fit(vel[None, None, :].expand(plan.image_shape).contiguous())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J, grad = image_power(plan, data, bs, coef)
        torch.cuda.synchronize()
        t_grad = time.perf_counter() - t0
        delta = torch.randn(coef.shape, generator=torch.Generator(device="cuda").manual_seed(9), device="cuda")
        delta /= delta.abs().max()

        def J_at(c):
            with torch.no_grad():
                img, _ = plan.forward(data, bs.evaluate(c), with_fold=False, flags=0)
            return float((img.double() ** 2).sum())

        h = 1.0
        fd = (J_at(coef + h * delta) - J_at(coef - h * delta)) / (2 * h)
        an = float((grad.double() * delta.double()).sum())
        grad_check = {"seconds": round(t_grad, 3), "lattice": [32, 32, 64], "fd_h_mps": h, "fd": fd, "analytic": an,
                      "rel_diff": abs(fd - an) / max(abs(fd), 1e-300),
                      "note": "GPU forward J(c +- h delta); the analytic gradient holds the aperture masks fixed"}
        print(f"[bench] image-power gradient {t_grad:.2f} s, FD rel diff {grad_check['rel_diff']:.2e}",
              file=sys.stderr, flush=True)


    cpu = None
    if not args.no_cpu_baseline:
        from oracle.ref_torch import demigrate_kirchhoff_full_cpu, migrate_kirchhoff_full_cpu
        threads = cpu_threads()
        torch.set_num_threads(threads)
        stride = max(args.cpu_tau_stride, 1) * 4
        sel, geo, ox, oy, tau_s = _cpu_sample_grid(s, stride, torch)
        v_top = 1500.0  # the restated reference loop takes one velocity: its value is irrelevant to the rate
        depth = v_top * tau_s / 2.0
        tr = data[sel].t().contiguous().cpu()
        kw = (v_top, p.dt * 1000.0, 0.0, p.aperture_m, p.max_angle_deg, 10, 10)
        t0 = time.perf_counter()
        img, fold = migrate_kirchhoff_full_cpu(tr, *geo, ox, oy, depth, *kw)
        t_f = time.perf_counter() - t0
        t0 = time.perf_counter()
        demigrate_kirchhoff_full_cpu(img, p.n_samples, *geo, ox, oy, depth, *kw)
        t_a = time.perf_counter() - t0
        c = float(fold.sum())
        cpu = {"value": 2.0 * c / (t_f + t_a) / 1e9, "unit": "Gcontrib/s", **host_cpu(threads), "kind": "port",
               "sample": f"restated reference loop and its transpose (torch CPU FP32, oracle/ref_torch.py) on 10x10 "
                         f"image columns x every {stride}th of {p.n_tau} times, {len(sel)} traces in reach; "
                         f"{c:.3e} contributions per operator, forward {t_f:.1f} s + transpose {t_a:.1f} s"}
    line = {
        "metric": "LSM iteration Gcontrib/s (adjoint + forward per CG iteration)",
        "value": round(rate, 3),
        "unit": "Gcontrib/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_it * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (sparse reflectivity x Ricker, data modelled by the adjoint kernel + noise at SNR 10)",
        "config": {"workload": "lsm3d", "traces": s.n_traces, "samples": p.n_samples, "classes": len(s.class_sizes),
                   "image": list(plan.image_shape), "bin_m": p.il_dx, "aperture_m": p.aperture_m,
                   "max_angle_deg": p.max_angle_deg, "velocity": "v_rms 1500 -> 3500 m/s over the record"},
        "contributions_per_launch": contrib,
        "cg": {"iterations": res.iterations, "residual_first": r[0], "residual_last": r[-1],
               "residual_ratio": r[-1] / r[0], "model_rel_error": model_err,
               "residuals": [float(x) for x in r],
               "seconds_per_iteration": [round(x, 3) for x in per_it]},
        "velocity_gradient": grad_check,
        "kernel_s_per_launch": {k: round(v, 4) for k, v in k_s.items()},
        "roofline": dict(roofline(contrib, k_s["adjoint"], 8, load_counters("lsm3d", "adjoint", len(s.class_sizes))),
                         kernel="k_tile adjoint (one launch: 16 offset-band dispatches)"),
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu"] = round(rate / cpu["value"], 1)
    print(json.dumps(line))


def _end_to_end(s, form, vel, torch):
    """One migration as the engine call sees it (MigrationEngine.
    migrate_bin_full, reference processors/migration/migration_engine.py:
    225-360): FP32 traces and FP64 coordinates in host memory -> device copy
    -> plan build (FP64 reduction, binning, sort) -> quantisation to the
    trace form -> forward.  Timed from pinned and from pageable host memory;
    seconds per stage, each stage synchronised."""
    from seisproc_amd.processors.migration.hip_ops import PstmPlan, quantize_traces
    host_pageable = s.traces.cpu()
    host_pinned = torch.empty(host_pageable.shape, dtype=torch.float32, pin_memory=True)
    host_pinned.copy_(host_pageable)
    src, rcv = s.src_xy.cpu(), s.rcv_xy.cpu()
    out = {}
    for kind, host in (("pinned", host_pinned), ("pageable", host_pageable)):
        torch.cuda.synchronize()
        t = {}
        t0 = time.perf_counter()
        dev = host.to("cuda", non_blocking=kind == "pinned")
        torch.cuda.synchronize()
        t["h2d_traces_s"] = time.perf_counter() - t0
        t1 = time.perf_counter()
        plan = PstmPlan(s.params, src, rcv)
        torch.cuda.synchronize()
        t["plan_build_s"] = time.perf_counter() - t1
        t1 = time.perf_counter()
        tr, sc = (dev, None) if form == "fp32" else quantize_traces(
            dev, torch.bfloat16 if form == "bf16" else torch.float16)
        torch.cuda.synchronize()
        t["quantize_s"] = time.perf_counter() - t1
        t1 = time.perf_counter()
        img, fold = plan.forward(tr, vel, scale=sc, comp=torch.zeros(plan.image_shape, device="cuda"))
        torch.cuda.synchronize()
        t["forward_s"] = time.perf_counter() - t1
        t["total_s"] = time.perf_counter() - t0
        t["h2d_GBs"] = host.numel() * 4 / t["h2d_traces_s"] / 1e9
        t["gcontrib_s"] = float(fold.double().sum()) / t["total_s"] / 1e9
        out[kind] = {k: round(v, 4) for k, v in t.items()}
        del dev, tr, sc, img, fold, plan
    out["note"] = f"{form} traces quantised on the device after an FP32 host-to-device copy"
    return out


def _bench_job(args, torch):
    """The bin-by-bin job driver on the precision workload held in host
    memory: 25 offset bins (one per class) read from a host array, gathered
    into pinned buffers and copied on a second stream while the previous bin
    migrates (BF16 traces, compensated accumulation, running stack; no
    per-bin files, no checkpoint).  Reports the whole job's rate and how much
    of the host-side staging the migration hides."""
    import numpy as np
    from seisproc_amd import synthetic as S
    from seisproc_amd.models.binning import create_uniform_offset_binning
    from seisproc_amd.processors.migration.job_driver import ArrayBinSource, MigrationJobRunner
    t_start = time.perf_counter()
    s = S.precision_3d(n_classes=args.classes or 25, device="cuda")
    host = s.traces.cpu().numpy()
    sx, sy, rx, ry = s.numpy_geometry()
    del s.traces
    torch.cuda.empty_cache()
    n_cls = len(s.class_sizes)
    binning = create_uniform_offset_binning(50.0, 50.0 + 100.0 * n_cls, n_cls)
    source = ArrayBinSource(host, sx, sy, rx, ry, binning)
    print(f"[bench] job: {host.shape[0]} host traces in {n_cls} bins, setup {time.perf_counter() - t_start:.1f} s",
          file=sys.stderr, flush=True)
    runs = []
    for i in range(args.warmup + args.steps):
        runner = MigrationJobRunner(s.params, s.vel.contiguous(), Path("/tmp") / f"pstm_job_bench_{os.getpid()}",
                                    trace_dtype="bf16", write_bins=False, checkpoint=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = runner.run(source, resume=False)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        contrib = float(res.stack_fold.double().sum())
        stage = sum(v["stage_host_s"] for v in res.timings.values())
        gpu = sum(v["gpu_s"] for v in res.timings.values())
        runs.append((wall, contrib, stage, gpu))
        print(f"[bench] job run {i + 1}: {wall:.2f} s, {contrib / wall / 1e9:.1f} Gcontrib/s, host staging {stage:.2f} s, "
              f"GPU work {gpu:.2f} s", file=sys.stderr, flush=True)
    wall, contrib, stage, gpu = runs[-1] if not args.steps else min(runs[args.warmup:])
    line = {
        "metric": "PSTM job Gcontrib/s (host traces -> bins -> stack)", "value": round(contrib / wall / 1e9, 3),
        "unit": "Gcontrib/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16-traces/f32-compute", "data": "synthetic (the precision workload's traces in host memory)",
        "config": {"workload": "job", "traces": int(host.shape[0]), "bins": n_cls, "samples": int(host.shape[1]),
                   "image": [s.params.n_il, s.params.n_xl, s.params.n_tau], "write_bins": False},
        "contributions": contrib,
        "job": {"wall_s": round(wall, 3), "host_staging_s": round(stage, 3), "gpu_s": round(gpu, 3),
                "serial_s": round(stage + gpu, 3),
                "staging_hidden_frac": round(max(0.0, min(1.0, (stage + gpu - wall) / max(stage, 1e-9))), 3),
                "note": "host_staging_s: the bins' gathers into pinned memory; gpu_s: each bin's plan, "
                        "quantisation and forward (events on the compute stream); the next bin's gather and "
                        "host-to-device copy run while a bin migrates, so the job takes less than their sum"},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line))


def main():
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py: the engine is measured on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is design-only", file=sys.stderr)
        sys.exit(2)
    import torch
    from seisproc_amd.processors.migration.hip_ops import PstmPlan, quantize_traces

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe_gib > 0:
        n = int(args.copy_probe_gib * (1 << 30)) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.workload == "lsm3d":
        _bench_lsm(args, torch)
        return
    if args.workload == "job":
        _bench_job(args, torch)
        return
    t_start = time.perf_counter()
    s, cfg, forms = _workload(args, torch)
    if args.forms:
        forms = [f for f in args.forms.split(",") if f]
    plan = PstmPlan(s.params, s.src_xy, s.rcv_xy)
    vel = s.vel.contiguous()
    eta = s.eta
    tr32 = s.traces
    data = {"fp32": (tr32, None)}
    for f, dt in (("bf16", torch.bfloat16), ("fp16", torch.float16)):
        if f in forms:
            data[f] = quantize_traces(tr32, dt)
    torch.cuda.synchronize()
    print(f"[bench] {cfg['workload']}: {s.n_traces} traces, kept {plan.kept()}, setup "
          f"{time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)

    image = torch.zeros(plan.image_shape, dtype=torch.float32, device="cuda")
    comp = torch.zeros_like(image)
    fold = torch.zeros_like(image)

    def step(form):
        tr, sc = data[form]
        image.zero_()
        comp.zero_()
        fold.zero_()
        plan.forward(tr, vel, scale=sc, eta=eta, image=image, comp=comp, fold=fold)

    main_form = forms[0]
    results, images, folds = {}, {}, {}
    stream = torch.cuda.current_stream()
    for form in forms:
        w = args.warmup if form == main_form else 0
        k = args.steps if form == main_form else 1
        tr, sc = data[form]
        for i in range(w):
            t_w = time.perf_counter()
            step(form)
            torch.cuda.synchronize()
            print(f"[bench] {form} warm-up {i + 1}/{w}: {time.perf_counter() - t_w:.2f} s", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(k)]
        t0 = time.perf_counter()
        for i, (e0, e1) in enumerate(ev):
            image.zero_()
            comp.zero_()
            fold.zero_()
            # a step of the largest workload runs for minutes: its launches go one
            # offset band at a time and each finished band prints a line
            prog = None
            if args.workload == "pstm3d-vti-largest":
                t_s = time.perf_counter()
                prog = (lambda f, msg, i=i, t_s=t_s: print(
                    f"[bench] {form} step {i + 1}/{k}: {msg} ({time.perf_counter() - t_s:.1f} s)",
                    file=sys.stderr, flush=True))
            e0.record(stream)
            plan.forward(tr, vel, scale=sc, eta=eta, image=image, comp=comp, fold=fold, progress_callback=prog)
            e1.record(stream)
            e1.synchronize()  # one line per step (long launches)
            print(f"[bench] {form} step {i + 1}/{k}: kernel {e0.elapsed_time(e1):.1f} ms", file=sys.stderr,
                  flush=True)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / k
        kts = sorted(e0.elapsed_time(e1) / 1e3 for e0, e1 in ev)
        kern = sum(kts) / k
        contrib = float(fold.double().sum())
        images[form] = (image + comp).cpu()
        folds[form] = fold.cpu()
        results[form] = {"wall_s": wall, "kernel_s": kern, "kernel_median_s": kts[len(kts) // 2],
                         "contributions": contrib, "gcontrib_s": contrib / wall / 1e9}
        print(f"[bench] {form}: {wall * 1e3:.1f} ms/step, kernel {kern * 1e3:.1f} ms, "
              f"{contrib / wall / 1e9:.1f} Gcontrib/s", file=sys.stderr, flush=True)

    err, err_desc = ({}, None)
    if not args.no_error:
        err, err_desc = _error_columns(s, images, folds, args.error_columns, torch)
        print(f"[bench] errors vs FP64 oracle: {err}", file=sys.stderr, flush=True)
        for k, e in err.items():
            if not e["error_ok"]:
                print(f"[bench] WARNING: the {k} form misses its error bar (fold-agreeing {e['rel_rms_fold_agree']:.3e}, "
                      f"all {e['rel_rms']:.3e} against {e['error_bar']}; {e['fold_flips']:.0f} fold flips)",
                      file=sys.stderr, flush=True)
    del images, folds

    e2e = None
    if not args.no_e2e and args.workload == "pstm3d-precision":
        e2e = _end_to_end(s, main_form, vel, torch)
        print(f"[bench] end to end: {e2e}", file=sys.stderr, flush=True)

    m = results[main_form]
    aa = bool(s.params.flags & 1)
    taps = 6 if aa else 2
    sbytes = 4 if main_form == "fp32" else 2
    B = taps * sbytes
    classes = cfg.get("offset_classes", cfg.get("classes", 1))  # single-offset: one class
    # the largest VTI workload is profiled on 12 of its classes (a counter pass
    # over all 60 takes minutes per counter set); its totals scale by contributions
    rl = roofline(m["contributions"], m["kernel_s"], B,
                  load_counters(cfg["workload"] + ("-vt" if args.vt else ""), main_form, classes,
                                any_classes=cfg["workload"] == "pstm3d-vti-largest"))
    rl["kernel"] = "k_tile forward (one pstm_forward call: 16 offset-band dispatches)"
    cpu = None
    if not args.no_cpu_baseline:
        cpu = _cpu_baseline(s, args.cpu_tau_stride, torch)
    line = {
        "metric": "PSTM Gcontrib/s (trace×image pt), BF16-trace vs FP32 at 1e-3 rel err",
        "value": round(m["gcontrib_s"], 3),
        "unit": "Gcontrib/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["wall_s"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16": "bf16-traces/f32-compute", "fp16": "f16-traces/f32-compute",
                  "fp32": "f32"}[main_form],
        "data": "synthetic (seeded Ricker events at exact DSR times + noise)",
        "config": cfg,
        "contributions": m["contributions"],
        "forms": {k: {"gcontrib_s": round(v["gcontrib_s"], 3), "ms_per_step": round(v["wall_s"] * 1e3, 3),
                      "kernel_ms": round(v["kernel_s"] * 1e3, 3),
                      "kernel_median_ms": round(v["kernel_median_s"] * 1e3, 3), **err.get(k, {})}
                  for k, v in results.items()},
        "error_check": err_desc,
        # the headline form's error against the FP64 oracle is within its bar
        # (1e-3 for BF16/FP16 traces, 1e-4 for FP32) with at most 1e-4 fold flips
        "error_ok": err.get(main_form, {}).get("error_ok") if err else None,
        "roofline": rl,
        "end_to_end": e2e,
        "cpu_baseline": cpu,
    }
    if "fp32" in results and main_form != "fp32":
        line["speedup_vs_fp32_form"] = round(m["gcontrib_s"] / results["fp32"]["gcontrib_s"], 3)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
