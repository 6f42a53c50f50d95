"""GPU: the benchmark workloads at their full sizes, through properties that
hold at any size (the FP64 oracle checks sampled image columns of the same
workloads in bench.py; it cannot sum these images whole in seconds).

* splitting the traces over calls that accumulate into one image gives the
  same fold exactly (each contribution's masks depend only on its own trace
  and image sample) and the same image to FP32 rounding;
* a call repeated on the same input gives a bit-identical image and fold (no
  atomics in the forward);
* the antialias forward from traces integrated once per call equals the one
  that integrates each staged window (to 1e-5: FP32 rounding of the
  integrals differs between them);
* the forward and the adjoint are transposes: <L d, m> = <d, L^T m>.
"""
import pytest
import torch

from seisproc_amd import synthetic as S

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _plan(p, src, rcv):
    from seisproc_amd.processors.migration.hip_ops import PstmPlan
    return PstmPlan(p, src, rcv)


def _zeros(shape, n):
    return [torch.zeros(shape, device=DEV) for _ in range(n)]


def _rel(a, b):
    a, b = a.double(), b.double()
    return float((a - b).norm() / b.norm())


def _split_forward(s, traces, scale, cuts, **kw):
    """Forward over trace ranges [cuts[i], cuts[i+1]), one plan each, summed
    into one (image, comp, fold)."""
    p = s.params
    shape = (p.n_il, p.n_xl, p.n_tau)
    img, comp, fold = _zeros(shape, 3)
    for a, b in zip(cuts[:-1], cuts[1:]):
        pl = _plan(p.replace(n_traces=int(b - a)), s.src_xy[a:b], s.rcv_xy[a:b])
        pl.forward(traces[a:b], s.vel, scale=None if scale is None else scale[a:b], eta=s.eta,
                   image=img, comp=comp, fold=fold, **kw)
    torch.cuda.synchronize()
    return img, comp, fold


def test_precision_full_size_split_and_repeat():
    # 1,000,000 traces x 1001 samples -> 400 x 400 x 1001, BF16 traces with
    # per-trace scales, compensated accumulation (the headline workload)
    from seisproc_amd.processors.migration.hip_ops import quantize_traces
    s = S.precision_3d(device=DEV)
    p = s.params
    assert (p.n_traces, p.n_samples, p.n_il, p.n_xl, p.n_tau) == (1_000_000, 1001, 400, 400, 1001)
    q, sc = quantize_traces(s.traces.to(DEV), torch.bfloat16)
    n = p.n_traces
    whole = _split_forward(s, q, sc, [0, n])
    again = _split_forward(s, q, sc, [0, n])
    for a, b in zip(whole, again):
        assert torch.equal(a, b)
    halves = _split_forward(s, q, sc, [0, 412_345, n])
    contrib = float(whole[2].double().sum())
    print(f"[full-precision] contributions {contrib:.4e}", flush=True)
    assert contrib > 1.2e13
    assert torch.equal(whole[2], halves[2])
    err = _rel(halves[0].double() + halves[1].double(), whole[0].double() + whole[1].double())
    print(f"[full-precision] split vs whole rel_rms {err:.2e}", flush=True)
    assert err < 1e-6


def test_single_offset_full_size_integrated_windows_split_and_repeat():
    # 40,000 traces x 1501 samples -> 200 x 200 x 1501, lateral v_rms, triangle
    # antialias, FP32 traces
    s = S.single_offset_3d(device=DEV)
    p = s.params
    assert (p.n_traces, p.n_samples, p.n_il, p.n_xl, p.n_tau) == (40_000, 1501, 200, 200, 1501)
    assert p.flags & 1
    tr = s.traces.to(DEV)
    n = p.n_traces
    whole = _split_forward(s, tr, None, [0, n])
    again = _split_forward(s, tr, None, [0, n])
    for a, b in zip(whole, again):
        assert torch.equal(a, b)
    windows = _split_forward(s, tr, None, [0, n], integrated=False)
    halves = _split_forward(s, tr, None, [0, 17_001, n])
    tot = whole[0].double() + whole[1].double()
    assert torch.equal(whole[2], windows[2]) and torch.equal(whole[2], halves[2])
    e_win = _rel(windows[0].double() + windows[1].double(), tot)
    e_split = _rel(halves[0].double() + halves[1].double(), tot)
    print(f"[full-single-offset] contributions {float(whole[2].double().sum()):.4e}; per-window vs integrated "
          f"{e_win:.2e}, split vs whole {e_split:.2e}", flush=True)
    # the integrals of whole 1501-sample traces round C in FP32 at the
    # magnitude C reaches along the record, a window's own integrals at the
    # window's: the paths differ by ~1e-6, a hundredth of the 1e-4 parity bar
    assert e_win < 1e-5 and e_split < 1e-6


def test_lsm_full_size_dot_product():
    # 1,048,576 traces x 751 samples <-> 256 x 256 x 751 (16 offset classes,
    # v_rms 1500 -> 3500 m/s, 2000 m aperture)
    s = S.lsm_3d(device=DEV)
    plan, vel = s.info["plan"], s.vel
    p = s.params
    assert (p.n_traces, p.n_samples, p.n_il, p.n_xl, p.n_tau) == (1_048_576, 751, 256, 256, 751)
    g = torch.Generator(device=DEV).manual_seed(7)
    d = torch.randn(s.traces.shape, generator=g, device=DEV)
    m = torch.randn(plan.image_shape, generator=g, device=DEV)
    Ld, _ = plan.forward(d, vel, with_fold=False, flags=0)
    Ltm = plan.adjoint(m, vel, flags=0)
    torch.cuda.synchronize()
    a = float(torch.dot(Ld.double().flatten(), m.double().flatten()))
    b = float(torch.dot(d.double().flatten(), Ltm.double().flatten()))
    print(f"[full-lsm-dot] <Ld,m>={a:.9e} <d,Ltm>={b:.9e} rel={abs(a - b) / abs(a):.2e}", flush=True)
    assert abs(a - b) <= 1e-5 * abs(a)
