"""Tensor-level wrappers of the HIP migration kernels (libpstm_hip.so).

``PstmPlan`` owns the per-job device workspace (binned trace geometry) and
exposes the three operators on torch tensors that live on the GPU:

  forward   image[il][xl][tau] (+ fold)   <- traces      (Kirchhoff sum)
  adjoint   traces                        <- image       (demigration)
  vel_grad  d<g, image>/dv_rms per image sample (masks held fixed)

All buffers are caller-owned torch tensors; the library allocates nothing.
Work is enqueued on the current torch stream.  There is no CPU path: a
missing GPU or library raises.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional, Tuple

import torch

from ...abi_types import (FLAG_AA, FLAG_KAHAN, FLAG_VTI, PSTM_BF16, PSTM_F16, PSTM_F32, VEL_1D,
                          VEL_3D, PstmParams)
from ..kernel_backend import check, load_library

_DTYPE_CODE = {torch.float32: PSTM_F32, torch.bfloat16: PSTM_BF16, torch.float16: PSTM_F16}


def _stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _need_cuda(t: torch.Tensor, name: str, dtype=None, shape=None) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be on the GPU (got {t.device}); the migration path has no CPU fallback")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


# Largest squared isotropic leg time (samples^2) of a VTI call, scaled: the
# kernels evaluate each anelliptic leg as t = N rsq(N D) with
# N <= (1 + e) (tqs + p)^2 and D <= (1 + 2 e) (tqs + p), e = max(eta, 0), so
# N D < 2^126 (a quarter of the FP32 maximum) while
# ((1 + e) (1 + 2 e))^(1/3) (tqs + p) < 2^42.
VTI_MAX_LEG2 = float(2 ** 42)


def _check_eta(eta: Optional[torch.Tensor], vel: torch.Tensor, p: "PstmParams") -> None:
    """The kernels take eta as given: the anelliptic time and its window
    bounds need 1 + 2 eta > 0 at every sample, and the leg times must stay
    within the FP32 range of the anelliptic product (one device reduction and
    a synchronisation per call)."""
    if eta is None:
        raise ValueError("VTI migration needs eta")
    _need_cuda(eta, "eta", torch.float32, tuple(vel.shape))
    if not eta.numel():
        return
    lo, hi, vmin = (float(x) for x in torch.stack([eta.min(), eta.max(), vel.min()]).cpu())
    if not lo > -0.5:
        raise ValueError(f"eta must be > -0.5 everywhere (1 + 2 eta > 0), min is {lo}")
    if vmin > 0.0:
        leg2 = vti_leg2_bound(p, vmin, hi)
        if leg2 >= VTI_MAX_LEG2:
            raise ValueError(f"VTI leg times up to {leg2 ** 0.5:.3g} samples (aperture {p.aperture_m} m, "
                             f"v_min {vmin} m/s, dt {p.dt} s) exceed the anelliptic range of 2^21 samples")


def vti_leg2_bound(p: "PstmParams", vmin: float, eta_max: float) -> float:
    """((1 + e) (1 + 2 e))^(1/3), e = max(eta_max, 0), times the largest
    squared isotropic leg time (samples^2) of a call: latest output time, a
    leg at the aperture, the smallest velocity.  Below VTI_MAX_LEG2 the
    kernels' anelliptic product N D stays below 2^126."""
    tau_max = p.tau0 + (p.n_tau - 1) * p.dtau
    leg2 = (tau_max / (2.0 * p.dt)) ** 2 + (p.aperture_m / (vmin * p.dt)) ** 2
    e = max(eta_max, 0.0)
    return leg2 * ((1.0 + e) * (1.0 + 2.0 * e)) ** (1.0 / 3.0)


class PstmPlan:
    """Device plan of one migration job (geometry binned once, reused by every
    forward / adjoint / gradient call of the job).

    params:  PstmParams (or dict) describing traces, image grid and options.
    src_xy, rcv_xy: (n_traces, 2) float64 world coordinates, GPU or host;
             host arrays are copied once (FP64 is kept until the on-GPU
             reduction to local coordinates).
    """

    def __init__(self, params, src_xy, rcv_xy, device: Optional[torch.device] = None):
        self.lib = load_library()
        p = params if isinstance(params, PstmParams) else PstmParams.from_dict(params)
        self.params = p
        if not torch.cuda.is_available():
            raise RuntimeError("PstmPlan needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        src = torch.as_tensor(src_xy, dtype=torch.float64).to(self.device).contiguous()
        rcv = torch.as_tensor(rcv_xy, dtype=torch.float64).to(self.device).contiguous()
        n = int(p.n_traces)
        _need_cuda(src, "src_xy", torch.float64, (n, 2))
        _need_cuda(rcv, "rcv_xy", torch.float64, (n, 2))
        check(self.lib.pstm_validate(ctypes.byref(p)), "pstm_validate", self.lib)
        nbytes = int(self.lib.pstm_plan_bytes(ctypes.byref(p)))
        if nbytes == 0:
            check(-1, "pstm_plan_bytes", self.lib)
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        check(self.lib.pstm_plan_build(ctypes.byref(p), _ptr(src), _ptr(rcv), _ptr(self.workspace),
                                       nbytes, ctypes.c_void_p(_stream_handle())),
              "pstm_plan_build", self.lib)
        self._src, self._rcv = src, rcv  # kept alive until the build has run

    # ------------------------------------------------------------------
    @property
    def image_shape(self) -> Tuple[int, int, int]:
        p = self.params
        return (p.n_il, p.n_xl, p.n_tau)

    def kept(self) -> int:
        """Number of traces whose midpoint can reach the image (synchronises)."""
        return int(check(self.lib.pstm_plan_kept(_ptr(self.workspace), ctypes.c_void_p(_stream_handle())),
                         "pstm_plan_kept", self.lib))

    def _vel_mode(self, vel: torch.Tensor) -> int:
        p = self.params
        if vel.dim() == 1:
            _need_cuda(vel, "vel", torch.float32, (p.n_tau,))
            return VEL_1D
        _need_cuda(vel, "vel", torch.float32, self.image_shape)
        return VEL_3D

    def _common_layout(self, vel: torch.Tensor, eta: Optional[torch.Tensor]):
        """The kernels read v_rms and eta in one layout (pstm_params_t.vel_mode):
        a 1D field is broadcast to the image grid when the other is 3D."""
        if eta is None or vel.dim() == eta.dim():
            return vel, eta
        shape = self.image_shape
        if vel.dim() == 1:
            vel = vel[None, None, :].expand(shape).contiguous()
        else:
            eta = eta[None, None, :].expand(shape).contiguous()
        return vel, eta

    def _call_params(self, flags: Optional[int], dtype_code: int, vel_mode: int) -> PstmParams:
        p = self.params
        return p.replace(flags=p.flags if flags is None else flags, dtype=dtype_code, vel_mode=vel_mode)

    # ------------------------------------------------------------------
    def forward(self, traces: torch.Tensor, vel: torch.Tensor, *, scale: Optional[torch.Tensor] = None,
                eta: Optional[torch.Tensor] = None, image: Optional[torch.Tensor] = None,
                comp: Optional[torch.Tensor] = None, fold: Optional[torch.Tensor] = None,
                with_fold: bool = True, flags: Optional[int] = None,
                progress_callback: Optional[Callable[[float, str], None]] = None,
                integrated: Optional[bool] = None):
        """Accumulate the migrated image (and fold) of ``traces``.

        traces: (n_traces, n_samples) float32 / bfloat16 / float16 (trace-major)
        scale:  (n_traces,) float32 per-trace scale (low-precision traces)
        vel:    (n_tau,) or (n_il, n_xl, n_tau) float32 RMS velocity
        eta:    same layout as vel (PSTM_FLAG_VTI)
        image/comp/fold: optional accumulators (allocated as zeros if None).
        progress_callback: called as (fraction, message) after each offset
            band of the plan has been migrated (the launches are split per
            band and the stream is synchronised between them; the sums are
            the same as one call's), like the reference kernel's callback
            (processors/migration/kirchhoff_kernel.py:760-774).
        integrated: with the antialias filter, integrate every trace once per
            call (pstm_integrate_traces, 12 bytes per sample of device memory
            for the call) and migrate from the integrals
            (pstm_forward_integrated), or, with False, integrate each staged
            window per image tile (pstm_forward).  The default (None) takes the
            first unless the device cannot hold the integrals.
        Returns (image, fold) -- fold is None when with_fold is False.
        """
        p = self.params
        if traces.dtype not in _DTYPE_CODE:
            raise ValueError(f"unsupported trace dtype {traces.dtype}")
        _need_cuda(traces, "traces", traces.dtype, (p.n_traces, p.n_samples))
        flags_eff = p.flags if flags is None else flags
        vel, eta = self._common_layout(vel, eta if flags_eff & FLAG_VTI else None)
        vm = self._vel_mode(vel)
        cp = self._call_params(flags, _DTYPE_CODE[traces.dtype], vm)
        if scale is not None:
            _need_cuda(scale, "scale", torch.float32, (p.n_traces,))
        if cp.flags & FLAG_VTI:
            _check_eta(eta, vel, cp)
        if image is None:
            image = torch.zeros(self.image_shape, dtype=torch.float32, device=self.device)
        _need_cuda(image, "image", torch.float32, self.image_shape)
        if comp is not None:
            _need_cuda(comp, "comp", torch.float32, self.image_shape)
        if fold is None and with_fold:
            fold = torch.zeros(self.image_shape, dtype=torch.float32, device=self.device)
        if fold is not None:
            _need_cuda(fold, "fold", torch.float32, self.image_shape)
        stream = ctypes.c_void_p(_stream_handle())
        outs = (_ptr(vel), _ptr(eta if cp.flags & FLAG_VTI else None), _ptr(image), _ptr(comp), _ptr(fold), stream)
        int_a = int_c = None
        if (cp.flags & FLAG_AA) and integrated is not False and p.n_traces > 0:
            try:
                int_a = torch.empty((p.n_traces, p.n_samples), dtype=torch.float64, device=self.device)
                int_c = torch.empty((p.n_traces, p.n_samples), dtype=torch.float32, device=self.device)
            except torch.cuda.OutOfMemoryError:
                if integrated:  # asked for explicitly
                    raise
                int_a = int_c = None  # no room for the integrals: integrate per window instead
        if int_a is not None:
            check(self.lib.pstm_integrate_traces(ctypes.byref(cp), _ptr(traces), _ptr(scale), _ptr(int_a),
                                                 _ptr(int_c), stream), "pstm_integrate_traces", self.lib)

            def run(b0, b1):
                check(self.lib.pstm_forward_integrated(ctypes.byref(cp), _ptr(self.workspace), b0, b1, _ptr(int_a),
                                                       _ptr(int_c), *outs), "pstm_forward_integrated", self.lib)
        else:
            def run(b0, b1):
                check(self.lib.pstm_forward_bands(ctypes.byref(cp), _ptr(self.workspace), b0, b1, _ptr(traces),
                                                  _ptr(scale), *outs), "pstm_forward_bands", self.lib)
        n_bands = self.lib.pstm_plan_bands(ctypes.byref(cp))
        check(min(n_bands, 0), "pstm_plan_bands", self.lib)
        if progress_callback is None:
            run(0, n_bands)
            return image, fold
        for band in range(n_bands):
            run(band, band + 1)
            torch.cuda.current_stream(self.device).synchronize()
            progress_callback((band + 1) / n_bands, f"offset band {band + 1}/{n_bands} migrated")
        return image, fold

    def adjoint(self, image: torch.Tensor, vel: torch.Tensor, *, eta: Optional[torch.Tensor] = None,
                out: Optional[torch.Tensor] = None, flags: Optional[int] = None) -> torch.Tensor:
        """Demigrate ``image`` into FP32 traces (accumulates into ``out``)."""
        p = self.params
        _need_cuda(image, "image", torch.float32, self.image_shape)
        flags_eff = p.flags if flags is None else flags
        vel, eta = self._common_layout(vel, eta if flags_eff & FLAG_VTI else None)
        vm = self._vel_mode(vel)
        cp = self._call_params(flags, PSTM_F32, vm)
        if cp.flags & FLAG_VTI:
            _check_eta(eta, vel, cp)
        if out is None:
            out = torch.zeros((p.n_traces, p.n_samples), dtype=torch.float32, device=self.device)
        _need_cuda(out, "out", torch.float32, (p.n_traces, p.n_samples))
        check(self.lib.pstm_adjoint(ctypes.byref(cp), _ptr(self.workspace), _ptr(image), _ptr(vel),
                                    _ptr(eta if cp.flags & FLAG_VTI else None), _ptr(out),
                                    ctypes.c_void_p(_stream_handle())),
              "pstm_adjoint", self.lib)
        return out

    def vel_grad(self, traces: torch.Tensor, vel: torch.Tensor, image_grad: torch.Tensor, *,
                 scale: Optional[torch.Tensor] = None, out: Optional[torch.Tensor] = None,
                 flags: Optional[int] = None) -> torch.Tensor:
        """d<image_grad, forward(traces; vel)>/d vel per image sample."""
        p = self.params
        if traces.dtype not in _DTYPE_CODE:
            raise ValueError(f"unsupported trace dtype {traces.dtype}")
        _need_cuda(traces, "traces", traces.dtype, (p.n_traces, p.n_samples))
        _need_cuda(vel, "vel", torch.float32, self.image_shape)
        _need_cuda(image_grad, "image_grad", torch.float32, self.image_shape)
        if scale is not None:
            _need_cuda(scale, "scale", torch.float32, (p.n_traces,))
        cp = self._call_params(flags, _DTYPE_CODE[traces.dtype], VEL_3D)
        if out is None:
            out = torch.zeros(self.image_shape, dtype=torch.float32, device=self.device)
        _need_cuda(out, "out", torch.float32, self.image_shape)
        check(self.lib.pstm_vel_grad(ctypes.byref(cp), _ptr(self.workspace), _ptr(traces), _ptr(scale),
                                     _ptr(vel), _ptr(image_grad), _ptr(out),
                                     ctypes.c_void_p(_stream_handle())),
              "pstm_vel_grad", self.lib)
        return out


def quantize_traces(traces: torch.Tensor, dtype: torch.dtype) -> Tuple[torch.Tensor, torch.Tensor]:
    """FP32 traces -> (bf16/fp16 traces, per-trace FP32 scale) on the GPU
    (scale = max|trace|, round to nearest even)."""
    lib = load_library()
    if dtype not in (torch.bfloat16, torch.float16):
        raise ValueError("dtype must be torch.bfloat16 or torch.float16")
    _need_cuda(traces, "traces", torch.float32)
    if traces.dim() != 2:
        raise ValueError("traces must be (n_traces, n_samples)")
    n, ns = traces.shape
    out = torch.empty((n, ns), dtype=dtype, device=traces.device)
    scale = torch.empty((n,), dtype=torch.float32, device=traces.device)
    check(lib.pstm_quantize_traces(_ptr(traces), n, ns, _DTYPE_CODE[dtype], _ptr(out), _ptr(scale),
                                   ctypes.c_void_p(_stream_handle())),
          "pstm_quantize_traces", lib)
    return out, scale


__all__ = ["PstmPlan", "quantize_traces", "FLAG_AA", "FLAG_KAHAN", "FLAG_VTI"]
