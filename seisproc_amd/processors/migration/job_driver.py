"""Bin-by-bin migration job: binned trace reading, pinned-host staging with
asynchronous host-to-device copies overlapped with the previous bin's
migration, per-bin outputs in the reference's format, a running stack, and a
per-bin checkpoint from which an interrupted job resumes bit for bit.

Follows the job loop of views/migration_monitor_dialog.py:177-449 (one
``migrate_bin_full`` per enabled offset-azimuth bin of a BinnedDataset,
utils/dataset_indexer.py:634-703, each bin written by MigrationOutputStorage
into ``output_dir/<bin name>`` and finalised with fold normalisation).  The
reference keeps no checkpoint of this loop (processors/migration/
checkpoint.py:221-611 is not wired to it); here the state after every bin is
``output_dir/checkpoint/state.json``, which names the FP32 stack file of the
same generation and is replaced last (the single commit point), so a job
stopped at any point restarts after its last completed bin and produces the
same files as an uninterrupted run.  The job identity covers the parameters,
binning, velocity, the trace geometry and a sample of the trace data.

Data flow per bin (one GPU):
  host gather of the bin's traces (numpy array or memory map, trace-major)
    -> pinned host buffer (two, used alternately)
    -> device, on a copy stream (default priority), recorded by an event
    -> the compute stream waits for the event, builds the bin's plan and
       runs the forward kernel; meanwhile the host gathers and stages the
       next bin.
"""
from __future__ import annotations

import hashlib
import json
import logging
import os
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Callable, Dict, Iterator, List, Optional, Tuple

import numpy as np
import torch

from ...abi_types import FLAG_KAHAN, PstmParams
from ...models.binning import BinningTable

logger = logging.getLogger(__name__)


# ---------------------------------------------------------------------------
# binned trace source

class ArrayBinSource:
    """Binned access to traces held as a (n_traces, n_samples) array (numpy
    or numpy memory map) with FP64 source/receiver coordinates.  Offsets and
    azimuths (degrees clockwise from north, [0, 360)) decide bin membership as
    in BinnedDataset (utils/dataset_indexer.py:634-703)."""

    def __init__(self, traces: np.ndarray, source_x, source_y, receiver_x, receiver_y, binning: BinningTable):
        if traces.ndim != 2:
            raise ValueError("traces must be (n_traces, n_samples)")
        self.traces = traces
        self.sx, self.sy, self.rx, self.ry = (np.asarray(a, dtype=np.float64) for a in
                                              (source_x, source_y, receiver_x, receiver_y))
        if not all(a.shape == (traces.shape[0],) for a in (self.sx, self.sy, self.rx, self.ry)):
            raise ValueError("one coordinate per trace")
        self.binning = binning
        dx, dy = self.rx - self.sx, self.ry - self.sy
        self.offsets = np.hypot(dx, dy)
        self.azimuths = np.mod(np.degrees(np.arctan2(dx, dy)), 360.0)
        masks = binning.assign_traces_batch(self.offsets, self.azimuths)
        self._members = {name: np.nonzero(m)[0] for name, m in masks.items()}

    @property
    def n_samples(self) -> int:
        return int(self.traces.shape[1])

    def bin_summary(self) -> Dict[str, int]:
        return {b.name: int(self._members[b.name].size) for b in self.binning.enabled_bins}

    # BinnedDataset's queries (reference utils/dataset_indexer.py:677-736)
    def get_bin_trace_numbers(self, bin_name: str) -> np.ndarray:
        return self._members.get(bin_name, np.zeros(0, dtype=np.int64))

    def get_bin_trace_count(self, bin_name: str) -> int:
        return int(self.get_bin_trace_numbers(bin_name).size)

    def get_bin_summary(self) -> Dict[str, int]:
        return self.bin_summary()

    def get_coverage_report(self) -> Dict[str, Any]:
        """Traces per bin, assigned in total, in no bin and in several bins."""
        n = int(self.traces.shape[0])
        hits = np.zeros(n, dtype=np.int64)
        for idx in self._members.values():
            hits[idx] += 1
        summary = self.bin_summary()
        unassigned = int(np.count_nonzero(hits == 0))
        return {"n_traces": n, "n_bins": len(summary), "bin_counts": summary, "total_assigned": sum(summary.values()),
                "n_unassigned": unassigned, "n_multi_bin": int(np.count_nonzero(hits > 1)),
                "coverage_percent": 100.0 * (n - unassigned) / n if n else 0.0}

    def identity(self) -> str:
        """Digest of the input: every source/receiver coordinate, the trace
        array's shape and dtype, its file (path, size, mtime) when it is a
        memory map, and a strided sample of its rows (at most 4096 rows,
        each read whole)."""
        h = hashlib.sha256()
        for a in (self.sx, self.sy, self.rx, self.ry):
            h.update(np.ascontiguousarray(a).tobytes())
        h.update(repr((self.traces.shape, str(self.traces.dtype))).encode())
        fn = getattr(self.traces, "filename", None)
        if fn:
            st = os.stat(fn)
            h.update(repr((str(fn), st.st_size, st.st_mtime_ns)).encode())
        n = self.traces.shape[0]
        if n:
            rows = np.unique(np.linspace(0, n - 1, min(n, 4096)).astype(np.int64))
            h.update(np.ascontiguousarray(self.traces[rows]).tobytes())
        return h.hexdigest()

    def iterate_bins(self) -> Iterator[Tuple[str, np.ndarray]]:
        """(bin name, trace numbers) of every enabled, non-empty bin."""
        for b in self.binning.enabled_bins:
            idx = self._members[b.name]
            if idx.size:
                yield b.name, idx

    def read(self, trace_numbers: np.ndarray, out: Optional[np.ndarray] = None) -> np.ndarray:
        """Gather rows (into `out`, e.g. a pinned buffer view)."""
        if out is None:
            return np.ascontiguousarray(self.traces[trace_numbers], dtype=np.float32)
        if self.traces.dtype == np.float32:
            np.take(self.traces, trace_numbers, axis=0, out=out)
        else:
            out[:] = self.traces[trace_numbers]
        return out

    def geometry(self, trace_numbers: np.ndarray) -> np.ndarray:
        """(n, 4) FP64: source x, y, receiver x, y."""
        return np.stack([self.sx[trace_numbers], self.sy[trace_numbers], self.rx[trace_numbers],
                         self.ry[trace_numbers]], 1)


# ---------------------------------------------------------------------------
# staging

@dataclass
class StagedBin:
    name: str
    traces: torch.Tensor     # (n, n_samples) float32 on the device
    geometry: torch.Tensor   # (n, 4) float64 on the device
    ready: Optional[torch.cuda.Event]
    host_seconds: float      # gather into pinned memory


class TraceStager:
    """Two pinned host buffers used alternately and a copy stream: `stage`
    fills a buffer on the host and enqueues its host-to-device copy, which
    runs while the compute stream works on the previous bin."""

    def __init__(self, n_samples: int, device: torch.device):
        self.n_samples = int(n_samples)
        self.device = torch.device(device)
        self.copy_stream = torch.cuda.Stream(device=self.device)
        self._pinned: List[Optional[torch.Tensor]] = [None, None]
        self._pinned_geo: List[Optional[torch.Tensor]] = [None, None]
        self._done: List[Optional[torch.cuda.Event]] = [None, None]
        self._slot = 0

    def stage(self, name: str, source: ArrayBinSource, trace_numbers: np.ndarray) -> StagedBin:
        i = self._slot
        self._slot ^= 1
        if self._done[i] is not None:
            self._done[i].synchronize()  # the buffer's previous copy has finished
        n = int(trace_numbers.size)
        if self._pinned[i] is None or self._pinned[i].shape[0] < n:
            self._pinned[i] = torch.empty((max(n, 1), self.n_samples), dtype=torch.float32, pin_memory=True)
            self._pinned_geo[i] = torch.empty((max(n, 1), 4), dtype=torch.float64, pin_memory=True)
        t0 = time.perf_counter()
        host = self._pinned[i][:n]
        source.read(trace_numbers, out=host.numpy())
        geo = self._pinned_geo[i][:n]
        geo.numpy()[:] = source.geometry(trace_numbers)
        host_s = time.perf_counter() - t0
        with torch.cuda.stream(self.copy_stream):
            d_tr = torch.empty((n, self.n_samples), dtype=torch.float32, device=self.device)
            d_tr.copy_(host, non_blocking=True)
            d_geo = torch.empty((n, 4), dtype=torch.float64, device=self.device)
            d_geo.copy_(geo, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.copy_stream)
        self._done[i] = ev
        return StagedBin(name, d_tr, d_geo, ev, host_s)


# ---------------------------------------------------------------------------
# checkpoint

def _safe_name(name: str) -> str:
    return name.replace(" ", "_").replace("/", "-")


def _atomic_save_npz(path: Path, **arrays: np.ndarray) -> None:
    tmp = path.with_name(path.name + ".tmp")
    with open(tmp, "wb") as fh:
        np.savez(fh, **arrays)
        fh.flush()
        os.fsync(fh.fileno())
    os.replace(tmp, path)


def _atomic_write_text(path: Path, text: str) -> None:
    tmp = path.with_name(path.name + ".tmp")
    with open(tmp, "w") as fh:
        fh.write(text)
        fh.flush()
        os.fsync(fh.fileno())
    os.replace(tmp, path)


@dataclass
class JobResult:
    completed: List[str] = field(default_factory=list)
    skipped: List[str] = field(default_factory=list)
    stack_image: Optional[torch.Tensor] = None  # [il][xl][tau]
    stack_fold: Optional[torch.Tensor] = None
    timings: Dict[str, Dict[str, float]] = field(default_factory=dict)
    seconds: float = 0.0


MigrateFn = Callable[[PstmParams, torch.Tensor, torch.Tensor], Tuple[torch.Tensor, torch.Tensor]]


class MigrationJobRunner:
    """Runs the enabled bins of a job one by one (see the module docstring).

    params:    job parameters; n_traces is set per bin
    velocity:  v_rms, (n_tau,) or (n_il, n_xl, n_tau) float32 on the device
    eta:       anisotropy in the same layout (with PSTM_FLAG_VTI), or None
    trace_dtype: 'fp32', 'bf16' or 'fp16' (quantised on the device after the copy)
    migrate_fn: replaces the HIP migration of one bin, (params, traces, geometry)
               -> (image, fold); for tests of the driver logic.
    """

    STATE_VERSION = 2

    def __init__(self, params: PstmParams, velocity: torch.Tensor, output_dir, *, eta: Optional[torch.Tensor] = None,
                 trace_dtype: str = "fp32", write_bins: bool = True, normalize: bool = True, stack: bool = True,
                 checkpoint: bool = True, device: Optional[torch.device] = None,
                 migrate_fn: Optional[MigrateFn] = None, grid_info: Optional[Dict] = None):
        self.params = params
        self.velocity = velocity
        self.eta = eta
        self.trace_dtype = trace_dtype
        self.output_dir = Path(output_dir)
        self.write_bins, self.normalize, self.stack, self.checkpoint = write_bins, normalize, stack, checkpoint
        self.migrate_fn = migrate_fn
        self.grid_info = dict(grid_info or {})
        if device is None:
            device = velocity.device
        self.device = torch.device(device)
        if migrate_fn is None and self.device.type != "cuda":
            raise RuntimeError("the migration job runs the HIP kernels: it needs a ROCm device")
        self.ckpt_dir = self.output_dir / "checkpoint"

    # -- job identity: a checkpoint only resumes the same job
    def _job_key(self, source: ArrayBinSource) -> str:
        h = hashlib.sha256()
        h.update(json.dumps(self.params.to_dict(), sort_keys=True).encode())
        h.update(json.dumps(source.binning.to_dict(), sort_keys=True).encode())
        h.update(json.dumps(source.bin_summary(), sort_keys=True).encode())
        h.update(repr((self.trace_dtype, bool(self.stack))).encode())
        h.update(np.ascontiguousarray(self.velocity.detach().float().cpu().numpy()).tobytes())
        if self.eta is not None:
            h.update(np.ascontiguousarray(self.eta.detach().float().cpu().numpy()).tobytes())
        # the input itself: the trace geometry and an identity of the trace data
        h.update(source.identity().encode())
        return h.hexdigest()

    def _load_state(self, key: str):
        f = self.ckpt_dir / "state.json"
        if not f.exists():
            return [], None, None
        st = json.loads(f.read_text())
        if st.get("version") != self.STATE_VERSION or st.get("job_key") != key:
            raise RuntimeError(f"{f} belongs to a different job; remove the checkpoint directory to restart")
        img = fold = None
        if self.stack and st["completed"]:
            with np.load(self.ckpt_dir / st["stack_file"], allow_pickle=False) as z:
                img = torch.from_numpy(z["image"]).to(self.device)
                fold = torch.from_numpy(z["fold"]).to(self.device)
        return list(st["completed"]), img, fold

    def _save_state(self, key: str, completed: List[str], img: Optional[torch.Tensor], fold: Optional[torch.Tensor]):
        """``state.json`` is the single commit point: the stack of this
        generation goes to its own file (``stack.<n completed>.npz``, image and
        fold together), then the state naming that file replaces the old one,
        then older generations are removed.  A stop at any point leaves a
        state whose stack file exists and holds exactly its completed bins."""
        self.ckpt_dir.mkdir(parents=True, exist_ok=True)
        stack_file = None
        if self.stack and img is not None:
            stack_file = f"stack.{len(completed)}.npz"
            _atomic_save_npz(self.ckpt_dir / stack_file, image=img.cpu().numpy(), fold=fold.cpu().numpy())
        _atomic_write_text(self.ckpt_dir / "state.json",
                           json.dumps({"version": self.STATE_VERSION, "job_key": key, "completed": completed,
                                       "stack_file": stack_file}, indent=2))
        for old in self.ckpt_dir.glob("stack.*.npz"):
            if old.name != stack_file:
                old.unlink()

    # -- one bin on the GPU
    def _migrate_hip(self, p: PstmParams, traces: torch.Tensor, geo: torch.Tensor):
        from .hip_ops import PstmPlan, quantize_traces
        plan = PstmPlan(p, geo[:, 0:2].contiguous(), geo[:, 2:4].contiguous(), device=self.device)
        scale = None
        if self.trace_dtype in ("bf16", "fp16"):
            traces, scale = quantize_traces(traces, torch.bfloat16 if self.trace_dtype == "bf16" else torch.float16)
        comp = torch.zeros(plan.image_shape, dtype=torch.float32, device=self.device) if p.flags & FLAG_KAHAN else None
        img, fold = plan.forward(traces, self.velocity, scale=scale, eta=self.eta, comp=comp)
        if comp is not None:
            img += comp
        return img, fold

    def _write_bin(self, name: str, img: torch.Tensor, fold: torch.Tensor, extra: Dict):
        from ...utils.migration_output_storage import MigrationOutputStorage
        p = self.params
        g = self.grid_info
        st = MigrationOutputStorage(str(self.output_dir / _safe_name(name)))
        st.initialize_output(n_time=p.n_tau, n_inline=p.n_il, n_xline=p.n_xl, dt_ms=p.dtau * 1000.0,
                             t0_ms=p.tau0 * 1000.0, d_inline=g.get("d_inline", float(np.hypot(p.il_dx, p.il_dy))),
                             d_xline=g.get("d_xline", float(np.hypot(p.xl_dx, p.xl_dy))),
                             inline_start=g.get("inline_start", 1), xline_start=g.get("xline_start", 1),
                             x_origin=p.origin_x, y_origin=p.origin_y,
                             grid_azimuth_deg=g.get("grid_azimuth_deg", 0.0), source_file=g.get("source_file"))
        st.write_device_cube(img, fold)
        st.finalize(normalize=self.normalize, extra_metadata=extra)

    # -- the job
    def run(self, source: ArrayBinSource, resume: bool = True,
            progress_callback: Optional[Callable[[int, int, str], None]] = None) -> JobResult:
        """Migrate every enabled bin not yet completed.  progress_callback
        (current, total, message) runs after each completed bin (the
        reference's processor progress signature); an exception it raises
        stops the job after that bin's checkpoint."""
        t_job = time.perf_counter()
        self.output_dir.mkdir(parents=True, exist_ok=True)
        key = self._job_key(source)
        completed, st_img, st_fold = self._load_state(key) if resume else ([], None, None)
        if not resume and self.ckpt_dir.exists():
            for f in self.ckpt_dir.iterdir():
                f.unlink()
        res = JobResult(skipped=list(completed))
        todo = [(n, idx) for n, idx in source.iterate_bins() if n not in completed]
        total = len(completed) + len(todo)
        if self.stack and st_img is None:
            shape = (self.params.n_il, self.params.n_xl, self.params.n_tau)
            st_img = torch.zeros(shape, dtype=torch.float32, device=self.device)
            st_fold = torch.zeros(shape, dtype=torch.float32, device=self.device)
        use_stager = self.device.type == "cuda"
        stager = TraceStager(source.n_samples, self.device) if use_stager else None

        def stage(k):
            name, idx = todo[k]
            if stager is not None:
                return stager.stage(name, source, idx)
            t0 = time.perf_counter()
            tr = torch.from_numpy(source.read(idx))
            geo = torch.from_numpy(source.geometry(idx))
            return StagedBin(name, tr, geo, None, time.perf_counter() - t0)

        nxt = stage(0) if todo else None
        for k in range(len(todo)):
            cur = nxt
            name, idx = todo[k]
            t0 = time.perf_counter()
            p = self.params.replace(n_traces=int(idx.size), n_samples=source.n_samples)
            if cur.ready is not None:
                cs = torch.cuda.current_stream(self.device)
                cs.wait_event(cur.ready)
                cur.traces.record_stream(cs)
                cur.geometry.record_stream(cs)
            fn = self.migrate_fn or self._migrate_hip
            ev = None
            if self.device.type == "cuda":
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                ev[0].record(torch.cuda.current_stream(self.device))
            img, fold = fn(p, cur.traces, cur.geometry)
            if self.stack:
                st_img += img
                st_fold += fold
            if ev is not None:
                ev[1].record(torch.cuda.current_stream(self.device))
            # the next bin's gather and copy overlap this bin's migration
            nxt = stage(k + 1) if k + 1 < len(todo) else None
            if self.device.type == "cuda":
                torch.cuda.synchronize(self.device)
            t_mig = time.perf_counter() - t0
            gpu_s = ev[0].elapsed_time(ev[1]) / 1e3 if ev is not None else None
            extra = {"bin_name": name, "input_traces": int(idx.size), "max_aperture_m": p.aperture_m,
                     "max_angle_deg": p.max_angle_deg, "trace_dtype": self.trace_dtype}
            t1 = time.perf_counter()
            if self.write_bins:
                self._write_bin(name, img, fold, extra)
            completed.append(name)
            if self.checkpoint:
                self._save_state(key, completed, st_img if self.stack else None, st_fold if self.stack else None)
            res.completed.append(name)
            res.timings[name] = {"stage_host_s": cur.host_seconds, "migrate_s": t_mig, "gpu_s": gpu_s,
                                 "write_s": time.perf_counter() - t1, "traces": int(idx.size)}
            del img, fold, cur
            if progress_callback:
                progress_callback(len(completed), total, f"bin {name} done")
        res.stack_image, res.stack_fold = (st_img, st_fold) if self.stack else (None, None)
        res.seconds = time.perf_counter() - t_job
        return res
