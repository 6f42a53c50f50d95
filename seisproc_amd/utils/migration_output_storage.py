"""Migration output in the reference's on-disk format.

Mirrors ``MigrationOutputStorage`` of utils/migration_output_storage.py:
30-480 (same directory layout, file names, array shapes, Parquet columns and
metadata keys), so a migrated image written here loads in SeisProc like one
of its own:

    output_dir/
        traces.zarr/            image, (n_time, n_traces) float32, il-major
                                traces (trace = il * n_xline + xl), Zarr v2,
                                uncompressed, chunks (n_time, chunk_size)
        fold.npy                fold, (n_time, n_traces) float32
        headers.parquet         trace_index, inline, xline, cdp_x, cdp_y,
                                ensemble_number, trace_in_ensemble
        ensemble_index.parquet  one row per inline
        trace_index.parquet     trace_index, global_trace_id
        metadata.json           grid, processing stats, storage info

Differences from the reference, in behaviour a caller can see: none in the
files; ``write_device_cube`` additionally takes the engine's device layout
([il][xl][tau] torch tensors) and streams it to disk in trace chunks without
building the (n_time, n_il, n_xl) host cube first.  The Zarr store is written
by seisproc_amd.utils.zarr_v2 (the ``zarr`` package is not required).
"""
from __future__ import annotations

import gc
import json
import logging
import math
from datetime import datetime
from pathlib import Path
from typing import Dict, Optional

import numpy as np

from .zarr_v2 import ZarrV2Array

logger = logging.getLogger(__name__)


class MigrationOutputStorage:
    """Writes a migrated cube as traces.zarr + fold.npy + Parquet headers."""

    def __init__(self, output_dir: str):
        self.output_dir = Path(output_dir)
        self.output_dir.mkdir(parents=True, exist_ok=True)
        self.traces_path = self.output_dir / "traces.zarr"
        self.fold_path = self.output_dir / "fold.npy"
        self.headers_path = self.output_dir / "headers.parquet"
        self.ensemble_index_path = self.output_dir / "ensemble_index.parquet"
        self.trace_index_path = self.output_dir / "trace_index.parquet"
        self.metadata_path = self.output_dir / "metadata.json"
        self._source_file = None
        self._n_time = self._n_inline = self._n_xline = self._n_traces = 0
        self._dt_ms, self._t0_ms = 4.0, 0.0
        self._d_inline = self._d_xline = 25.0
        self._inline_start = self._xline_start = 1
        self._x_origin = self._y_origin = 0.0
        self._grid_azimuth_deg = 0.0
        self._chunk = 1000
        self._traces_zarr: Optional[ZarrV2Array] = None
        self._fold_array: Optional[np.ndarray] = None
        self._stats = {"inlines_written": 0, "traces_written": 0, "max_fold": 0, "start_time": None,
                       "end_time": None}

    # ------------------------------------------------------------------ set-up
    def initialize_output(self, n_time: int, n_inline: int, n_xline: int, dt_ms: float = 4.0, t0_ms: float = 0.0,
                          d_inline: float = 25.0, d_xline: float = 25.0, inline_start: int = 1,
                          xline_start: int = 1, x_origin: float = 0.0, y_origin: float = 0.0,
                          grid_azimuth_deg: float = 0.0, chunk_size: int = 1000,
                          source_file: Optional[str] = None):
        """Create the image store (zero-filled) and the in-memory fold."""
        self._n_time, self._n_inline, self._n_xline = int(n_time), int(n_inline), int(n_xline)
        self._n_traces = self._n_inline * self._n_xline
        self._dt_ms, self._t0_ms = float(dt_ms), float(t0_ms)
        self._d_inline, self._d_xline = float(d_inline), float(d_xline)
        self._inline_start, self._xline_start = int(inline_start), int(xline_start)
        self._x_origin, self._y_origin = float(x_origin), float(y_origin)
        self._grid_azimuth_deg = float(grid_azimuth_deg)
        self._source_file = source_file
        self._chunk = max(1, min(int(chunk_size), self._n_traces))
        self._stats["start_time"] = datetime.now().isoformat()
        logger.info("Initializing migration output: %dx%dx%d = %d traces", n_inline, n_xline, n_time,
                    self._n_traces)
        self._traces_zarr = ZarrV2Array.create(self.traces_path, (self._n_time, self._n_traces),
                                               (self._n_time, self._chunk))
        self._fold_array = np.zeros((self._n_time, self._n_traces), dtype=np.float32)

    def _require(self):
        if self._traces_zarr is None:
            raise RuntimeError("Output not initialized. Call initialize_output first.")

    # ------------------------------------------------------------------ writes
    def _write_traces(self, t0: int, image_2d: np.ndarray):
        """Image columns [t0, t0 + n) (n_time, n); read-modify-write of the
        chunks the range touches."""
        z, cw = self._traces_zarr, self._chunk
        t1 = t0 + image_2d.shape[1]
        for c in range(t0 // cw, (t1 - 1) // cw + 1):
            c0, c1 = c * cw, min((c + 1) * cw, self._n_traces)
            if t0 <= c0 and c1 <= t1:
                blk = image_2d[:, c0 - t0:c1 - t0]
            else:
                blk = z.read_chunk((0, c))[:, :c1 - c0].copy()
                a, b = max(c0, t0), min(c1, t1)
                blk[:, a - c0:b - c0] = image_2d[:, a - t0:b - t0]
            z.write_chunk((0, c), np.ascontiguousarray(blk, dtype=np.float32))

    def write_inline(self, inline_idx: int, image_data: np.ndarray, fold_data: np.ndarray):
        """One inline: image and fold (n_time, n_xline)."""
        self._require()
        s = inline_idx * self._n_xline
        self._write_traces(s, np.asarray(image_data, dtype=np.float32))
        self._fold_array[:, s:s + self._n_xline] = fold_data
        self._stats["inlines_written"] += 1
        self._stats["traces_written"] += self._n_xline
        self._stats["max_fold"] = max(self._stats["max_fold"], float(np.max(fold_data)))

    def write_full_cube(self, image_cube: np.ndarray, fold_cube: np.ndarray):
        """Whole cube, (n_time, n_inline, n_xline) each."""
        self._require()
        n_time, n_inline, n_xline = image_cube.shape
        if (n_time, n_inline, n_xline) != (self._n_time, self._n_inline, self._n_xline):
            raise ValueError(f"Shape mismatch: expected ({self._n_time}, {self._n_inline}, {self._n_xline}), "
                             f"got ({n_time}, {n_inline}, {n_xline})")
        self._write_traces(0, np.asarray(image_cube, dtype=np.float32).reshape(n_time, -1))
        self._fold_array[:] = np.asarray(fold_cube).reshape(n_time, -1)
        self._stats["inlines_written"] = n_inline
        self._stats["traces_written"] = n_inline * n_xline
        self._stats["max_fold"] = float(np.max(fold_cube))

    def write_device_cube(self, image, fold=None, traces_per_copy: int = 65536):
        """Whole cube from the engine's device layout: torch tensors
        [il][xl][tau] (image and fold).  Copies chunk-aligned trace ranges to
        the host and transposes them to (n_time, n) on the way."""
        self._require()
        if tuple(image.shape) != (self._n_inline, self._n_xline, self._n_time):
            raise ValueError(f"expected device image ({self._n_inline}, {self._n_xline}, {self._n_time}), "
                             f"got {tuple(image.shape)}")
        flat = image.reshape(self._n_traces, self._n_time)
        fflat = fold.reshape(self._n_traces, self._n_time) if fold is not None else None
        step = max(self._chunk, (traces_per_copy // self._chunk) * self._chunk)
        mx = 0.0
        for t0 in range(0, self._n_traces, step):
            t1 = min(t0 + step, self._n_traces)
            self._write_traces(t0, flat[t0:t1].float().cpu().numpy().T)
            if fflat is not None:
                fb = fflat[t0:t1].float().cpu().numpy()
                self._fold_array[:, t0:t1] = fb.T
                mx = max(mx, float(fb.max()) if fb.size else 0.0)
        self._stats["inlines_written"] = self._n_inline
        self._stats["traces_written"] = self._n_traces
        self._stats["max_fold"] = mx

    # ------------------------------------------------------------------ tables
    def _build_headers(self):
        import pandas as pd
        az = math.radians(self._grid_azimuth_deg)
        il = np.repeat(np.arange(self._n_inline), self._n_xline)
        xl = np.tile(np.arange(self._n_xline), self._n_inline)
        di, dx = il * self._d_inline, xl * self._d_xline
        return pd.DataFrame({
            "trace_index": np.arange(self._n_traces, dtype=np.int64),
            "inline": (self._inline_start + il).astype(np.int64),
            "xline": (self._xline_start + xl).astype(np.int64),
            # inline direction at azimuth from north, crossline at azimuth + 90 deg
            "cdp_x": self._x_origin + di * math.sin(az) + dx * math.cos(az),
            "cdp_y": self._y_origin + di * math.cos(az) - dx * math.sin(az),
            "ensemble_number": (self._inline_start + il).astype(np.int64),
            "trace_in_ensemble": xl.astype(np.int64),
        })

    def _build_ensemble_index(self):
        import pandas as pd
        il = np.arange(self._n_inline, dtype=np.int64)
        return pd.DataFrame({
            "ensemble_id": il,
            "ensemble_value": self._inline_start + il,
            "inline": self._inline_start + il,
            "start_trace": il * self._n_xline,
            "end_trace": il * self._n_xline + self._n_xline - 1,
            "n_traces": np.full(self._n_inline, self._n_xline, dtype=np.int64),
        })

    def _build_metadata(self, extra_metadata: Optional[Dict] = None) -> Dict:
        self._stats["end_time"] = datetime.now().isoformat()
        md = {
            "shape": [self._n_time, self._n_traces],
            "sample_rate": self._dt_ms,
            "n_samples": self._n_time,
            "n_traces": self._n_traces,
            "duration_ms": (self._n_time - 1) * self._dt_ms,
            "nyquist_freq": 1000.0 / (2 * self._dt_ms),
            "data_type": "migration",
            "migration_info": {
                "n_inline": self._n_inline, "n_xline": self._n_xline, "n_time": self._n_time,
                "dt_ms": self._dt_ms, "t0_ms": self._t0_ms, "d_inline": self._d_inline, "d_xline": self._d_xline,
                "inline_start": self._inline_start, "xline_start": self._xline_start,
                "inline_end": self._inline_start + self._n_inline - 1,
                "xline_end": self._xline_start + self._n_xline - 1,
                "x_origin": self._x_origin, "y_origin": self._y_origin,
                "grid_azimuth_deg": self._grid_azimuth_deg,
            },
            "processing_stats": dict(self._stats),
            "storage_info": {
                "zarr_chunks": f"({self._n_time}, chunk_size)",
                "parquet_compression": "snappy",
                "zarr_compression": "none",
                "fold_available": True,
            },
            "seismic_metadata": {
                "source_file": self._source_file or str(self.output_dir),
                "source_type": "migration",
                "processing_history": ["PSTM Kirchhoff Migration"],
                "file_info": {
                    "filename": self._source_file or "PSTM Migration Output",
                    "n_traces": self._n_traces,
                    "n_samples": self._n_time,
                    "sample_interval": self._dt_ms,
                    "trace_length_ms": (self._n_time - 1) * self._dt_ms,
                },
            },
        }
        if extra_metadata:
            md["migration_params"] = extra_metadata
        return md

    # ------------------------------------------------------------------ finish
    def finalize(self, normalize: bool = True, extra_metadata: Optional[Dict] = None) -> Dict:
        """Optionally divide the image by fold (fold == 0 -> 0, the
        reference's :410-420), then write fold, headers, indexes, metadata."""
        import pandas as pd
        self._require()
        if normalize:
            z, cw = self._traces_zarr, self._chunk
            for c in range(z.n_chunks()[1]):
                c0, c1 = c * cw, min((c + 1) * cw, self._n_traces)
                img = z.read_chunk((0, c))[:, :c1 - c0]
                fold = self._fold_array[:, c0:c1]
                with np.errstate(divide="ignore", invalid="ignore"):
                    norm = np.where(fold > 0, img / fold, 0.0).astype(np.float32)
                z.write_chunk((0, c), norm)
            gc.collect()
        np.save(self.fold_path, self._fold_array)
        self._build_headers().to_parquet(self.headers_path, engine="pyarrow", compression="snappy", index=False)
        self._build_ensemble_index().to_parquet(self.ensemble_index_path, engine="pyarrow", compression="snappy",
                                                index=False)
        pd.DataFrame({"trace_index": np.arange(self._n_traces), "global_trace_id": np.arange(self._n_traces)}) \
            .to_parquet(self.trace_index_path, engine="pyarrow", compression="snappy", index=False)
        md = self._build_metadata(extra_metadata)
        with open(self.metadata_path, "w") as fh:
            json.dump(md, fh, indent=2)
        logger.info("Migration output saved to %s (%dx%dx%d, max fold %.0f)", self.output_dir, self._n_inline,
                    self._n_xline, self._n_time, self._stats["max_fold"])
        return md

    def get_output_path(self) -> Path:
        return self.output_dir


def create_migration_output_storage(output_dir: str, n_time: int, n_inline: int, n_xline: int, dt_ms: float = 4.0,
                                    **kwargs) -> MigrationOutputStorage:
    """Create and initialize a storage (reference factory of the same name)."""
    s = MigrationOutputStorage(output_dir)
    s.initialize_output(n_time=n_time, n_inline=n_inline, n_xline=n_xline, dt_ms=dt_ms, **kwargs)
    return s
