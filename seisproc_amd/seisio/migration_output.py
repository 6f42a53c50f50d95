"""Per-bin, stack and fold output volumes of a migration job (reference
seisio/migration_output.py:24-371).

``MigrationOutputManager`` holds one (n_time, n_inline, n_xline) FP32 volume
per enabled bin of a binning table, optionally their int32 fold volumes and a
full stack with its fold, accumulates migrated bins into them, normalises by
fold and saves each volume as ``<base_name>_<volume>.npy``, a Zarr v2 array
(``.zarr``, one chunk per inline) or SEG-Y (``.sgy``, IEEE float samples with
inline/crossline and CDP coordinates in the trace headers); fold volumes go
beside them with a ``_fold`` suffix.  Migrated images come either in that
layout (``add_migrated_data``) or as the kernels' device tensors
([il][xl][tau], ``add_device_volume``).
"""
from __future__ import annotations

import logging
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Dict, Optional

import numpy as np

from ..models.binning import BinningTable
from ..models.migration_config import OutputGrid

logger = logging.getLogger(__name__)

_SUFFIX = {"npy": ".npy", "zarr": ".zarr", "segy": ".sgy"}


@dataclass
class OutputVolumeInfo:
    name: str
    filepath: str
    bin_name: Optional[str] = None  # None for the stack
    offset_center: Optional[float] = None
    azimuth_center: Optional[float] = None
    n_traces_migrated: int = 0  # additions to the volume
    is_complete: bool = False


@dataclass
class MigrationOutputManager:
    output_directory: str
    output_grid: OutputGrid
    binning_table: BinningTable
    base_name: str = "migrated"
    output_format: str = "npy"  # 'npy', 'zarr' or 'segy'
    create_stack: bool = True
    create_fold: bool = False
    compress: bool = False  # kept for the interface: every format here is written uncompressed

    _volumes: Dict[str, np.ndarray] = field(default_factory=dict, repr=False)
    _fold_volumes: Dict[str, np.ndarray] = field(default_factory=dict, repr=False)
    _stack_volume: Optional[np.ndarray] = field(default=None, repr=False)
    _stack_fold: Optional[np.ndarray] = field(default=None, repr=False)
    _volume_info: Dict[str, OutputVolumeInfo] = field(default_factory=dict)
    _initialized: bool = False

    def __post_init__(self):
        if self.output_format not in _SUFFIX:
            raise ValueError(f"Unknown output format: {self.output_format}")
        Path(self.output_directory).mkdir(parents=True, exist_ok=True)

    @property
    def shape(self):
        g = self.output_grid
        return (g.n_time, g.n_inline, g.n_xline)

    def _path(self, name: str) -> str:
        return str(Path(self.output_directory) / f"{self.base_name}_{name}{_SUFFIX[self.output_format]}")

    def initialize(self) -> None:
        """Allocate the volumes: one per enabled bin (and its fold with
        create_fold), and the stack with its fold (with create_stack)."""
        if self._initialized:
            logger.warning("Output manager already initialized")
            return
        for b in self.binning_table.enabled_bins:
            self._volumes[b.name] = np.zeros(self.shape, dtype=np.float32)
            if self.create_fold:
                self._fold_volumes[b.name] = np.zeros(self.shape, dtype=np.int32)
            self._volume_info[b.name] = OutputVolumeInfo(
                b.name, self._path(b.name), b.name, 0.5 * (b.offset_min + b.offset_max),
                0.5 * (b.azimuth_min + b.azimuth_max))
        if self.create_stack:
            self._stack_volume = np.zeros(self.shape, dtype=np.float32)
            self._stack_fold = np.zeros(self.shape, dtype=np.int32)
            self._volume_info["stack"] = OutputVolumeInfo("stack", self._path("stack"))
        self._initialized = True

    def _require(self, bin_name: str) -> None:
        if not self._initialized:
            raise RuntimeError("Output manager not initialized")
        if bin_name not in self._volumes:
            raise ValueError(f"Unknown bin: {bin_name}")

    def add_migrated_data(self, bin_name: str, data: np.ndarray, fold: Optional[np.ndarray] = None) -> None:
        """Add a migrated image (n_time, n_inline, n_xline) of the bin, and
        its fold, to the bin's volumes and the stack."""
        self._require(bin_name)
        data = np.asarray(data, dtype=np.float32)
        if data.shape != self.shape:
            raise ValueError(f"expected data of shape {self.shape}, got {data.shape}")
        self._volumes[bin_name] += data
        if fold is not None:
            fold = np.rint(np.asarray(fold)).astype(np.int32)
            if self.create_fold:
                self._fold_volumes[bin_name] += fold
        if self.create_stack:
            self._stack_volume += data
            if fold is not None:
                self._stack_fold += fold
        self._volume_info[bin_name].n_traces_migrated += 1

    def add_device_volume(self, bin_name: str, image, fold=None) -> None:
        """add_migrated_data for the kernels' [il][xl][tau] device tensors."""
        to_host = lambda t: t.detach().permute(2, 0, 1).cpu().numpy()
        self.add_migrated_data(bin_name, to_host(image), None if fold is None else to_host(fold))

    def accumulate_to_bin(self, bin_name: str, time_idx: int, inline_idx: int, xline_idx: int,
                          value: float) -> None:
        """Add one sample value (and one unit of fold) to the bin and the stack."""
        if not self._initialized:
            raise RuntimeError("Output manager not initialized")
        if bin_name not in self._volumes:
            return
        at = (time_idx, inline_idx, xline_idx)
        self._volumes[bin_name][at] += value
        if self.create_fold:
            self._fold_volumes[bin_name][at] += 1
        if self.create_stack:
            self._stack_volume[at] += value
            self._stack_fold[at] += 1

    def get_volume(self, bin_name: str) -> Optional[np.ndarray]:
        return self._stack_volume if bin_name == "stack" else self._volumes.get(bin_name)

    def get_fold(self, bin_name: str) -> Optional[np.ndarray]:
        return self._stack_fold if bin_name == "stack" else self._fold_volumes.get(bin_name)

    @staticmethod
    def _normalize(volume: np.ndarray, fold: np.ndarray, min_fold: int) -> None:
        keep = fold >= min_fold
        volume[keep] /= fold[keep]
        volume[~keep] = 0.0

    def normalize_by_fold(self, min_fold: int = 1) -> None:
        """Divide every volume that has a fold by it; samples whose fold is
        below min_fold become 0."""
        for name, vol in self._volumes.items():
            if name in self._fold_volumes:
                self._normalize(vol, self._fold_volumes[name], min_fold)
        if self.create_stack and self._stack_fold is not None:
            self._normalize(self._stack_volume, self._stack_fold, min_fold)

    def finalize_bin(self, bin_name: str) -> None:
        if bin_name in self._volume_info:
            self._volume_info[bin_name].is_complete = True

    # -- files
    @staticmethod
    def _fold_path(path: str) -> str:
        p = Path(path)
        return str(p.with_name(p.stem + "_fold" + p.suffix))

    def _save_volume(self, volume: np.ndarray, path: str) -> None:
        if self.output_format == "npy":
            np.save(path, volume)
        elif self.output_format == "zarr":
            from ..utils.zarr_v2 import ZarrV2Array
            nt, nil, nxl = volume.shape
            z = ZarrV2Array.create(path, volume.shape, (nt, 1, nxl), dtype=volume.dtype.newbyteorder("<").str)
            for i in range(nil):
                z.write_chunk((0, i, 0), volume[:, i:i + 1, :])
        else:
            from ..utils.segy import write_segy
            g = self.output_grid
            il, xl = np.meshgrid(np.arange(g.n_inline), np.arange(g.n_xline), indexing="ij")
            xy = np.array([g.get_coordinates(int(i), int(j)) for i, j in zip(il.ravel(), xl.ravel())])
            traces = volume.reshape(volume.shape[0], -1).T  # one trace per (inline, crossline)
            write_segy(path, traces, g.dt, format_code=5 if volume.dtype.kind == "f" else 2,
                       inline=g.inline_start + il.ravel(), crossline=g.xline_start + xl.ravel(),
                       cdp=np.arange(1, traces.shape[0] + 1), cdp_x=np.rint(xy[:, 0]), cdp_y=np.rint(xy[:, 1]),
                       delay_ms=int(round(g.t0 * 1000.0)))

    def save_bin(self, bin_name: str) -> None:
        if bin_name not in self._volumes:
            raise ValueError(f"Unknown bin: {bin_name}")
        self._save_volume(self._volumes[bin_name], self._volume_info[bin_name].filepath)

    def save_all(self) -> None:
        """Write every volume (and fold volume with create_fold)."""
        for name, vol in self._volumes.items():
            path = self._volume_info[name].filepath
            self._save_volume(vol, path)
            if self.create_fold and name in self._fold_volumes:
                self._save_volume(self._fold_volumes[name], self._fold_path(path))
        if self.create_stack and self._stack_volume is not None:
            path = self._volume_info["stack"].filepath
            self._save_volume(self._stack_volume, path)
            if self.create_fold:
                self._save_volume(self._stack_fold, self._fold_path(path))

    def get_summary(self) -> Dict[str, Any]:
        return {"output_directory": self.output_directory, "n_bins": len(self._volumes),
                "output_format": self.output_format, "create_stack": self.create_stack,
                "create_fold": self.create_fold,
                "volumes": {n: {k: getattr(i, k) for k in ("filepath", "bin_name", "offset_center", "azimuth_center",
                                                           "n_traces_migrated", "is_complete")}
                            for n, i in self._volume_info.items()}}

    def get_output_headers(self, bin_name: str, inline_idx: int, xline_idx: int) -> Dict[str, Any]:
        """Trace header values of an output position: inline and crossline
        numbers, CDP world coordinates and the bin's offset/azimuth centres."""
        g = self.output_grid
        x, y = g.get_coordinates(inline_idx, xline_idx)
        h = {"INLINE_3D": g.inline_start + inline_idx, "CROSSLINE_3D": g.xline_start + xline_idx,
             "CDP_X": x, "CDP_Y": y}
        info = self._volume_info.get(bin_name)
        if info is not None:
            if info.offset_center is not None:
                h["OFFSET"] = info.offset_center
            if info.azimuth_center is not None:
                h["AZIMUTH"] = info.azimuth_center
        return h


def create_output_manager(job_config) -> MigrationOutputManager:
    """The manager of a MigrationJobConfig: its directory, grid, binning,
    name (lower case, underscores) and output options."""
    return MigrationOutputManager(
        output_directory=job_config.output_directory, output_grid=job_config.get_output_grid(),
        binning_table=job_config.get_binning_table(), base_name=job_config.name.replace(" ", "_").lower(),
        output_format=job_config.output_format, create_stack=job_config.create_stack_volume, create_fold=job_config.create_fold_volume,
        compress=job_config.compress_output)
