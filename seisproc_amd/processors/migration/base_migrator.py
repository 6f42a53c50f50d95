"""BaseMigrator and MigrationResult (reference processors/migration/base_migrator.py:24-386).

Same abstract interface (migrate_gather / migrate_dataset / estimate_memory_gb /
get_description), progress callbacks, memory helpers, status and
serialization, and the same MigrationResult container with
apply_fold_normalization semantics (fold clipped below at min_fold for the
division, then zeroed where fold < min_fold, :317-349).
"""
from __future__ import annotations

import logging
from abc import ABC, abstractmethod
from typing import Any, Callable, Dict, List, Optional

import numpy as np

from ...models.migration_config import MigrationConfig

logger = logging.getLogger(__name__)
ProgressCallback = Callable[[int, int, str], None]


class BaseMigrator(ABC):
    def __init__(self, velocity_model, config: MigrationConfig, **kwargs):
        self.velocity_model = velocity_model
        self.config = config
        # a migrator that is also a processor has set its processor params already
        self.params = {**getattr(self, "params", {}), **kwargs}
        self._progress_callback: Optional[ProgressCallback] = None
        self._validate_inputs()
        logger.info("Initialized %s", self.__class__.__name__)

    @abstractmethod
    def _validate_inputs(self): ...

    def set_progress_callback(self, callback: Optional[ProgressCallback]) -> "BaseMigrator":
        self._progress_callback = callback
        return self

    def _report_progress(self, current: int, total: int, message: str = ""):
        if self._progress_callback is not None:
            self._progress_callback(current, total, message)

    @abstractmethod
    def migrate_gather(self, gather, geometry): ...

    @abstractmethod
    def migrate_dataset(self, gathers: List, geometries: List, batch_size: int = 1) -> "MigrationResult": ...

    @abstractmethod
    def estimate_memory_gb(self, n_traces: int, n_samples: int) -> float: ...

    def can_process(self, n_traces: int, n_samples: int, available_memory_gb: float) -> bool:
        return self.estimate_memory_gb(n_traces, n_samples) < available_memory_gb * 0.8

    def calculate_batch_size(self, n_traces_per_gather: int, n_samples: int, available_memory_gb: float) -> int:
        per = self.estimate_memory_gb(n_traces_per_gather, n_samples)
        if per <= 0:
            return 1
        return max(1, int(available_memory_gb * 0.7 / per))

    @abstractmethod
    def get_description(self) -> str: ...

    def get_status(self) -> Dict[str, Any]:
        g = self.config.output_grid
        vt = getattr(self.velocity_model, "velocity_type", None)
        return {"class": self.__class__.__name__, "velocity_type": str(vt),
                "output_grid": f"{g.n_inline}x{g.n_xline}x{g.n_time}",
                "max_aperture": self.config.max_aperture_m, "max_angle": self.config.max_angle_deg}

    def to_dict(self) -> Dict[str, Any]:
        """What rebuilds this migrator: class, module, velocity model,
        configuration and keyword parameters."""
        vm = self.velocity_model
        return {"class_name": self.__class__.__name__, "module": self.__class__.__module__,
                "velocity_model": vm.to_dict() if hasattr(vm, "to_dict") else vm,
                "config": self.config.to_dict(), "params": dict(self.params)}

    def __repr__(self) -> str:
        return f"<{self.__class__.__name__}: {self.velocity_model!r}, aperture {self.config.max_aperture_m:g} m>"


class MigrationResult:
    """Migrated image (n_time, n_inline, n_xline), fold, config, metadata."""

    def __init__(self, image: np.ndarray, fold: Optional[np.ndarray] = None,
                 config: Optional[MigrationConfig] = None, metadata: Optional[Dict[str, Any]] = None):
        self.image = np.asarray(image).astype(np.float32)
        self.fold = np.asarray(fold).astype(np.int32) if fold is not None else None
        self.config = config
        self.metadata = metadata or {}

    @property
    def shape(self):
        return self.image.shape

    @property
    def n_time(self) -> int:
        return self.image.shape[0]

    @property
    def n_inline(self) -> int:
        return self.image.shape[1]

    @property
    def n_xline(self) -> int:
        return self.image.shape[2]

    def get_inline_section(self, xline_idx: int) -> np.ndarray:
        return self.image[:, :, xline_idx]

    def get_xline_section(self, inline_idx: int) -> np.ndarray:
        return self.image[:, inline_idx, :]

    def get_time_slice(self, time_idx: int) -> np.ndarray:
        return self.image[time_idx, :, :]

    def apply_fold_normalization(self, min_fold: int = 1) -> "MigrationResult":
        if self.fold is None:
            return self
        img = self.image / np.maximum(self.fold, min_fold)
        img[self.fold < min_fold] = 0.0
        return MigrationResult(img, self.fold.copy(), self.config,
                               {**self.metadata, "fold_normalized": True, "min_fold": min_fold})

    def get_statistics(self) -> Dict[str, Any]:
        f = self.fold
        return {"shape": self.shape, "image_min": float(np.min(self.image)), "image_max": float(np.max(self.image)),
                "image_mean": float(np.mean(self.image)), "image_std": float(np.std(self.image)),
                "image_rms": float(np.sqrt(np.mean(self.image.astype(np.float64) ** 2))),
                "fold_min": int(np.min(f)) if f is not None else None,
                "fold_max": int(np.max(f)) if f is not None else None,
                "fold_mean": float(np.mean(f)) if f is not None else None}

    def memory_gb(self) -> float:
        n = self.image.nbytes + (self.fold.nbytes if self.fold is not None else 0)
        return n / (1024 ** 3)

    def to_dict(self) -> Dict[str, Any]:
        return {"shape": self.shape, "has_fold": self.fold is not None,
                "config": self.config.to_dict() if self.config else None, "metadata": self.metadata,
                "statistics": self.get_statistics()}

    def __repr__(self) -> str:
        return f"<MigrationResult {'x'.join(map(str, self.shape))} ({self.memory_gb():.3g} GB)>"
