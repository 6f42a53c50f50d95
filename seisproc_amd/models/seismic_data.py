"""SeismicData: traces plus metadata (reference models/seismic_data.py:11-97).

traces are (n_samples, n_traces) like the reference; sample_rate is in ms.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Dict, Optional

import numpy as np


@dataclass
class SeismicData:
    traces: np.ndarray
    sample_rate: float  # milliseconds
    headers: Optional[Dict[str, np.ndarray]] = None
    metadata: Dict[str, Any] = field(default_factory=dict)

    def __post_init__(self):
        if self.traces.ndim != 2:
            raise ValueError(f"Traces must be 2D array, got shape {self.traces.shape}")
        if self.sample_rate <= 0:
            raise ValueError(f"Sample rate must be positive, got {self.sample_rate}")
        if not np.issubdtype(self.traces.dtype, np.floating):
            object.__setattr__(self, "traces", self.traces.astype(np.float32))

    @property
    def n_samples(self) -> int:
        return self.traces.shape[0]

    @property
    def n_traces(self) -> int:
        return self.traces.shape[1]

    @property
    def duration(self) -> float:
        return (self.n_samples - 1) * self.sample_rate

    @property
    def nyquist_freq(self) -> float:
        return 1000.0 / (2.0 * self.sample_rate)

    def get_time_axis(self) -> np.ndarray:
        return np.arange(self.n_samples) * self.sample_rate

    def get_trace_axis(self) -> np.ndarray:
        return np.arange(self.n_traces)

    def copy(self) -> "SeismicData":
        return SeismicData(traces=self.traces.copy(), sample_rate=self.sample_rate,
                           headers={k: v.copy() for k, v in self.headers.items()} if self.headers else None,
                           metadata=dict(self.metadata))

    def __repr__(self) -> str:
        # the dataclass default would print every sample
        return f"<SeismicData {self.n_traces} traces x {self.n_samples} samples at {self.sample_rate:g} ms>"
