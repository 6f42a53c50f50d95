"""BaseProcessor: the processor contract of the reference
(processors/base_processor.py:17-146) -- parameters in ``params``,
``_validate_params`` raising ValueError, ``process(SeismicData) ->
SeismicData`` returning a new object, (current, total, message) progress
callbacks and ``to_dict``/``from_dict`` reconstruction by module/class name.
"""
from __future__ import annotations

import importlib
from abc import ABC, abstractmethod
from typing import Any, Callable, Dict, Optional

from ..models.seismic_data import SeismicData

ProgressCallback = Callable[[int, int, str], None]


class BaseProcessor(ABC):
    def __init__(self, **params):
        self.params = params
        self._progress_callback: Optional[ProgressCallback] = None
        self._validate_params()

    def set_progress_callback(self, callback: Optional[ProgressCallback]) -> "BaseProcessor":
        self._progress_callback = callback
        return self

    def _report_progress(self, current: int, total: int, message: str = "") -> None:
        if self._progress_callback is not None:
            self._progress_callback(current, total, message)

    @abstractmethod
    def _validate_params(self) -> None:
        """Raise ValueError on invalid parameters."""

    @abstractmethod
    def process(self, data: SeismicData) -> SeismicData:
        """Return a new SeismicData; the input is not modified."""

    @abstractmethod
    def get_description(self) -> str:
        """Human-readable description of the processor and its parameters."""

    def __repr__(self) -> str:
        return f"{self.__class__.__name__}({self.params})"

    def to_dict(self) -> Dict[str, Any]:
        return {"class_name": self.__class__.__name__, "module": self.__class__.__module__,
                "params": dict(self.params)}

    @classmethod
    def from_dict(cls, config: Dict[str, Any]) -> "BaseProcessor":
        try:
            module = importlib.import_module(config["module"])
            proc_cls = getattr(module, config["class_name"])
            if not (isinstance(proc_cls, type) and issubclass(proc_cls, BaseProcessor)):
                raise ValueError(f"{config['class_name']} is not a BaseProcessor subclass")
            return proc_cls(**config["params"])
        except KeyError as e:
            raise ValueError(f"Invalid processor config - missing key: {e}")
        except ImportError as e:
            raise ValueError(f"Cannot import processor module '{config.get('module')}': {e}")
        except AttributeError as e:
            raise ValueError(f"Processor class '{config.get('class_name')}' not found: {e}")
