"""Processors of the migration path and their registry (reference
processors/__init__.py:58-78 registers 'KirchhoffMigrator'; the PSTM
processor is registered beside it)."""
from typing import Dict, Type

from .base_processor import BaseProcessor, ProgressCallback

PROCESSOR_REGISTRY: Dict[str, Type[BaseProcessor]] = {}


def _register():
    from .migration.kirchhoff_migrator import KirchhoffMigrator
    from .migration.pstm_processor import PSTMProcessor
    PROCESSOR_REGISTRY.update({"KirchhoffMigrator": KirchhoffMigrator, "PSTMProcessor": PSTMProcessor})


def get_processor_class(name: str) -> Type[BaseProcessor]:
    if not PROCESSOR_REGISTRY:
        _register()
    if name not in PROCESSOR_REGISTRY:
        raise ValueError(f"Unknown processor: {name}. Available: {sorted(PROCESSOR_REGISTRY)}")
    return PROCESSOR_REGISTRY[name]


def create_processor(name: str, **params) -> BaseProcessor:
    return get_processor_class(name)(**params)


_register()

__all__ = ["BaseProcessor", "ProgressCallback", "PROCESSOR_REGISTRY", "get_processor_class", "create_processor"]
