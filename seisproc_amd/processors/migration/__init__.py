"""Kirchhoff PSTM on gfx950 (drop-in for SeisProc's processors/migration)."""
from .base_migrator import BaseMigrator, MigrationResult
from .config_adapter import ConfigAdapter, MigrationParams, create_adapter_from_wizard
from .kirchhoff_kernel import migrate_kirchhoff_full, normalize_by_fold
from .kirchhoff_migrator import (KirchhoffMigrator, OptimizedKirchhoffMigrator, create_kirchhoff_migrator,
                                 create_optimized_kirchhoff_migrator)
from .migration_engine import MigrationEngine
from .pstm_processor import PSTMProcessor, migrate_arrays

__all__ = ["BaseMigrator", "MigrationResult", "ConfigAdapter", "MigrationParams", "create_adapter_from_wizard",
           "migrate_kirchhoff_full", "normalize_by_fold", "KirchhoffMigrator", "create_kirchhoff_migrator",
           "OptimizedKirchhoffMigrator", "create_optimized_kirchhoff_migrator",
           "MigrationEngine", "PSTMProcessor", "migrate_arrays"]
