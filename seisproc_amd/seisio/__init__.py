"""Seismic I/O on the migration path: the multi-volume output manager."""
from .migration_output import MigrationOutputManager, OutputVolumeInfo, create_output_manager

__all__ = ["MigrationOutputManager", "OutputVolumeInfo", "create_output_manager"]
