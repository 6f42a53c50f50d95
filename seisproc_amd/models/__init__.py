"""Data and configuration objects of the migration path (same names and
fields as SeisProc's models/ package)."""
from .seismic_data import SeismicData
from .migration_config import (AnisotropyMethod, InterpolationMode, MigrationConfig, OutputGrid,
                               TraveltimeMode, WeightMode, create_default_config)
from .velocity_model import VelocityModel, VelocityType
from .anisotropy_model import AnisotropyModel, AnisotropyType
from .migration_geometry import MigrationGeometry
from .migration_job import HeaderMapping, MigrationJobConfig

__all__ = ["SeismicData", "AnisotropyMethod", "InterpolationMode", "MigrationConfig", "OutputGrid",
           "TraveltimeMode", "WeightMode", "create_default_config", "VelocityModel", "VelocityType",
           "AnisotropyModel", "AnisotropyType", "MigrationGeometry", "HeaderMapping",
           "MigrationJobConfig"]
