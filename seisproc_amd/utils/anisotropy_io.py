"""Anisotropy (VTI epsilon/delta) model files -> AnisotropyModel -> device
eta tables.

Same readers and writers as utils/anisotropy_io.py of the reference
(``read_anisotropy_text`` :31-131: z, epsilon, delta columns, comments #
and //, z in s or ms, rows sorted by z; ``write_anisotropy_text`` :134-192;
JSON :195-241; NumPy .npz with epsilon / delta / z_axis and a JSON metadata
string :244-329; ``read_anisotropy_auto`` / ``write_anisotropy_auto``
:332-392).  eta = (epsilon - delta) / (1 + 2 delta) as in
models/anisotropy_model.py:133-155; ``AnisotropyModel.eta_on_grid`` gives the
kernel's table.  .npz files are read without pickle.
"""
from __future__ import annotations

import json
import logging
from pathlib import Path
from typing import Dict, Optional, Union

import numpy as np

from ..models.anisotropy_model import AnisotropyModel, AnisotropyType, create_1d_anisotropy

logger = logging.getLogger(__name__)

_TEXT_SUFFIXES = (".txt", ".ani", ".asc", ".dat", ".csv")


def read_anisotropy_text(filepath: Union[str, Path], z_column: int = 0, epsilon_column: int = 1,
                         delta_column: int = 2, z_unit: str = "s", delimiter: Optional[str] = None,
                         skip_header: int = 0) -> AnisotropyModel:
    path = Path(filepath)
    if not path.exists():
        raise FileNotFoundError(f"Anisotropy file not found: {path}")
    zs, es, ds = [], [], []
    need = max(z_column, epsilon_column, delta_column)
    with open(path) as fh:
        for i, line in enumerate(fh):
            if i < skip_header:
                continue
            s = line.strip()
            if not s or s.startswith("#") or s.startswith("//"):
                continue
            parts = s.split(delimiter) if delimiter else s.split()
            if len(parts) <= need:
                logger.warning("Skipping line %d: insufficient columns", i + 1)
                continue
            try:
                z, e, d = float(parts[z_column]), float(parts[epsilon_column]), float(parts[delta_column])
            except ValueError as err:
                logger.warning("Skipping line %d: %s", i + 1, err)
                continue
            zs.append(z)
            es.append(e)
            ds.append(d)
    if not zs:
        raise ValueError(f"No valid anisotropy data found in {path}")
    z = np.asarray(zs, dtype=np.float32)
    if z_unit == "ms":
        z = z / np.float32(1000.0)
    o = np.argsort(z, kind="stable")
    e, d = np.asarray(es, dtype=np.float32)[o], np.asarray(ds, dtype=np.float32)[o]
    m = AnisotropyModel(epsilon=e.astype(np.float64), delta=d.astype(np.float64), z_axis=z[o].astype(np.float64))
    m.metadata.update({"source_file": str(path), "original_z_unit": z_unit})
    return m


def write_anisotropy_text(model: AnisotropyModel, filepath: Union[str, Path], z_unit: str = "s",
                          delimiter: str = "\t", header: bool = True, precision: int = 4) -> None:
    path = Path(filepath)
    path.parent.mkdir(parents=True, exist_ok=True)
    zs = 1000.0 if z_unit == "ms" else 1.0
    with open(path, "w") as fh:
        if header:
            fh.write(f"# VTI anisotropy model ({model.anisotropy_type.value})\n")
            fh.write(f"# Z({z_unit}){delimiter}Epsilon{delimiter}Delta\n")
        if model.anisotropy_type is AnisotropyType.CONSTANT:
            for z in (0.0, 10.0):
                fh.write(f"{z * zs:.{precision}f}{delimiter}{model.epsilon:.{precision}f}{delimiter}"
                         f"{model.delta:.{precision}f}\n")
            return
        if model.anisotropy_type is not AnisotropyType.V_OF_Z:
            raise ValueError("text output holds constant or 1D models")
        for z, e, d in zip(np.asarray(model.z_axis) * zs, model.epsilon, model.delta):
            fh.write(f"{z:.{precision}f}{delimiter}{e:.{precision}f}{delimiter}{d:.{precision}f}\n")


def read_anisotropy_json(filepath: Union[str, Path]) -> AnisotropyModel:
    path = Path(filepath)
    if not path.exists():
        raise FileNotFoundError(f"Anisotropy file not found: {path}")
    m = AnisotropyModel.from_dict(json.loads(path.read_text()))
    m.metadata["source_file"] = str(path)
    return m


def write_anisotropy_json(model: AnisotropyModel, filepath: Union[str, Path], indent: int = 2) -> None:
    path = Path(filepath)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(model.to_dict(), indent=indent))


def read_anisotropy_npy(filepath: Union[str, Path]) -> AnisotropyModel:
    path = Path(filepath)
    if not path.exists():
        raise FileNotFoundError(f"Anisotropy file not found: {path}")
    with np.load(path, allow_pickle=False) as npz:
        e, d = npz["epsilon"], npz["delta"]
        z = npz["z_axis"] if "z_axis" in npz and npz["z_axis"].ndim > 0 else None
        md = {}
        if "metadata" in npz:
            try:
                md = json.loads(str(npz["metadata"]))
            except ValueError:
                md = {}
    if z is not None:
        m = AnisotropyModel(epsilon=e.astype(np.float64), delta=d.astype(np.float64), z_axis=z.astype(np.float64))
    elif e.ndim == 0:
        m = AnisotropyModel(epsilon=float(e), delta=float(d))
    else:
        raise ValueError("Invalid anisotropy data format")
    m.metadata.update(md)
    m.metadata["source_file"] = str(path)
    return m


def write_anisotropy_npy(model: AnisotropyModel, filepath: Union[str, Path]) -> None:
    path = Path(filepath)
    if not path.suffix:
        path = path.with_suffix(".npz")
    path.parent.mkdir(parents=True, exist_ok=True)
    arrays = {"epsilon": np.asarray(model.epsilon), "delta": np.asarray(model.delta),
              "metadata": np.asarray(json.dumps(model.metadata))}
    if model.z_axis is not None:
        arrays["z_axis"] = np.asarray(model.z_axis)
    np.savez(path, **arrays)


def read_anisotropy_auto(filepath: Union[str, Path], **kwargs) -> AnisotropyModel:
    path = Path(filepath)
    suf = path.suffix.lower()
    if suf == ".json":
        return read_anisotropy_json(path)
    if suf in (".npy", ".npz"):
        return read_anisotropy_npy(path)
    if suf not in _TEXT_SUFFIXES:
        logger.warning("Unknown extension %s, trying text format", suf)
    return read_anisotropy_text(path, **kwargs)


def write_anisotropy_auto(model: AnisotropyModel, filepath: Union[str, Path], **kwargs) -> None:
    path = Path(filepath)
    suf = path.suffix.lower()
    if suf in _TEXT_SUFFIXES:
        write_anisotropy_text(model, path, **kwargs)
    elif suf in (".npy", ".npz"):
        write_anisotropy_npy(model, path)
    else:
        write_anisotropy_json(model, path, **kwargs)


def create_anisotropy_from_picks(picks, dz: float = 0.004, z_max: Optional[float] = None) -> AnisotropyModel:
    """(z, epsilon, delta) picks -> epsilon(z), delta(z) on 0, dz, ..., z_max
    by linear interpolation, constant beyond the end picks (reference
    utils/anisotropy_io.py:395-437); eta follows from epsilon and delta."""
    if not picks:
        raise ValueError("No anisotropy picks provided")
    picks = sorted(picks, key=lambda x: x[0])
    zp = np.array([p[0] for p in picks], dtype=np.float64)
    ep = np.array([p[1] for p in picks], dtype=np.float64)
    dp = np.array([p[2] for p in picks], dtype=np.float64)
    z_max = float(zp[-1]) if z_max is None else float(z_max)
    n = int(np.floor(z_max / dz + 1e-9)) + 1
    z = dz * np.arange(n)
    m = AnisotropyModel(epsilon=np.interp(z, zp, ep).astype(np.float32), delta=np.interp(z, zp, dp).astype(np.float32),
                        z_axis=z)
    m.metadata.update(source="picks", n_picks=len(picks))
    return m


def estimate_anisotropy_from_wells(well_data: Dict[str, np.ndarray], method: str = "average") -> AnisotropyModel:
    """A v(z) anisotropy model from a well (reference utils/anisotropy_io.py:
    438-486): the logged epsilon and delta along 'z' where given, else, with
    only velocities ('vp'), the typical shale values epsilon 0.15 and
    delta 0.08 at every z.  The method name is recorded in the metadata."""
    z = well_data.get("z")
    if z is None:
        raise ValueError("Well data must include 'z' axis")
    z = np.asarray(z, dtype=np.float64)
    if "epsilon" in well_data and "delta" in well_data:
        eps, dl = well_data["epsilon"], well_data["delta"]
    elif well_data.get("vp") is not None:
        eps, dl = np.full(z.shape, 0.15), np.full(z.shape, 0.08)
    else:
        raise ValueError("Well data must include either (epsilon, delta) or velocities")
    model = create_1d_anisotropy(z, eps, dl)
    model.metadata.update(source="well_estimate", method=method)
    return model
