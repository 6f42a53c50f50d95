"""Velocity model files -> VelocityModel -> device v_rms tables.

Same readers, writers, column conventions, unit handling and format
detection as utils/velocity_io.py of the reference (``detect_velocity_format``
:93-164, ``read_velocity_text`` :515-628, ``write_velocity_text`` :631-693,
JSON :696-742, NumPy :745-836, ``read_velocity_auto`` / ``write_velocity_auto``
:839-899, ``read_velocity_cdptv`` :1007-1093 and the inline-xline-time-velocity
layout of :297-344):

  two columns      time, velocity                      -> v(t)
  three columns    CDP, time, velocity                 -> v(CDP, t)
  four columns     inline, xline, time, velocity       -> v(inline, xline, t)
  .json            VelocityModel.to_dict
  .npz             arrays data / z_axis / x_axis / y_axis / is_time and a
                   JSON metadata string

Differences: .npz files are read with ``allow_pickle=False`` (the metadata
is a plain string array, so nothing needs unpickling); the lateral axes of
CDP and inline/xline tables are kept in those numbers and the model records
it (``metadata['lateral_axes']``) so that ``VelocityModel.vrms_on_grid``
samples them at the image columns' CDP or inline/xline numbers.  SEG-Y
velocity cubes are out of scope.
"""
from __future__ import annotations

import json
import logging
import re
from enum import Enum
from pathlib import Path
from typing import Optional, Union

import numpy as np

from ..models.velocity_model import VelocityModel, VelocityType

logger = logging.getLogger(__name__)

_TEXT_SUFFIXES = (".txt", ".vel", ".asc", ".dat", ".csv")
_TIME_DIV = {"s": 1.0, "ms": 1000.0, "us": 1e6}  # file time unit -> s (divide)
_VEL_SCALE = {"m/s": 1.0, "km/s": 1000.0, "ft/s": 0.3048, "ft/ms": 304.8}


class VelocityFileFormat(Enum):
    ASCII_TV = "ascii_tv"
    ASCII_CDPTV = "ascii_cdptv"
    ASCII_ILXLTV = "ascii_ilxltv"
    SEGY = "segy"
    JSON = "json"
    NPZ = "npz"
    UNKNOWN = "unknown"


def _data_lines(path: Path, skip_header: int = 0):
    with open(path) as fh:
        for i, line in enumerate(fh):
            if i < skip_header:
                continue
            s = line.strip()
            if s and not s.startswith("#") and not s.startswith("//"):
                yield i, s


def _numbers(line: str):
    out = []
    for tok in re.split(r"[,\s\t]+", line):
        try:
            out.append(float(tok))
        except ValueError:
            pass
    return out


def detect_velocity_format(filepath: Union[str, Path]) -> VelocityFileFormat:
    """By extension (.sgy/.segy, .json, .npy/.npz), else by the most common
    count of numeric columns among the first 20 data lines."""
    path = Path(filepath)
    if not path.exists():
        return VelocityFileFormat.UNKNOWN
    suf = path.suffix.lower()
    if suf in (".sgy", ".segy"):
        return VelocityFileFormat.SEGY
    if suf == ".json":
        return VelocityFileFormat.JSON
    if suf in (".npy", ".npz"):
        return VelocityFileFormat.NPZ
    counts = []
    try:
        for k, (_, s) in enumerate(_data_lines(path)):
            if k >= 20:
                break
            n = len(_numbers(s))
            if n:
                counts.append(n)
    except (OSError, UnicodeDecodeError):
        return VelocityFileFormat.UNKNOWN
    if not counts:
        return VelocityFileFormat.UNKNOWN
    c = max(set(counts), key=counts.count)
    return {2: VelocityFileFormat.ASCII_TV, 3: VelocityFileFormat.ASCII_CDPTV}.get(
        c, VelocityFileFormat.ASCII_ILXLTV if c >= 4 else VelocityFileFormat.UNKNOWN)


def _scale(unit: str, table, what: str) -> float:
    if unit not in table:
        raise ValueError(f"unknown {what} unit {unit!r} (one of {sorted(table)})")
    return table[unit]


def read_velocity_text(filepath: Union[str, Path], velocity_type: str = "rms", time_column: int = 0,
                       velocity_column: int = 1, time_unit: str = "ms", velocity_unit: str = "m/s",
                       delimiter: Optional[str] = None, skip_header: int = 0) -> VelocityModel:
    """Time-velocity pairs (comments # or //, blank lines skipped, malformed
    lines skipped with a warning), sorted by time; times to s, velocities to
    m/s."""
    path = Path(filepath)
    if not path.exists():
        raise FileNotFoundError(f"Velocity file not found: {path}")
    ts, vs = [], []
    need = max(time_column, velocity_column)
    for i, s in _data_lines(path, skip_header):
        parts = s.split(delimiter) if delimiter else s.split()
        if len(parts) <= need:
            logger.warning("Skipping line %d: insufficient columns", i + 1)
            continue
        try:
            t, v = float(parts[time_column]), float(parts[velocity_column])
        except ValueError as e:
            logger.warning("Skipping line %d: %s", i + 1, e)
            continue
        ts.append(t)
        vs.append(v)
    if not ts:
        raise ValueError(f"No valid velocity data found in {path}")
    t = np.asarray(ts, dtype=np.float32) / np.float32(_scale(time_unit, _TIME_DIV, "time"))
    v = np.asarray(vs, dtype=np.float32) * np.float32(_scale(velocity_unit, _VEL_SCALE, "velocity"))
    o = np.argsort(t, kind="stable")
    return VelocityModel(data=v[o], z_axis=t[o], is_time=True,
                         metadata={"source_file": str(path), "velocity_type": velocity_type,
                                   "original_time_unit": time_unit, "original_velocity_unit": velocity_unit})


def write_velocity_text(model: VelocityModel, filepath: Union[str, Path], time_unit: str = "ms",
                        velocity_unit: str = "m/s", delimiter: str = "\t", header: bool = True,
                        precision: int = 2) -> None:
    """Constant models as two rows (t = 0 and 10000 in the time unit); 1D
    models one row per sample."""
    path = Path(filepath)
    path.parent.mkdir(parents=True, exist_ok=True)
    vs = 1.0 / _scale(velocity_unit, _VEL_SCALE, "velocity")
    ts = _scale(time_unit, _TIME_DIV, "time")
    with open(path, "w") as fh:
        if model.velocity_type is VelocityType.CONSTANT:
            if header:
                fh.write("# Constant velocity model\n")
                fh.write(f"# Time({time_unit}){delimiter}Velocity({velocity_unit})\n")
            v = model.data * vs
            fh.write(f"0.0{delimiter}{v:.{precision}f}\n10000.0{delimiter}{v:.{precision}f}\n")
            return
        if model.velocity_type is not VelocityType.V_OF_Z:
            raise ValueError("text output holds constant or 1D models")
        if header:
            fh.write(f"# Velocity model: {model.velocity_type.value}\n")
            fh.write(f"# v0={model.v0:.0f} m/s, gradient={model.gradient:.4f}\n")
            fh.write(f"# Time({time_unit}){delimiter}Velocity({velocity_unit})\n")
        for t, v in zip(model.z_axis * ts, model.data * vs):
            fh.write(f"{t:.{precision}f}{delimiter}{v:.{precision}f}\n")


def read_velocity_json(filepath: Union[str, Path]) -> VelocityModel:
    path = Path(filepath)
    if not path.exists():
        raise FileNotFoundError(f"Velocity file not found: {path}")
    m = VelocityModel.from_dict(json.loads(path.read_text()))
    m.metadata["source_file"] = str(path)
    return m


def write_velocity_json(model: VelocityModel, filepath: Union[str, Path], indent: int = 2) -> None:
    path = Path(filepath)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(model.to_dict(), indent=indent))


def read_velocity_npy(filepath: Union[str, Path]) -> VelocityModel:
    """.npz with 'data' and optional 'z_axis', 'x_axis', 'y_axis', 'is_time'
    and a JSON 'metadata' string (loaded without pickle)."""
    path = Path(filepath)
    if not path.exists():
        raise FileNotFoundError(f"Velocity file not found: {path}")
    obj = np.load(path, allow_pickle=False)
    if isinstance(obj, np.ndarray):
        if obj.ndim == 0:
            return VelocityModel(data=float(obj), metadata={"source_file": str(path)})
        raise ValueError(f"{path}: a .npy file holds no axes; store data and z_axis in an .npz")
    with obj as npz:
        data = npz["data"]
        data = float(data) if data.ndim == 0 else data
        ax = lambda k: (npz[k] if k in npz and npz[k].ndim > 0 else None)
        z, x, y = ax("z_axis"), ax("x_axis"), ax("y_axis")
        is_time = bool(npz["is_time"]) if "is_time" in npz else True
        md = {}
        if "metadata" in npz:
            try:
                md = json.loads(str(npz["metadata"]))
            except ValueError:
                md = {}
    md["source_file"] = str(path)
    return VelocityModel(data=data, z_axis=z, x_axis=x, y_axis=y, is_time=is_time, metadata=md)


def write_velocity_npy(model: VelocityModel, filepath: Union[str, Path]) -> None:
    path = Path(filepath)
    if not path.suffix:
        path = path.with_suffix(".npz")
    path.parent.mkdir(parents=True, exist_ok=True)
    arrays = {"data": np.asarray(model.data), "is_time": np.asarray(model.is_time),
              "metadata": np.asarray(json.dumps(model.metadata))}
    for k in ("z_axis", "x_axis", "y_axis"):
        if getattr(model, k) is not None:
            arrays[k] = getattr(model, k)
    np.savez(path, **arrays)


def read_velocity_cdptv(filepath: Union[str, Path], time_unit: str = "ms", velocity_unit: str = "m/s",
                        velocity_type: str = "rms") -> VelocityModel:
    """CDP, time, velocity rows -> v(t, CDP) on the first CDP's sorted time
    axis (other CDPs interpolated onto it); the lateral axis is the CDP
    number."""
    path = Path(filepath)
    if not path.exists():
        raise FileNotFoundError(f"Velocity file not found: {path}")
    rows = [r[:3] for _, s in _data_lines(path) if len(r := _numbers(s)) >= 3]
    if not rows:
        raise ValueError(f"No valid CDP-time-velocity data found in {path}")
    a = np.asarray(rows, dtype=np.float64)
    cdp, t = a[:, 0], a[:, 1] / _scale(time_unit, _TIME_DIV, "time")
    v = a[:, 2] * _scale(velocity_unit, _VEL_SCALE, "velocity")
    ucdp = np.unique(cdp)
    t_axis = np.sort(t[cdp == ucdp[0]])
    vel = np.empty((t_axis.size, ucdp.size), dtype=np.float32)
    for j, c in enumerate(ucdp):
        m = cdp == c
        o = np.argsort(t[m])
        vel[:, j] = np.interp(t_axis, t[m][o], v[m][o])
    return VelocityModel(data=vel, z_axis=t_axis, x_axis=ucdp, is_time=True,
                         metadata={"source_file": str(path), "velocity_type": velocity_type, "lateral_axes": "cdp",
                                   "cdp_range": (int(ucdp.min()), int(ucdp.max()))})


def read_velocity_ilxltv(filepath: Union[str, Path], time_unit: str = "ms", velocity_unit: str = "m/s",
                         velocity_type: str = "rms") -> VelocityModel:
    """Inline, xline, time, velocity rows on a regular (inline, xline)
    lattice -> v(t, inline, xline) on the sorted union time axis (each
    location's function interpolated onto it, constant beyond its ends)."""
    path = Path(filepath)
    if not path.exists():
        raise FileNotFoundError(f"Velocity file not found: {path}")
    rows = [r[:4] for _, s in _data_lines(path) if len(r := _numbers(s)) >= 4]
    if not rows:
        raise ValueError(f"No valid inline-xline-time-velocity data found in {path}")
    a = np.asarray(rows, dtype=np.float64)
    il, xl = a[:, 0].round().astype(np.int64), a[:, 1].round().astype(np.int64)
    t = a[:, 2] / _scale(time_unit, _TIME_DIV, "time")
    v = a[:, 3] * _scale(velocity_unit, _VEL_SCALE, "velocity")
    uil, uxl = np.unique(il), np.unique(xl)
    t_axis = np.unique(t)
    vel = np.full((t_axis.size, uil.size, uxl.size), np.nan, dtype=np.float32)
    for i, ii in enumerate(uil):
        for j, jj in enumerate(uxl):
            m = (il == ii) & (xl == jj)
            if not m.any():
                continue
            o = np.argsort(t[m])
            vel[:, i, j] = np.interp(t_axis, t[m][o], v[m][o])
    if np.isnan(vel).any():
        raise ValueError(f"{path}: the (inline, xline) locations do not cover a full lattice")
    return VelocityModel(data=vel, z_axis=t_axis, x_axis=uil.astype(np.float64), y_axis=uxl.astype(np.float64),
                         is_time=True, metadata={"source_file": str(path), "velocity_type": velocity_type,
                                                 "lateral_axes": "inline_xline"})


def read_velocity_auto(filepath: Union[str, Path], **kwargs) -> VelocityModel:
    """By extension (text, .json, .npy/.npz); text files by column count
    (2: time-velocity, 3: CDP-time-velocity, >= 4: inline-xline-time-velocity)."""
    path = Path(filepath)
    suf = path.suffix.lower()
    if suf == ".json":
        return read_velocity_json(path)
    if suf in (".npy", ".npz"):
        return read_velocity_npy(path)
    if suf in (".sgy", ".segy"):
        raise NotImplementedError("SEG-Y velocity cubes are not read here; load the cube with a SEG-Y "
                                  "reader and pass it as a 3D VelocityModel or a device v_rms array")
    fmt = detect_velocity_format(path)
    units = {k: kwargs[k] for k in ("time_unit", "velocity_unit", "velocity_type") if k in kwargs}
    if fmt is VelocityFileFormat.ASCII_CDPTV:
        return read_velocity_cdptv(path, **units)
    if fmt is VelocityFileFormat.ASCII_ILXLTV:
        return read_velocity_ilxltv(path, **units)
    if suf not in _TEXT_SUFFIXES:
        logger.warning("Unknown extension %s, trying text format", suf)
    return read_velocity_text(path, **kwargs)


def write_velocity_auto(model: VelocityModel, filepath: Union[str, Path], **kwargs) -> None:
    path = Path(filepath)
    suf = path.suffix.lower()
    if suf in _TEXT_SUFFIXES:
        write_velocity_text(model, path, **kwargs)
    elif suf in (".npy", ".npz"):
        write_velocity_npy(model, path)
    else:
        if suf != ".json":
            logger.warning("Unknown extension %s, using JSON format", suf)
        write_velocity_json(model, path, **kwargs)
