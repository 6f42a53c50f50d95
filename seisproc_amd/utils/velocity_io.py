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
velocity files are read by the hand-written reader of ``utils/segy.py`` (the
reference uses segyio, :1096-1174): one trace is a v(t); several traces are
v(t, CDP) as in the reference, or, when their inline/crossline headers
cover a full lattice, a v(t, inline, xline) cube.  ``preview_velocity_file``
/ ``velocity_info_to_model`` (:166-512, :1175-1290) and
``create_velocity_from_picks`` (:953-1006) follow the reference.
"""
from __future__ import annotations

import json
import logging
import re
from dataclasses import dataclass, field
from enum import Enum
from pathlib import Path
from typing import Any, Dict, List, Optional, Tuple, Union

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
        return read_velocity_segy(path, velocity_type=kwargs.get("velocity_type", "rms"))
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


def convert_velocity_file(input_path: Union[str, Path], output_path: Union[str, Path], input_type: str = "rms",
                          output_type: str = "interval", **read_kwargs) -> VelocityModel:
    """Read a 1D velocity function, convert between RMS and interval
    velocities (Dix, on the file's own time axis) and write it
    (reference utils/velocity_io.py:902-950); the output format follows the
    output path's extension.  Returns the written model."""
    from ..models.velocity_model import interval_to_rms_velocity, rms_to_interval_velocity
    kinds = ("rms", "interval")
    if input_type not in kinds or output_type not in kinds:
        raise ValueError(f"velocity types are 'rms' or 'interval', got {input_type!r} -> {output_type!r}")
    model = read_velocity_auto(input_path, velocity_type=input_type, **read_kwargs)
    if model.velocity_type is not VelocityType.V_OF_Z:
        raise ValueError("Velocity conversion only supported for 1D v(z) models")
    if input_type != output_type:
        conv = rms_to_interval_velocity if input_type == "rms" else interval_to_rms_velocity
        t_axis, v = conv(model.z_axis, model.data)
        model = VelocityModel(data=np.asarray(v, dtype=np.float32), z_axis=np.asarray(t_axis), is_time=model.is_time,
                              metadata={**model.metadata, "velocity_type": output_type})
    write_velocity_auto(model, output_path)
    return model


# ---------------------------------------------------------------------------
# SEG-Y, preview and picks

def read_velocity_segy(filepath: Union[str, Path], velocity_type: str = "rms", inline_byte: Optional[int] = None,
                       xline_byte: Optional[int] = None) -> VelocityModel:
    """SEG-Y velocity file (reference utils/velocity_io.py:1096-1174): each
    trace is a velocity function at regular times (the sample interval and
    the first trace's delay time).  One trace -> v(t); several traces whose
    inline/crossline headers (bytes 189/193, or inline_byte/xline_byte) form a
    full lattice of more than one inline and crossline -> v(t, inline,
    xline) with ``lateral_axes='inline_xline'``; otherwise -> v(t, CDP) with
    the CDP numbers (trace numbers when every CDP header is 0) as lateral
    axis, as the reference builds it."""
    from .segy import read_segy
    path = Path(filepath)
    if not path.exists():
        raise FileNotFoundError(f"Velocity file not found: {path}")
    vol = read_segy(path, inline_byte, xline_byte)
    if vol.n_traces == 0:
        raise ValueError(f"{path}: no traces")
    times = vol.times_s.astype(np.float64)
    md = {"source_file": str(path), "velocity_type": velocity_type, "format": "segy",
          "sample_format": vol.format_code}
    if vol.n_traces == 1:
        return VelocityModel(data=vol.samples[0].astype(np.float32), z_axis=times, is_time=True, metadata=md)
    il, xl = vol.headers["inline"], vol.headers["crossline"]
    uil, uxl = np.unique(il), np.unique(xl)
    if (np.all(il != 0) and np.all(xl != 0) and uil.size > 1 and uxl.size > 1
            and uil.size * uxl.size == vol.n_traces
            and np.unique(il * (int(xl.max()) + 1) + xl).size == vol.n_traces):
        cube = np.empty((times.size, uil.size, uxl.size), dtype=np.float32)
        cube[:, np.searchsorted(uil, il), np.searchsorted(uxl, xl)] = vol.samples.T
        md.update(lateral_axes="inline_xline", inline_range=(int(uil.min()), int(uil.max())),
                  xline_range=(int(uxl.min()), int(uxl.max())))
        return VelocityModel(data=cube, z_axis=times, x_axis=uil.astype(np.float64), y_axis=uxl.astype(np.float64),
                             is_time=True, metadata=md)
    cdp = vol.headers["cdp"]
    x = cdp.astype(np.float64) if np.any(cdp != 0) else np.arange(vol.n_traces, dtype=np.float64)
    order = np.argsort(x, kind="stable")
    if np.unique(x).size != x.size:
        raise ValueError(f"{path}: repeated CDP numbers and no inline/crossline lattice")
    md.update(lateral_axes="cdp", cdp_range=(int(cdp.min()), int(cdp.max())) if np.any(cdp != 0) else None)
    return VelocityModel(data=vol.samples[order].T.astype(np.float32), z_axis=times, x_axis=x[order], is_time=True,
                         metadata=md)


@dataclass
class VelocityFileInfo:
    """What a velocity file holds, for preview (reference
    utils/velocity_io.py:59-91): sizes, ranges and the parsed arrays
    (``raw_data``) that velocity_info_to_model turns into a model."""

    path: Path
    format: VelocityFileFormat = VelocityFileFormat.UNKNOWN
    n_locations: int = 0
    n_time_samples: int = 0
    time_range: Tuple[float, float] = (0.0, 0.0)
    velocity_range: Tuple[float, float] = (0.0, 0.0)
    cdp_range: Optional[Tuple[int, int]] = None
    inline_range: Optional[Tuple[int, int]] = None
    xline_range: Optional[Tuple[int, int]] = None
    sample_interval: Optional[float] = None
    is_valid: bool = False
    error_message: Optional[str] = None
    raw_data: Dict[str, Any] = field(default_factory=dict)


def _regular_interval(t: np.ndarray) -> Optional[float]:
    d = np.diff(np.unique(t))
    return float(d[0]) if d.size and np.allclose(d, d[0], rtol=1e-6) else None


def preview_velocity_file(filepath: Union[str, Path], inline_byte: Optional[int] = None,
                          xline_byte: Optional[int] = None) -> VelocityFileInfo:
    """Detect the format and parse the file without unit conversion (times
    and velocities as written; SEG-Y times in s).  Errors are reported in
    ``error_message`` with ``is_valid`` False, never raised."""
    path = Path(filepath)
    info = VelocityFileInfo(path=path)
    if not path.exists():
        info.error_message = f"File not found: {path}"
        return info
    info.format = detect_velocity_format(path)
    try:
        rows = None
        if info.format in (VelocityFileFormat.ASCII_TV, VelocityFileFormat.ASCII_CDPTV,
                           VelocityFileFormat.ASCII_ILXLTV):
            need = {VelocityFileFormat.ASCII_TV: 2, VelocityFileFormat.ASCII_CDPTV: 3}.get(info.format, 4)
            rows = np.asarray([r[:need] for _, s in _data_lines(path) if len(r := _numbers(s)) >= need])
            if rows.size == 0:
                info.error_message = "No valid data found"
                return info
        if info.format is VelocityFileFormat.ASCII_TV:
            t, v = rows[:, 0], rows[:, 1]
            info.n_locations, info.n_time_samples = 1, t.size
            info.raw_data = {"times": t, "velocities": v, "format": "tv"}
        elif info.format is VelocityFileFormat.ASCII_CDPTV:
            cdp, t, v = rows[:, 0], rows[:, 1], rows[:, 2]
            u = np.unique(cdp)
            info.n_locations, info.n_time_samples = u.size, int(np.sum(cdp == u[0]))
            info.cdp_range = (int(u.min()), int(u.max()))
            info.raw_data = {"cdps": cdp, "times": t, "velocities": v, "unique_cdps": u, "format": "cdp_tv"}
        elif info.format is VelocityFileFormat.ASCII_ILXLTV:
            il, xl, t, v = rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3]
            uil, uxl = np.unique(il), np.unique(xl)
            info.n_locations = int(np.unique(np.stack([il, xl], 1), axis=0).shape[0])
            info.n_time_samples = int(np.sum((il == il[0]) & (xl == xl[0])))
            info.inline_range, info.xline_range = (int(uil.min()), int(uil.max())), (int(uxl.min()), int(uxl.max()))
            info.raw_data = {"inlines": il, "xlines": xl, "times": t, "velocities": v, "unique_inlines": uil,
                             "unique_xlines": uxl, "format": "ilxl_tv"}
        elif info.format is VelocityFileFormat.SEGY:
            from .segy import read_segy
            vol = read_segy(path, inline_byte, xline_byte)
            h = vol.headers
            t, v = vol.times_s, vol.samples
            info.n_locations, info.n_time_samples = vol.n_traces, t.size
            info.sample_interval = vol.interval_s
            for key, name in (("cdp", "cdp_range"), ("inline", "inline_range"), ("crossline", "xline_range")):
                if np.any(h[key] != 0):
                    setattr(info, name, (int(h[key].min()), int(h[key].max())))
            info.raw_data = {"format": "segy", "filepath": str(path), "inline_byte": inline_byte,
                             "xline_byte": xline_byte}
        elif info.format is VelocityFileFormat.JSON:
            m = read_velocity_json(path)
            t, v = np.atleast_1d(m.z_axis if m.z_axis is not None else [0.0]), np.atleast_1d(m.data)
            info.n_locations = int(np.prod(np.shape(m.data)[1:])) if np.ndim(m.data) > 1 else 1
            info.n_time_samples = int(np.shape(m.data)[0]) if np.ndim(m.data) else 1
            info.raw_data = {"format": "json", "data": m.to_dict()}
        elif info.format is VelocityFileFormat.NPZ:
            m = read_velocity_npy(path)
            t, v = np.atleast_1d(m.z_axis if m.z_axis is not None else [0.0]), np.atleast_1d(m.data)
            info.n_locations = int(np.prod(np.shape(m.data)[1:])) if np.ndim(m.data) > 1 else 1
            info.n_time_samples = int(np.shape(m.data)[0]) if np.ndim(m.data) else 1
            info.raw_data = {"format": "npz", "filepath": str(path)}
        else:
            info.error_message = "Unknown file format"
            return info
        info.time_range = (float(np.min(t)), float(np.max(t)))
        info.velocity_range = (float(np.min(v)), float(np.max(v)))
        if info.sample_interval is None:
            info.sample_interval = _regular_interval(np.asarray(t, dtype=np.float64))
        info.is_valid = True
    except Exception as e:  # the preview reports, the caller decides
        info.error_message = f"Preview error: {e}"
    return info


def velocity_info_to_model(info: VelocityFileInfo, velocity_type: str = "rms", time_unit: str = "ms",
                           velocity_unit: str = "m/s") -> VelocityModel:
    """Turn a preview into a model (reference :1175-1290): text formats with
    the given units (time-velocity -> v(t); CDP tables -> v(t, CDP);
    inline-xline tables -> v(t, inline) on the first inline's time axis, as
    the reference builds it); SEG-Y, JSON and .npz through their readers."""
    if not info.is_valid:
        raise ValueError(f"Invalid velocity file: {info.error_message}")
    raw = info.raw_data
    fmt = raw.get("format", "")
    tdiv = _scale(time_unit, _TIME_DIV, "time")
    vmul = _scale(velocity_unit, _VEL_SCALE, "velocity")
    if fmt == "tv":
        t = np.asarray(raw["times"], dtype=np.float64) / tdiv
        v = np.asarray(raw["velocities"], dtype=np.float64) * vmul
        o = np.argsort(t, kind="stable")
        return VelocityModel(data=v[o].astype(np.float32), z_axis=t[o], is_time=True,
                             metadata={"source_file": str(info.path), "velocity_type": velocity_type})
    if fmt == "cdp_tv":
        return read_velocity_cdptv(info.path, time_unit=time_unit, velocity_unit=velocity_unit,
                                   velocity_type=velocity_type)
    if fmt == "ilxl_tv":
        il = np.asarray(raw["inlines"])
        t = np.asarray(raw["times"], dtype=np.float64) / tdiv
        v = np.asarray(raw["velocities"], dtype=np.float64) * vmul
        uil = np.asarray(raw["unique_inlines"])
        t_axis = np.sort(t[il == uil[0]])
        vel = np.empty((t_axis.size, uil.size), dtype=np.float32)
        for i, u in enumerate(uil):
            m = il == u
            o = np.argsort(t[m])
            vel[:, i] = np.interp(t_axis, t[m][o], v[m][o])
        return VelocityModel(data=vel, z_axis=t_axis, x_axis=uil.astype(np.float64), is_time=True,
                             metadata={"source_file": str(info.path), "velocity_type": velocity_type,
                                       "lateral_axes": "cdp", "inline_range": info.inline_range,
                                       "xline_range": info.xline_range})
    if fmt == "segy":
        return read_velocity_segy(info.path, velocity_type=velocity_type, inline_byte=raw.get("inline_byte"),
                                  xline_byte=raw.get("xline_byte"))
    if fmt == "json":
        return read_velocity_json(info.path)
    if fmt == "npz":
        return read_velocity_npy(info.path)
    raise ValueError(f"Unknown velocity format: {fmt}")


def create_velocity_from_picks(picks: List[Tuple[float, float]], dt: float = 0.004, t_max: Optional[float] = None,
                               velocity_type: str = "rms") -> VelocityModel:
    """(time s, velocity m/s) picks -> v(t) on 0, dt, ..., t_max by linear
    interpolation, constant beyond the first and last pick (reference
    :953-1006)."""
    if not picks:
        raise ValueError("No velocity picks provided")
    picks = sorted(picks, key=lambda x: x[0])
    tp = np.array([p[0] for p in picks], dtype=np.float64)
    vp = np.array([p[1] for p in picks], dtype=np.float64)
    t_max = float(tp[-1]) if t_max is None else float(t_max)
    n = int(np.floor(t_max / dt + 1e-9)) + 1
    t_axis = dt * np.arange(n)
    return VelocityModel(data=np.interp(t_axis, tp, vp).astype(np.float32), z_axis=t_axis, is_time=True,
                         metadata={"velocity_type": velocity_type, "source": "picks", "n_picks": len(picks)})


def get_velocity_at_cdp(model: VelocityModel, cdp: int, times: Optional[np.ndarray] = None):
    """(times, velocities) of the model's velocity function at a CDP
    (reference utils/velocity_io.py:1291-1330): constant and v(t) models
    have one function (a constant one is given on 0 and 6 s unless times are
    asked for); a v(t, CDP) table is interpolated to the CDP number."""
    t = model.velocity_type
    if t is VelocityType.CONSTANT:
        times = np.array([0.0, 6.0], dtype=np.float32) if times is None else np.asarray(times)
        return times, np.full(np.shape(times), model.data, dtype=np.float32)
    if t is VelocityType.V_OF_Z:
        if times is None:
            return model.z_axis, model.data
        return times, np.interp(times, model.z_axis, model.data).astype(np.float32)
    if t is VelocityType.V_OF_XZ:
        times = model.z_axis if times is None else np.asarray(times)
        return times, model.get_velocity_at(times, x=float(cdp))
    raise ValueError(f"Unsupported velocity type: {t}")


def get_velocity_summary(model: VelocityModel) -> str:
    """A few lines describing the model: kind, extents, velocity range,
    velocity type and source file where known."""
    t = model.velocity_type
    rows = [("kind", t.value)]
    if t is VelocityType.CONSTANT:
        rows.append(("velocity", f"{float(model.data):g} m/s"))
    else:
        d = np.asarray(model.data)
        rows.append(("samples", " x ".join(map(str, d.shape))))
        rows.append(("times" if model.is_time else "depths", f"{float(model.z_axis[0]):g}..{float(model.z_axis[-1]):g}"))
        if model.x_axis is not None:
            rows.append(("x", f"{float(model.x_axis[0]):g}..{float(model.x_axis[-1]):g}"))
        rows.append(("velocity", f"{float(d.min()):g}..{float(d.max()):g} m/s"))
    if model.metadata.get("velocity_type"):
        rows.append(("type", str(model.metadata["velocity_type"])))
    if model.metadata.get("source_file"):
        rows.append(("source", Path(model.metadata["source_file"]).name))
    return "velocity model\n" + "\n".join(f"  {k:<9} {v}" for k, v in rows)
