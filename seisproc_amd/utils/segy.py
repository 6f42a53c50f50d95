"""Minimal SEG-Y reading and writing for velocity and anisotropy volumes.

The reference reads SEG-Y velocity files with segyio (utils/velocity_io.py:
1096-1174), which is not available here; this module reads the same
information from the file itself, following the SEG-Y revision 1/2 layout:

  bytes 1-3200      textual header (skipped; EBCDIC or ASCII)
  bytes 3201-3600   binary header, big-endian unless the revision-2
                    byte-order word (bytes 3297-3300) reads 0x04030201
                      3217-3218  sample interval (microseconds, unsigned)
                      3221-3222  samples per trace (unsigned)
                      3225-3226  sample format code
                      3505-3506  number of extended textual headers
  then, per trace, a 240-byte trace header and the samples
                      21-24    CDP ensemble number
                      109-110  delay recording time (ms)
                      115-116  samples in this trace, 117-118 its interval (us);
                               used when the binary header holds zero
                      181-184  CDP X, 185-188 CDP Y
                      189-192  inline (3D), 193-196 crossline (3D)

Sample formats: 1 IBM hexadecimal float, 2 int32, 3 int16, 5 IEEE float32,
6 IEEE float64, 8 int8.  IBM floats are (-1)^s 0.F x 16^(e - 64) with a
7-bit exponent e and a 24-bit fraction F (IBM System/360 format).
Traces are read whole with numpy; every trace has the binary header's sample
count (fixed-length traces).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, Optional, Union

import numpy as np

TEXT_BYTES = 3200
BIN_BYTES = 400
TRACE_HEADER_BYTES = 240
_FORMATS = {1: ("ibm", 4), 2: (">i4", 4), 3: (">i2", 2), 5: (">f4", 4), 6: (">f8", 8), 8: ("i1", 1)}
# trace header fields (1-based byte, numpy type)
_FIELDS = {"cdp": (21, "i4"), "delay_ms": (109, "i2"), "n_samples": (115, "u2"), "interval_us": (117, "u2"),
           "cdp_x": (181, "i4"), "cdp_y": (185, "i4"), "inline": (189, "i4"), "crossline": (193, "i4")}


def ibm_to_float(words: np.ndarray) -> np.ndarray:
    """IBM System/360 single-precision floats (as uint32 words) to float64."""
    u = np.asarray(words, dtype=np.uint32).astype(np.uint64)
    sign = np.where((u >> 31) & 1, -1.0, 1.0)
    exp = ((u >> 24) & 0x7F).astype(np.int64)
    frac = (u & 0xFFFFFF).astype(np.float64)
    return sign * np.ldexp(frac, 4 * (exp - 64) - 24)


def float_to_ibm(x: np.ndarray) -> np.ndarray:
    """float values to IBM single-precision words (uint32), fraction
    truncated to 24 bits (exact for values with <= 21 significant bits)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros(x.shape, dtype=np.uint32)
    nz = x != 0
    a = np.abs(x[nz])
    m, e2 = np.frexp(a)                       # a = m 2^e2, m in [0.5, 1)
    e16 = np.ceil(e2 / 4.0).astype(np.int64)  # a = f 16^e16 with f in [1/16, 1)
    f = a / np.power(16.0, e16)
    frac = np.floor(f * (1 << 24)).astype(np.uint64)
    over = frac >= (1 << 24)
    frac[over] >>= 4
    e16[over] += 1
    words = (np.where(x[nz] < 0, 1, 0).astype(np.uint64) << 31) | ((e16 + 64).astype(np.uint64) << 24) | frac
    out[nz] = words.astype(np.uint32)
    return out


@dataclass
class SegyVolume:
    """Traces of a SEG-Y file: samples (n_traces, n_samples) float64, the
    time of every sample (s), and the header fields per trace."""

    samples: np.ndarray
    times_s: np.ndarray
    interval_s: float
    format_code: int
    headers: Dict[str, np.ndarray] = field(default_factory=dict)
    little_endian: bool = False

    @property
    def n_traces(self) -> int:
        return int(self.samples.shape[0])


def read_segy(path: Union[str, Path], inline_byte: Optional[int] = None,
              xline_byte: Optional[int] = None) -> SegyVolume:
    """Read every trace and its header fields.  inline_byte / xline_byte:
    1-based positions of 4-byte inline / crossline numbers when they are not
    at the standard 189 / 193."""
    path = Path(path)
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size < TEXT_BYTES + BIN_BYTES:
        raise ValueError(f"{path}: too short for a SEG-Y file ({raw.size} bytes)")
    binh = raw[TEXT_BYTES:TEXT_BYTES + BIN_BYTES]
    order = binh[96:100].view(">u4")[0]
    little = order == 0x04030201
    e = "<" if little else ">"
    i2 = lambda off: int(binh[off:off + 2].view(e + "i2")[0])
    u2 = lambda buf, off: int(buf[off:off + 2].view(e + "u2")[0])
    interval_us, n_samples, fmt = u2(binh, 16), u2(binh, 20), i2(24)
    n_ext = i2(304)
    if fmt not in _FORMATS:
        raise ValueError(f"{path}: unsupported SEG-Y sample format code {fmt}")
    kind, nbytes = _FORMATS[fmt]
    start = TEXT_BYTES + BIN_BYTES + max(n_ext, 0) * TEXT_BYTES
    if (n_samples == 0 or interval_us == 0) and raw.size >= start + TRACE_HEADER_BYTES:
        # zero in the binary header: take the first trace header's values
        th0 = raw[start:start + TRACE_HEADER_BYTES]
        n_samples = n_samples or u2(th0, 114)
        interval_us = interval_us or u2(th0, 116)
    if n_samples <= 0:
        raise ValueError(f"{path}: neither the binary header nor the first trace header gives a sample count")
    rec = TRACE_HEADER_BYTES + n_samples * nbytes
    body = raw[start:]
    n_tr = body.size // rec
    if n_tr * rec != body.size:
        raise ValueError(f"{path}: {body.size} bytes of traces is not a whole number of {rec}-byte traces")
    body = body.reshape(n_tr, rec)
    th, data = body[:, :TRACE_HEADER_BYTES], body[:, TRACE_HEADER_BYTES:]
    hdr = {}
    fields = dict(_FIELDS)
    if inline_byte is not None:
        fields["inline"] = (int(inline_byte), "i4")
    if xline_byte is not None:
        fields["crossline"] = (int(xline_byte), "i4")
    for name, (b, t) in fields.items():
        n = np.dtype(t).itemsize
        hdr[name] = np.ascontiguousarray(th[:, b - 1:b - 1 + n]).view(e + t).reshape(n_tr).astype(np.int64)
    if kind == "ibm":
        words = np.ascontiguousarray(data).view(e + "u4").reshape(n_tr, n_samples)
        samples = ibm_to_float(words)
    else:
        dt = np.dtype(kind).newbyteorder("<" if little else ">") if kind != "i1" else np.dtype("i1")
        samples = np.ascontiguousarray(data).view(dt).reshape(n_tr, n_samples).astype(np.float64)
    t0 = (hdr["delay_ms"][0] if n_tr else 0) / 1000.0
    dt_s = interval_us / 1e6
    return SegyVolume(samples=samples, times_s=t0 + dt_s * np.arange(n_samples), interval_s=dt_s,
                      format_code=fmt, headers=hdr, little_endian=bool(little))


def write_segy(path: Union[str, Path], samples: np.ndarray, interval_s: float, *, format_code: int = 1,
               cdp=None, inline=None, crossline=None, cdp_x=None, cdp_y=None, delay_ms: int = 0,
               little_endian: bool = False, n_extended_text: int = 0) -> None:
    """Write traces (n_traces, n_samples) as SEG-Y (revision 2 byte-order
    word, fixed-length traces) with the given header fields."""
    samples = np.asarray(samples, dtype=np.float64)
    n_tr, n_s = samples.shape
    if format_code not in _FORMATS:
        raise ValueError(f"unsupported format code {format_code}")
    e = "<" if little_endian else ">"
    text = np.full(TEXT_BYTES, 0x40, dtype=np.uint8)  # EBCDIC blanks
    binh = np.zeros(BIN_BYTES, dtype=np.uint8)

    def put(buf, off, val, t):
        buf[off:off + np.dtype(t).itemsize] = np.frombuffer(np.asarray(val, dtype=e + t).tobytes(), dtype=np.uint8)

    put(binh, 16, int(round(interval_s * 1e6)), "u2")
    put(binh, 20, n_s, "u2")
    put(binh, 24, format_code, "i2")
    put(binh, 96, 0x01020304, "u4")
    put(binh, 300, 0x0200, "u2")  # revision 2.0
    put(binh, 302, 1, "i2")       # fixed-length traces
    put(binh, 304, n_extended_text, "i2")
    kind, nbytes = _FORMATS[format_code]
    if kind == "ibm":
        body = float_to_ibm(samples).astype(e + "u4")
    else:
        dt = np.dtype(kind).newbyteorder("<" if little_endian else ">") if kind != "i1" else np.dtype("i1")
        body = np.round(samples).astype(dt) if dt.kind == "i" else samples.astype(dt)
    rec = np.zeros((n_tr, TRACE_HEADER_BYTES + n_s * nbytes), dtype=np.uint8)
    rec[:, TRACE_HEADER_BYTES:] = np.frombuffer(np.ascontiguousarray(body).tobytes(), dtype=np.uint8).reshape(n_tr, -1)
    cols = {"cdp": cdp, "inline": inline, "crossline": crossline, "cdp_x": cdp_x, "cdp_y": cdp_y}
    for i in range(n_tr):
        h = rec[i]
        put(h, 108, delay_ms, "i2")
        put(h, 114, n_s, "u2")
        put(h, 116, int(round(interval_s * 1e6)), "u2")
        for name, vals in cols.items():
            if vals is not None:
                b, t = _FIELDS[name]
                put(h, b - 1, int(vals[i]), t)
    with open(path, "wb") as fh:
        fh.write(text.tobytes())
        fh.write(binh.tobytes())
        for _ in range(n_extended_text):
            fh.write(text.tobytes())
        fh.write(rec.tobytes())
