"""Minimal Zarr (format 2) array store: uncompressed, C order, one file per
chunk.

The reference writes its migrated image with the ``zarr`` package
(utils/migration_output_storage.py:163-171: ``zarr.open(..., shape=(n_time,
n_traces), chunks=(n_time, chunk), dtype=float32, compressor=None,
zarr_format=2)``).  That package is not a dependency here, so this module
writes the same on-disk format directly: a directory holding a ``.zarray``
JSON document and one raw little-endian file per chunk named by its chunk
coordinates joined with ``.``; edge chunks are stored at full chunk shape,
padded with the fill value.  Any Zarr v2 reader opens the result.
"""
from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Sequence, Tuple

import numpy as np


class ZarrV2Array:
    """Uncompressed Zarr v2 array on a local directory."""

    def __init__(self, path, shape: Sequence[int], chunks: Sequence[int], dtype="<f4", fill_value=0.0):
        self.path = Path(path)
        self.shape = tuple(int(s) for s in shape)
        self.chunks = tuple(int(c) for c in chunks)
        self.dtype = np.dtype(dtype).newbyteorder("<")
        self.fill_value = fill_value
        if len(self.shape) != len(self.chunks) or any(c <= 0 for c in self.chunks):
            raise ValueError("chunks must be positive and match the array rank")

    # -- creation / opening
    @classmethod
    def create(cls, path, shape, chunks, dtype="<f4", fill_value=0.0) -> "ZarrV2Array":
        a = cls(path, shape, chunks, dtype, fill_value)
        a.path.mkdir(parents=True, exist_ok=True)
        for f in a.path.iterdir():
            if f.is_file():
                f.unlink()
        meta = {"zarr_format": 2, "shape": list(a.shape), "chunks": list(a.chunks), "dtype": a.dtype.str,
                "compressor": None, "fill_value": fill_value, "order": "C", "filters": None}
        (a.path / ".zarray").write_text(json.dumps(meta, indent=4))
        return a

    @classmethod
    def open(cls, path) -> "ZarrV2Array":
        meta = json.loads((Path(path) / ".zarray").read_text())
        if meta.get("zarr_format") != 2 or meta.get("compressor") is not None or meta.get("order") != "C" \
                or meta.get("filters"):
            raise ValueError("only uncompressed, unfiltered C-order Zarr v2 arrays are supported")
        return cls(path, meta["shape"], meta["chunks"], meta["dtype"], meta.get("fill_value", 0.0))

    # -- chunk I/O
    def n_chunks(self) -> Tuple[int, ...]:
        return tuple(-(-s // c) for s, c in zip(self.shape, self.chunks))

    def _chunk_file(self, idx: Sequence[int]) -> Path:
        return self.path / ".".join(str(i) for i in idx)

    def write_chunk(self, idx: Sequence[int], block: np.ndarray) -> None:
        """Write chunk `idx`; `block` may be the clipped edge shape."""
        full = np.full(self.chunks, self.fill_value, dtype=self.dtype)
        full[tuple(slice(0, b) for b in block.shape)] = block
        tmp = self._chunk_file(idx).with_suffix(".tmp")
        with open(tmp, "wb") as fh:
            fh.write(full.tobytes(order="C"))
        os.replace(tmp, self._chunk_file(idx))

    def read_chunk(self, idx: Sequence[int]) -> np.ndarray:
        f = self._chunk_file(idx)
        if not f.exists():
            return np.full(self.chunks, self.fill_value, dtype=self.dtype)
        return np.frombuffer(f.read_bytes(), dtype=self.dtype).reshape(self.chunks)

    # -- whole-array helpers (2D: the image layout (n_time, n_traces))
    def write_columns(self, col0: int, block: np.ndarray) -> None:
        """Write block (n_rows, n) at columns [col0, col0 + n) of a 2D array
        whose row chunk is the full height; col0 and n must be chunk-aligned
        except at the array end."""
        if len(self.shape) != 2 or self.chunks[0] != self.shape[0]:
            raise ValueError("write_columns needs a 2D array chunked over columns only")
        cw = self.chunks[1]
        if col0 % cw:
            raise ValueError("col0 must be a multiple of the column chunk")
        n = block.shape[1]
        for j0 in range(0, n, cw):
            j1 = min(j0 + cw, n)
            if j1 - j0 < cw and col0 + j1 != self.shape[1]:
                raise ValueError("partial chunk before the array end")
            self.write_chunk((0, (col0 + j0) // cw), block[:, j0:j1])

    def read(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        for idx in np.ndindex(*self.n_chunks()):
            sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, self.chunks, self.shape))
            blk = self.read_chunk(idx)
            out[sl] = blk[tuple(slice(0, x.stop - x.start) for x in sl)]
        return out
