"""Write golden input/output vectors of the reference's own Kirchhoff PSTM
function, ``migrate_kirchhoff_full`` (processors/migration/kirchhoff_kernel.py:459-854).

Run in a container that holds the SeisProc tree (read-only) at REFERENCE:

    python tests/golden/make_reference_vectors.py [REFERENCE]

The end-to-end image of that function (tile pre-filter, epsilons, clamp to
n - 1.001, masks and weights acting together) is the one part of the path
that no known answer of the reference's tests fixes, so it is pinned by
running the function itself, on the CPU, on small seeded inputs.  Only that
module (and geometry_preprocessor.py, which it imports at module level) is
loaded by file path; bare package objects are registered for ``processors``
and ``processors.migration`` so that the package ``__init__`` files, which
import every other processor and their third-party dependencies, are not run.

Output: ``tests/golden/ref_migrate_kirchhoff_full.npz`` -- for each case the
inputs exactly as handed to the reference (FP32 traces as (n_samples,
n_traces), FP32 coordinates, the output grid, the depth axis and the scalar
arguments) and its outputs (image, fold as (n_depths, n_il, n_xl) FP32).
Nothing at test time imports or reads the reference; the tests read only this
file.
"""
from __future__ import annotations

import importlib.util
import json
import sys
import types
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
OUT = HERE / "ref_migrate_kirchhoff_full.npz"


def _load_reference(ref: Path):
    for pkg, sub in (("processors", "processors"), ("processors.migration", "processors/migration")):
        if pkg not in sys.modules:
            m = types.ModuleType(pkg)
            m.__path__ = [str(ref / sub)]
            sys.modules[pkg] = m
    mods = {}
    for name in ("geometry_preprocessor", "kirchhoff_kernel"):
        full = f"processors.migration.{name}"
        spec = importlib.util.spec_from_file_location(full, ref / "processors" / "migration" / f"{name}.py")
        mod = importlib.util.module_from_spec(spec)
        sys.modules[full] = mod
        spec.loader.exec_module(mod)
        mods[name] = mod
    return mods["kirchhoff_kernel"]


def _ricker(t, f):
    a = (np.pi * f * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def _survey(rng, n_mid, mid_lo, mid_hi, offsets, n_samples, dt, v, line=False):
    """Sources/receivers around random midpoints, Ricker events at the exact
    DSR times of a few point diffractors plus one flat reflector, noise."""
    mx = rng.uniform(mid_lo[0], mid_hi[0], n_mid)
    my = np.zeros(n_mid) if line else rng.uniform(mid_lo[1], mid_hi[1], n_mid)
    sx, sy, rx, ry = [], [], [], []
    for o in offsets:
        az = np.zeros(n_mid) if line else rng.uniform(0.0, np.pi, n_mid)
        hx, hy = 0.5 * o * np.cos(az), 0.5 * o * np.sin(az)
        sx.append(mx - hx)
        sy.append(my - hy)
        rx.append(mx + hx)
        ry.append(my + hy)
    sx, sy, rx, ry = (np.concatenate(a) for a in (sx, sy, rx, ry))
    n_tr = sx.size
    t = np.arange(n_samples) * dt
    tmax = t[-1]
    tr = np.zeros((n_tr, n_samples))
    for _ in range(4):
        px = rng.uniform(mid_lo[0], mid_hi[0])
        py = 0.0 if line else rng.uniform(mid_lo[1], mid_hi[1])
        z = 0.5 * v * rng.uniform(0.25 * tmax, 0.8 * tmax)
        amp = rng.normal()
        tt = (np.sqrt((sx - px) ** 2 + (sy - py) ** 2 + z * z) + np.sqrt((rx - px) ** 2 + (ry - py) ** 2 + z * z)) / v
        tr += amp * _ricker(t[None, :] - tt[:, None], 25.0)
    t0 = 0.5 * tmax
    off = np.hypot(rx - sx, ry - sy)
    tr += 0.5 * _ricker(t[None, :] - np.sqrt(t0 ** 2 + (off / v) ** 2)[:, None], 25.0)
    tr += rng.normal(scale=0.05 * np.abs(tr).max(), size=tr.shape)
    return tr.astype(np.float32), (sx, sy, rx, ry)


CASES = [
    # name, seed, geometry, record, image, migration arguments
    dict(name="grid3d", seed=101, n_mid=60, mid=((-100.0, -100.0), (325.0, 275.0)),
         offsets=(0.0, 300.0, 600.0, 900.0), n_samples=201, dt_ms=4.0, v=2000.0,
         n_il=12, n_xl=10, d_il=25.0, d_xl=25.0, x0=10.0, y0=-5.0, tau0_ms=0.0, dtau_ms=4.0, n_tau=201,
         t_min_ms=0.0, aperture=450.0, angle=60.0, time_aperture=False),
    # three reference tiles of 100 points that end mid-row; time-dependent aperture
    dict(name="grid3d_wide_time_aperture", seed=404, n_mid=70, mid=((-150.0, -150.0), (450.0, 350.0)),
         offsets=(100.0, 500.0, 1100.0), n_samples=176, dt_ms=4.0, v=2200.0,
         n_il=16, n_xl=13, d_il=25.0, d_xl=20.0, x0=-20.0, y0=15.0, tau0_ms=0.0, dtau_ms=4.0, n_tau=176,
         t_min_ms=0.0, aperture=400.0, angle=75.0, time_aperture=True),
    # record shorter than the traveltimes of deep / far pairs: clamp to n - 1.001
    dict(name="line2d_clamp", seed=202, n_mid=40, mid=((0.0, 0.0), (500.0, 0.0)),
         offsets=(0.0, 250.0, 500.0, 750.0, 1000.0, 1250.0), n_samples=126, dt_ms=4.0, v=1800.0,
         n_il=41, n_xl=1, d_il=12.5, d_xl=12.5, x0=0.0, y0=0.0, tau0_ms=0.0, dtau_ms=4.0, n_tau=201,
         t_min_ms=0.0, aperture=800.0, angle=70.0, time_aperture=False, line=True),
    # record start after time zero, output times off the input sample grid
    dict(name="grid3d_tmin", seed=303, n_mid=50, mid=((-50.0, -50.0), (200.0, 200.0)),
         offsets=(200.0, 700.0), n_samples=251, dt_ms=2.0, v=2500.0,
         n_il=8, n_xl=7, d_il=20.0, d_xl=20.0, x0=0.0, y0=0.0, tau0_ms=50.0, dtau_ms=3.0, n_tau=167,
         t_min_ms=40.0, aperture=600.0, angle=45.0, time_aperture=False),
]


def main(ref: Path) -> None:
    import torch

    kk = _load_reference(ref)
    torch.set_num_threads(1)
    arrays, meta = {}, []
    for c in CASES:
        rng = np.random.default_rng(c["seed"])
        tr, (sx, sy, rx, ry) = _survey(rng, c["n_mid"], c["mid"][0], c["mid"][1], c["offsets"], c["n_samples"],
                                       c["dt_ms"] / 1000.0, c["v"], line=c.get("line", False))
        il = np.arange(c["n_il"])
        xl = np.arange(c["n_xl"])
        ox = (c["x0"] + il[:, None] * c["d_il"] + 0.0 * xl[None, :]).reshape(-1).astype(np.float32)
        oy = (c["y0"] + 0.0 * il[:, None] + xl[None, :] * c["d_xl"]).reshape(-1).astype(np.float32)
        tau_s = (c["tau0_ms"] + np.arange(c["n_tau"]) * c["dtau_ms"]) / 1000.0
        depth = (c["v"] * tau_s / 2.0).astype(np.float32)
        f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32))
        traces_ref = np.ascontiguousarray(tr.T)  # (n_samples, n_traces), the reference layout
        img, fold = kk.migrate_kirchhoff_full(
            f32(traces_ref), f32(sx), f32(sy), f32(rx), f32(ry), f32(ox), f32(oy), f32(depth),
            float(c["v"]), float(c["dt_ms"]), float(c["t_min_ms"]), float(c["aperture"]), float(c["angle"]),
            int(c["n_il"]), int(c["n_xl"]), use_time_dependent_aperture=bool(c["time_aperture"]))
        n = c["name"]
        arrays[f"{n}/traces"] = traces_ref
        for k, a in (("sx", sx), ("sy", sy), ("rx", rx), ("ry", ry), ("output_x", ox), ("output_y", oy),
                     ("depth", depth)):
            arrays[f"{n}/{k}"] = np.asarray(a, dtype=np.float32)
        arrays[f"{n}/image"] = img.numpy().astype(np.float32)
        arrays[f"{n}/fold"] = fold.numpy().astype(np.float32)
        meta.append({k: c[k] for k in ("name", "seed", "v", "dt_ms", "t_min_ms", "aperture", "angle", "n_il",
                                       "n_xl", "d_il", "d_xl", "x0", "y0", "tau0_ms", "dtau_ms", "n_tau",
                                       "time_aperture")})
        print(f"{n}: traces {traces_ref.shape}, image {tuple(img.shape)}, contributions {float(fold.sum()):.0f}")
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1024:.0f} KiB)")


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
