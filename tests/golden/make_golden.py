"""Writes the golden fixtures under tests/golden/ (run on the CPU):

    python tests/golden/make_golden.py

Each fixture is a compressed .npz (no pickles) holding one complete case:
the sensor description (scalars and tables), the ray hits, and the expected
results computed by the FP64 oracle:
  records64      streamed records (directivity x round-trip phase), FP64
  records32      the same rounded to FP32 -- the binning and ping inputs
  order, keys, segments, windows, overlap
                 the Binner chain of records32 (integers, bit-exact)
  ping           the FP64 plain-form ping of records32 (complex128)
The cases are small enough for the oracle to finish in seconds.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle as O  # noqa: E402
import rtac_simulation_amd as R  # noqa: E402
from rtac_simulation_amd import configs  # noqa: E402

HERE = Path(__file__).resolve().parent

SCALARS = ("kind", "range_min", "range_max", "range_count", "psf_range_span",
           "psf_bearing_span", "margin", "bearing_axis", "dir_xscale", "dir_yscale",
           "frequency", "sound_speed", "spreading")


def cases():
    # 1D single beam (the 38 kHz echo sounder) with a small piston table
    cfg1, scene1, _, seed1 = configs.single_beam_38k()
    lam = configs.SOUND_SPEED / cfg1.frequency
    cfg1.directivity = R.directivity_disk(0.3, lam, True, 1)
    yield "single_beam_38k_2000", cfg1, R.generate_hits(scene1, seed1, 2000, nthreads=4)
    # 2D: 48 bearings x 96 ranges, 31 x 17 PSF, seabed scene
    cfg2 = configs._fls(48, 96, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "fls_48x96")
    yield "fls_48x96_8000", cfg2, R.generate_hits(configs.fls_scene(), 21, 8000, nthreads=4)
    # 2D with nulls and exact travel ties (identical hits repeated)
    cfg3 = configs._fls(32, 64, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "fls_32x64_ties")
    h = R.generate_hits(configs.fls_scene(), 22, 3000, nthreads=4)
    idx = np.random.default_rng(0).integers(0, h.n, 1000)
    for name in ("range", "dir_x", "dir_y", "dir_z"):
        a = getattr(h, name)
        a[2000:3000] = a[idx]
    h.amplitude[4000:6000] = h.amplitude.reshape(-1, 2)[idx].ravel()
    h.range[::97] = 0.0          # r < 1e-4: null samples
    yield "fls_32x64_ties_3000", cfg3, h


def main():
    for name, cfg, hits in cases():
        rec64 = O.stream(cfg, hits)
        rec32 = rec64.astype(np.float32)
        b = O.binning(cfg, rec32)
        ping = O.conv(cfg, b["sorted"], b["windows"])
        arrays = {k: np.asarray(getattr(cfg, k)) for k in SCALARS}
        arrays.update(
            psf=cfg.psf.astype(np.complex64), rotation=np.asarray(cfg.rotation, np.float32),
            bearings=np.asarray(cfg.bearings if cfg.bearings is not None else [], np.float32),
            directivity=np.asarray(cfg.directivity if cfg.directivity is not None
                                   else np.zeros((0, 0)), np.float32),
            hit_range=hits.range, hit_dir_x=hits.dir_x, hit_dir_y=hits.dir_y,
            hit_dir_z=hits.dir_z, hit_amplitude=hits.amplitude,
            records64=rec64, records32=rec32, order=b["order"].astype(np.uint32),
            keys=b["keys"].astype(np.uint32), segments=b["segments"].astype(np.uint32),
            windows=b["windows"].astype(np.uint32), overlap=np.asarray(b["overlap"]),
            ping=ping)
        path = HERE / f"{name}.npz"
        np.savez_compressed(path, **arrays)
        print(f"wrote {path.name}: {hits.n} samples, {path.stat().st_size / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
