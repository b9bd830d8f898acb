"""Loader of the golden fixtures written by make_golden.py."""
from __future__ import annotations

from pathlib import Path

import numpy as np

import rtac_simulation_amd as R
from rtac_simulation_amd import configs

HERE = Path(__file__).resolve().parent
NAMES = ("single_beam_38k_2000", "fls_48x96_8000", "fls_32x64_ties_3000")


def load(name: str):
    """(SensorConfig, HostHits, dict of expected arrays) of one fixture."""
    z = np.load(HERE / f"{name}.npz", allow_pickle=False)
    cfg = configs.SensorConfig(
        kind=int(z["kind"]), range_min=float(z["range_min"]), range_max=float(z["range_max"]),
        range_count=int(z["range_count"]), psf=z["psf"],
        psf_range_span=float(z["psf_range_span"]), psf_bearing_span=float(z["psf_bearing_span"]),
        margin=float(z["margin"]),
        bearings=z["bearings"] if z["bearings"].size else None,
        bearing_axis=int(z["bearing_axis"]),
        directivity=z["directivity"] if z["directivity"].size else None,
        dir_xscale=float(z["dir_xscale"]), dir_yscale=float(z["dir_yscale"]),
        frequency=float(z["frequency"]), sound_speed=float(z["sound_speed"]),
        spreading=int(z["spreading"]), rotation=z["rotation"], name=name)
    h = R.HostHits(len(z["hit_range"]))
    h.range[:] = z["hit_range"]
    h.dir_x[:] = z["hit_dir_x"]
    h.dir_y[:] = z["hit_dir_y"]
    h.dir_z[:] = z["hit_dir_z"]
    h.amplitude[:] = z["hit_amplitude"]
    keys = ("records64", "records32", "order", "keys", "segments", "windows", "overlap", "ping")
    return cfg, h, {k: z[k] for k in keys}
