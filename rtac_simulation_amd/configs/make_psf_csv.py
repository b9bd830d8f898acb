"""Writes the perturbed PSF tap files the workload YAML configs point to
(psf_31x17.csv, psf_63x33.csv) from rtac_simulation_amd.configs."""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from rtac_simulation_amd import configs  # noqa: E402


def write(path: Path, taps: np.ndarray) -> None:
    t = taps.astype(np.complex64)
    with open(path, "w") as f:
        for row in t:
            f.write(",".join(f"{np.float32(v.real)!r},{np.float32(v.imag)!r}"
                             .replace("np.float32(", "").replace(")", "") for v in row) + "\n")


if __name__ == "__main__":
    write(HERE / "psf_31x17.csv", configs.psf_31x17())
    write(HERE / "psf_63x33.csv", configs.psf_63x33())
