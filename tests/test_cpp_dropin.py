"""The C++ drop-in header used from a compiled program (tests/cpp/
dropin_example.cpp, built by build()).  `info` runs on the CPU; `ping` runs on
the GPU and must equal the same ping computed through the Python binding."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import rtac_simulation_amd as R
from rtac_simulation_amd import build as B
from rtac_simulation_amd import configs

GOLDEN = Path(__file__).resolve().parent / "golden" / "reference_configs"


def _run(*args):
    assert B.EXAMPLE_BIN.exists(), "run __graft_entry__.build() first"
    return subprocess.run([str(B.EXAMPLE_BIN), *map(str, args)], capture_output=True, text=True,
                          timeout=120)


def test_info_workload_and_reference_configs():
    r = _run("info", R.CONFIG_DIR / "fls_512x1024.yaml")
    assert r.returncode == 0, r.stderr
    assert "ranges=1024" in r.stdout and "bearings=512" in r.stdout and "taps=31x17" in r.stdout
    r = _run("info", GOLDEN / "EA400_38KHz_receiver.yaml")
    assert r.returncode == 0 and "kind=1" in r.stdout and "ranges=1049" in r.stdout


def test_errors_map_to_reference_exceptions(tmp_path):
    assert _run("info", "/nonexistent.yaml").returncode == 4          # FileError
    p = tmp_path / "bad.yaml"
    p.write_text("type: sidescan\n")
    r = _run("info", p)
    assert r.returncode == 3 and "ConfigError" in r.stderr             # ConfigError


@pytest.mark.gpu
@pytest.mark.parametrize("form", [0, 2])
def test_cpp_ping_equals_python_binding(gpu, tmp_path, form):
    R_ = gpu
    out = tmp_path / "ping.bin"
    n, seed = 40_000, 77
    sink = tmp_path / "sink" / "ping.rtpg"
    r = _run("ping", R_.CONFIG_DIR / "fls_512x1024.yaml", n, seed, form, out, sink)
    assert r.returncode == 0, r.stderr
    cpp = np.fromfile(out, np.complex64).reshape(1024, 512)
    assert "pings=1" in r.stdout
    assert np.array_equal(R_.read_ping(sink, 0)["data"], cpp)
    cfg = R_.config_from_yaml(R_.CONFIG_DIR / "fls_512x1024.yaml")
    s = R_.make_sensor(cfg)
    s.set_sample_count(n)
    s.write_samples(R_.generate_hits(configs.fls_scene(), seed, n).to_device())
    s.compute_output(form)
    assert O.rel_l2(cpp, s.output()) < 1e-6
