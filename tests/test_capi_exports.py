"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute call: these run without a GPU)."""
import ctypes
import re
from pathlib import Path

import rtac_simulation_amd as R

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(rtsim_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(R.LIB_PATH))
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_declared_symbols():
    declared = declared_functions()
    bound = set(R.exported_names())
    assert declared <= bound, f"not bound in python: {sorted(declared - bound)}"


def test_version_and_errors_without_gpu():
    L = R.lib()
    assert L.rtsim_version().decode().startswith("rtsim")
    # invalid arguments are reported, not crashed on
    rc = L.rtsim_sensor_create(None, ctypes.byref(ctypes.c_void_p()))
    assert rc == R.ERR_INVALID
    assert "desc" in L.rtsim_last_error().decode()


def test_generator_is_deterministic_and_splittable():
    from rtac_simulation_amd import configs
    cfg, scene, S, seed = configs.workload("fls_512x1024", samples=10_000)
    a = R.generate_hits(scene, seed, 10_000)
    b = R.generate_hits(scene, seed, 4_000, first=6_000, nthreads=3)
    assert (a.range[6000:] == b.range).all()
    assert (a.amplitude[12000:] == b.amplitude).all()
    c = R.generate_hits(scene, seed + 1, 10_000)
    assert not (a.range == c.range).all()
