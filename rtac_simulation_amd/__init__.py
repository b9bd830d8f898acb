"""rtac_simulation_amd: MI355X-native sample-to-ping path.

Python binding (ctypes) of the C ABI in ``include/rtsim.h`` /
``include/rtsim_scene.h``, used by the tests and ``bench.py``.  The compute
path is the in-tree HIP library ``librtsim.so`` built for gfx950; importing
this package never falls back to anything else: if the library is missing
every call raises.

Class names follow the reference's sensor-instance API
(``SensorInstance1D_Complex`` / ``SensorInstance2D_Complex``,
SensorInstance1D.h:72-110, SensorInstance2D.h:153-205).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "librtsim.so"

RTSIM_OK = 0
ERR_INVALID, ERR_HIP, ERR_UNSUPPORTED, ERR_CONFIG, ERR_IO = 1, 2, 3, 4, 5
SENSOR_1D, SENSOR_2D = 1, 2
FORM_PLAIN_FP32, FORM_MFMA_BF16, FORM_MFMA_FP16 = 0, 1, 2
FORM_NAMES = {FORM_PLAIN_FP32: "plain_fp32", FORM_MFMA_BF16: "mfma_bf16",
              FORM_MFMA_FP16: "mfma_fp16"}
AXIS_ANGLE, AXIS_SINE = 0, 1
SPREAD_NONE, SPREAD_INV_R2 = 0, 1
SCENE_CONE, SCENE_SEABED = 1, 2
KEY_OUT_OF_RANGE = 0xFFFFFFFF


class RtsimError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rtsim error {code}: {msg}")
        self.code = code


class WaveformSine(C.Structure):
    _fields_ = [("frequency", C.c_float), ("duration", C.c_float), ("fixed", C.c_int),
                ("oversampling", C.c_uint32), ("beam", C.POINTER(C.c_float)),
                ("beam_taps", C.c_uint32)]


class SensorDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("range_min", C.c_float), ("range_max", C.c_float), ("range_count", C.c_uint32),
        ("bearings", C.POINTER(C.c_float)), ("bearing_count", C.c_uint32),
        ("bearing_axis", C.c_int),
        ("psf", C.POINTER(C.c_float)),
        ("psf_range_taps", C.c_uint32), ("psf_bearing_taps", C.c_uint32),
        ("psf_range_span", C.c_float), ("psf_bearing_span", C.c_float),
        ("margin", C.c_float),
        ("directivity", C.POINTER(C.c_float)),
        ("dir_width", C.c_uint32), ("dir_height", C.c_uint32),
        ("dir_xscale", C.c_float), ("dir_yscale", C.c_float),
        ("frequency", C.c_float), ("sound_speed", C.c_float),
        ("spreading", C.c_int),
        ("rotation", C.c_float * 9),
        ("waveform", C.POINTER(WaveformSine)),
    ]


class Hits(C.Structure):
    _fields_ = [("range", C.c_void_p), ("dir_x", C.c_void_p), ("dir_y", C.c_void_p),
                ("dir_z", C.c_void_p), ("amplitude", C.c_void_p)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("range_min", C.c_float), ("range_max", C.c_float),
        ("off_axis_max", C.c_float),
        ("bearing_max", C.c_float), ("elevation_max", C.c_float),
        ("plane_count", C.c_uint32), ("planes", (C.c_float * 4) * 4),
        ("height_noise", C.c_float),
        ("sphere_count", C.c_uint32),
        ("sphere_radius_min", C.c_float), ("sphere_radius_max", C.c_float),
        ("sphere_range_min", C.c_float), ("sphere_range_max", C.c_float),
        ("sphere_fraction", C.c_float), ("dynamic_range_db", C.c_float),
        ("rayleigh_sigma", C.c_float), ("incidence_power", C.c_int),
    ]


# (name, restype, argtypes) of every exported entry point
_SIGNATURES = [
    ("rtsim_last_error", C.c_char_p, []),
    ("rtsim_version", C.c_char_p, []),
    ("rtsim_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("rtsim_stream_sync", C.c_int, [C.c_void_p]),
    ("rtsim_sensor_create", C.c_int, [C.POINTER(SensorDesc), C.POINTER(C.c_void_p)]),
    ("rtsim_sensor_destroy", C.c_int, [C.c_void_p]),
    ("rtsim_sensor_set_ranges", C.c_int,
     [C.c_void_p, C.c_float, C.c_float, C.c_uint32, C.c_float]),
    ("rtsim_sensor_set_sample_count", C.c_int, [C.c_void_p, C.c_size_t]),
    ("rtsim_sensor_sample_count", C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    ("rtsim_sensor_samples", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("rtsim_receiver_write_samples", C.c_int,
     [C.c_void_p, C.POINTER(Hits), C.c_size_t, C.c_size_t, C.c_void_p]),
    ("rtsim_sensor_compute", C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ("rtsim_measure_peaks", C.c_int,
     [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("rtsim_sensor_compute_noalloc", C.c_int,
     [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]),
    ("rtsim_sensor_autotune", C.c_int,
     [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    ("rtsim_sensor_set_tile_shape", C.c_int, [C.c_void_p, C.c_int]),
    ("rtsim_sensor_set_tile_waves", C.c_int, [C.c_void_p, C.c_int]),
    ("rtsim_sensor_tile_waves", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("rtsim_sensor_set_pose", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("rtsim_sensor_tile_shape", C.c_int,
     [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("rtsim_sensor_output", C.c_int,
     [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("rtsim_sensor_sorted_samples", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("rtsim_sensor_binner_debug", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int),
      C.c_void_p]),
    ("rtsim_sensor_upload_samples", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    ("rtsim_sensor_download_samples", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]),
    ("rtsim_sensor_download_output", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtsim_sensor_set_profiling", C.c_int, [C.c_void_p, C.c_int]),
    ("rtsim_sensor_stage_times", C.c_int,
     [C.c_void_p, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int)]),
    ("rtsim_stream_create", C.c_int, [C.POINTER(C.c_void_p)]),
    ("rtsim_stream_destroy", C.c_int, [C.c_void_p]),
    ("rtsim_timer_create", C.c_int, [C.POINTER(C.c_void_p)]),
    ("rtsim_timer_destroy", C.c_int, [C.c_void_p]),
    ("rtsim_timer_start", C.c_int, [C.c_void_p, C.c_void_p]),
    ("rtsim_timer_stop", C.c_int, [C.c_void_p, C.c_void_p]),
    ("rtsim_timer_elapsed", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("rtsim_device_alloc", C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    ("rtsim_device_free", C.c_int, [C.c_void_p]),
    ("rtsim_memcpy_h2d", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("rtsim_memcpy_d2h", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("rtsim_generate_hits", C.c_int,
     [C.POINTER(SceneDesc), C.c_uint64, C.c_uint64, C.c_size_t, C.c_void_p, C.c_void_p,
      C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    ("rtsim_directivity_disk", C.c_int,
     [C.c_float, C.c_float, C.c_int, C.c_uint32, C.c_void_p, C.c_size_t,
      C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("rtsim_directivity_rectangle", C.c_int,
     [C.c_float, C.c_float, C.c_float, C.c_int, C.c_uint32, C.c_void_p, C.c_size_t,
      C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("rtsim_directivity_uniform", C.c_int,
     [C.c_uint32, C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("rtsim_emitter_polar_directions", C.c_int,
     [C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("rtsim_emitter_initial_values", C.c_int,
     [C.c_void_p, C.c_size_t, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
      C.c_void_p]),
    ("rtsim_generate_emitter_hits", C.c_int,
     [C.POINTER(SceneDesc), C.c_uint64, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
      C.c_uint64, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
      C.c_int]),
    ("rtsim_info_from_yaml", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("rtsim_info_desc", C.c_int, [C.c_void_p, C.POINTER(C.POINTER(SensorDesc))]),
    ("rtsim_info_destroy", C.c_int, [C.c_void_p]),
    ("rtsim_sensor_create_from_yaml", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("rtsim_config_add_search_path", C.c_int, [C.c_char_p]),
    ("rtsim_config_find", C.c_int, [C.c_char_p, C.c_char_p, C.c_size_t]),
    ("rtsim_batch_create", C.c_int,
     [C.POINTER(C.c_void_p), C.POINTER(C.c_int), C.c_uint32, C.POINTER(C.c_void_p)]),
    ("rtsim_batch_bind_hits", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(Hits), C.c_size_t]),
    ("rtsim_batch_compute", C.c_int, [C.c_void_p, C.c_void_p]),
    ("rtsim_batch_autotune", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_float)]),
    ("rtsim_batch_set_persistent", C.c_int, [C.c_void_p, C.c_int]),
    ("rtsim_batch_persistent", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("rtsim_batch_check", C.c_int, [C.c_void_p]),
    ("rtsim_batch_destroy", C.c_int, [C.c_void_p]),
    # ping files (rtsim_sink.h)
    ("rtsim_sink_open", C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]),
    ("rtsim_sink_write", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rtsim_sink_write_host", C.c_int,
     [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_void_p,
      C.c_void_p]),
    ("rtsim_sink_count", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    ("rtsim_sink_close", C.c_int, [C.c_void_p]),
    ("rtsim_ping_file_count", C.c_int, [C.c_char_p, C.POINTER(C.c_uint64)]),
    ("rtsim_ping_file_read", C.c_int, [C.c_char_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]),
]

_lib = None


def lib() -> C.CDLL:
    """Load ``librtsim.so`` (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in _SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_names() -> list[str]:
    return [name for name, _, _ in _SIGNATURES]


def check(rc: int) -> None:
    if rc != RTSIM_OK:
        raise RtsimError(rc, lib().rtsim_last_error().decode())


def device_count() -> int:
    c = C.c_int(0)
    check(lib().rtsim_device_count(C.byref(c)))
    return c.value


def measure_peaks(stream: "Stream | None" = None) -> dict:
    """This GPU's own HBM copy bandwidth (GB/s, read + write) and dense MFMA
    rates (TFLOP/s, FP16 and BF16) from rtsim_measure_peaks."""
    h, f16, b16 = C.c_double(0), C.c_double(0), C.c_double(0)
    check(lib().rtsim_measure_peaks(stream.handle if stream else None, C.byref(h), C.byref(f16),
                                    C.byref(b16)))
    return {"hbm_copy_gbs": h.value, "mfma_fp16_tflops": f16.value, "mfma_bf16_tflops": b16.value}


def _ptr(a: np.ndarray) -> C.c_void_p:
    assert a.flags.c_contiguous
    return C.c_void_p(a.ctypes.data)


def _fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


class DeviceBuffer:
    """A hipMalloc'd buffer owned through the C ABI."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().rtsim_device_alloc(C.byref(p), max(1, self.nbytes)))
        self.ptr = p

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(lib().rtsim_memcpy_h2d(b.ptr, _ptr(a), a.nbytes))
        return b

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().rtsim_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().rtsim_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = C.c_void_p()
        check(lib().rtsim_stream_create(C.byref(p)))
        self.handle = p

    def sync(self):
        check(lib().rtsim_stream_sync(self.handle))

    def __del__(self):
        try:
            lib().rtsim_stream_destroy(self.handle)
        except Exception:
            pass


class Timer:
    """hipEvent pair recorded on a stream."""

    def __init__(self):
        p = C.c_void_p()
        check(lib().rtsim_timer_create(C.byref(p)))
        self.handle = p

    def start(self, stream=None):
        check(lib().rtsim_timer_start(self.handle, stream.handle if stream else None))

    def stop(self, stream=None):
        check(lib().rtsim_timer_stop(self.handle, stream.handle if stream else None))

    def elapsed_ms(self) -> float:
        ms = C.c_float()
        check(lib().rtsim_timer_elapsed(self.handle, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            lib().rtsim_timer_destroy(self.handle)
        except Exception:
            pass


# ----------------------------------------------------------------------------
# synthetic hits
# ----------------------------------------------------------------------------
class HostHits:
    """SoA ray hits on the host."""

    def __init__(self, n: int):
        self.n = int(n)
        self.range = np.zeros(n, np.float32)
        self.dir_x = np.zeros(n, np.float32)
        self.dir_y = np.zeros(n, np.float32)
        self.dir_z = np.zeros(n, np.float32)
        self.amplitude = np.zeros(2 * n, np.float32)

    def to_device(self) -> "DeviceHits":
        return DeviceHits(self)


class DeviceHits:
    def __init__(self, h: HostHits):
        self.n = h.n
        self.bufs = [DeviceBuffer.from_array(a) for a in
                     (h.range, h.dir_x, h.dir_y, h.dir_z, h.amplitude)]
        self.struct = Hits(*[b.ptr for b in self.bufs])

    def prefix(self, n: int) -> "DeviceHits":
        """The first n hits, sharing this object's device buffers (with the
        counter-based generator these are exactly the hits a run of n gives)."""
        v = DeviceHits.__new__(DeviceHits)
        v.n = int(min(n, self.n))
        v.bufs = self.bufs
        v.struct = self.struct
        return v


def generate_hits(scene: SceneDesc, seed: int, n: int, first: int = 0,
                  nthreads: int = 8) -> HostHits:
    h = HostHits(n)
    check(lib().rtsim_generate_hits(C.byref(scene), seed, first, n, _ptr(h.range),
                                    _ptr(h.dir_x), _ptr(h.dir_y), _ptr(h.dir_z),
                                    _ptr(h.amplitude), nthreads))
    return h


def emitter_polar_directions(resolution_deg: float, bearing_aperture_deg: float,
                             elevation_aperture_deg: float) -> np.ndarray:
    """Emitter::generate_polar_directions (Emitter.cu:59-80): (rays, 3) float32."""
    cnt = C.c_size_t()
    check(lib().rtsim_emitter_polar_directions(resolution_deg, bearing_aperture_deg,
                                               elevation_aperture_deg, None, 0, C.byref(cnt)))
    out = np.zeros((cnt.value, 3), np.float32)
    check(lib().rtsim_emitter_polar_directions(resolution_deg, bearing_aperture_deg,
                                               elevation_aperture_deg, _ptr(out), cnt.value,
                                               C.byref(cnt)))
    return out


def emitter_initial_values(dirs: np.ndarray, table: np.ndarray | None = None,
                           xscale: float = 1.0 / np.pi, yscale: float = 1.0 / np.pi) -> np.ndarray:
    """Emitter::load_initial_values (Emitter.cu:38-55): D(dir) per ray, complex64."""
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    out = np.zeros(d.shape[0], np.complex64)
    if table is None:
        check(lib().rtsim_emitter_initial_values(_ptr(d), d.shape[0], None, 0, 0, 0.0, 0.0,
                                                 _ptr(out)))
    else:
        t = np.ascontiguousarray(table, np.float32)
        check(lib().rtsim_emitter_initial_values(_ptr(d), d.shape[0], _ptr(t), t.shape[1],
                                                 t.shape[0], xscale, yscale, _ptr(out)))
    return out


def generate_emitter_hits(scene: SceneDesc, seed: int, dirs: np.ndarray, values: np.ndarray,
                          n: int, first: int = 0, rotation: np.ndarray | None = None,
                          nthreads: int = 8) -> HostHits:
    """Hits of an emitter's rays (ray i along rotation @ dirs[i % rays], value values[i % rays])."""
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    v = np.ascontiguousarray(values, np.complex64)
    if v.size != d.shape[0]:
        raise ValueError("one initial value per emitter ray")
    rot = None if rotation is None else np.ascontiguousarray(rotation, np.float32).reshape(9)
    h = HostHits(n)
    check(lib().rtsim_generate_emitter_hits(C.byref(scene), seed, _ptr(d), _ptr(v), d.shape[0],
                                            None if rot is None else _ptr(rot), first, n,
                                            _ptr(h.range), _ptr(h.dir_x), _ptr(h.dir_y),
                                            _ptr(h.dir_z), _ptr(h.amplitude), nthreads))
    return h


def directivity_disk(diameter: float, wavelength: float, single_sided: bool = True,
                     oversampling: int = 8) -> np.ndarray:
    w, h = C.c_uint32(), C.c_uint32()
    check(lib().rtsim_directivity_disk(diameter, wavelength, int(single_sided), oversampling,
                                       None, 0, C.byref(w), C.byref(h)))
    out = np.zeros((h.value, w.value), np.float32)
    check(lib().rtsim_directivity_disk(diameter, wavelength, int(single_sided), oversampling,
                                       _ptr(out), out.size, C.byref(w), C.byref(h)))
    return out


def directivity_rectangle(width: float, height: float, wavelength: float,
                          single_sided: bool = True, oversampling: int = 8) -> np.ndarray:
    w, h = C.c_uint32(), C.c_uint32()
    check(lib().rtsim_directivity_rectangle(width, height, wavelength, int(single_sided),
                                            oversampling, None, 0, C.byref(w), C.byref(h)))
    out = np.zeros((h.value, w.value), np.float32)
    check(lib().rtsim_directivity_rectangle(width, height, wavelength, int(single_sided),
                                            oversampling, _ptr(out), out.size, C.byref(w),
                                            C.byref(h)))
    return out


# ----------------------------------------------------------------------------
# sensor instances
# ----------------------------------------------------------------------------
class SensorInstance:
    """Owner of one rtsim_sensor handle (SensorInstance, SensorInstance.h:21-72)."""

    def __init__(self, cfg: "SensorConfig"):
        self.cfg = cfg
        d = cfg.desc()
        h = C.c_void_p()
        check(lib().rtsim_sensor_create(C.byref(d), C.byref(h)))
        self.handle = h
        self.kind = cfg.kind

    def __del__(self):
        try:
            if self.handle:
                lib().rtsim_sensor_destroy(self.handle)
        except Exception:
            pass

    # -- producer side -----------------------------------------------------------
    def set_sample_count(self, n: int):
        check(lib().rtsim_sensor_set_sample_count(self.handle, n))

    def sample_count(self) -> int:
        c = C.c_size_t()
        check(lib().rtsim_sensor_sample_count(self.handle, C.byref(c)))
        return c.value

    def set_ranges(self, lo: float, hi: float, n: int, margin: float = -1.0):
        check(lib().rtsim_sensor_set_ranges(self.handle, lo, hi, n, margin))

    def write_samples(self, hits: DeviceHits, offset: int = 0, stream: Stream | None = None):
        check(lib().rtsim_receiver_write_samples(self.handle, C.byref(hits.struct), hits.n,
                                                 offset, stream.handle if stream else None))

    def upload_samples(self, rec: np.ndarray):
        rec = np.ascontiguousarray(rec, dtype=np.float32).reshape(-1, 4)
        check(lib().rtsim_sensor_upload_samples(self.handle, _ptr(rec), rec.shape[0], None))

    # -- consumer side -------------------------------------------------------------
    def compute_output(self, form: int = FORM_PLAIN_FP32, stream: Stream | None = None):
        check(lib().rtsim_sensor_compute(self.handle, form, stream.handle if stream else None))

    def autotune(self, form: int, iterations: int = 10, stream: Stream | None = None):
        """Pick the fastest MFMA correlation tile shape; returns (shape, ms)."""
        sh, ms = C.c_int(), C.c_float()
        check(lib().rtsim_sensor_autotune(self.handle, form, iterations,
                                          stream.handle if stream else None, C.byref(sh),
                                          C.byref(ms)))
        return sh.value, ms.value

    def set_pose(self, rotation: np.ndarray):
        """Receiver rotation R (3 x 3, row-major); directions map as R^T dir."""
        r = np.ascontiguousarray(np.asarray(rotation, np.float32).reshape(9))
        check(lib().rtsim_sensor_set_pose(self.handle, r.ctypes.data_as(C.POINTER(C.c_float))))

    def set_tile_shape(self, shape: int):
        check(lib().rtsim_sensor_set_tile_shape(self.handle, shape))

    def set_tile_waves(self, waves: int):
        check(lib().rtsim_sensor_set_tile_waves(self.handle, waves))

    def tile_waves(self) -> int:
        w = C.c_int()
        check(lib().rtsim_sensor_tile_waves(self.handle, C.byref(w)))
        return w.value

    def tile_shape(self):
        sh, r, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().rtsim_sensor_tile_shape(self.handle, C.byref(sh), C.byref(r), C.byref(c)))
        return sh.value, r.value, c.value

    def shape(self):
        h, w = C.c_uint32(), C.c_uint32()
        check(lib().rtsim_sensor_output(self.handle, None, C.byref(h), C.byref(w)))
        return h.value, w.value

    def output(self) -> np.ndarray:
        """The ping as complex64 [range][bearing] (1D: [range])."""
        H, W = self.shape()
        out = np.empty(2 * H * W, np.float32)
        check(lib().rtsim_sensor_download_output(self.handle, _ptr(out), None))
        c = out.view(np.complex64)
        return c.reshape(H, W) if self.kind == SENSOR_2D else c

    def samples(self, sorted: bool = False) -> np.ndarray:
        """The resident records, or (sorted) as the last sort left them: grouped
        by Binner key after a plain-form compute, by travel after binner_debug."""
        n = self.sample_count()
        out = np.empty((n, 4), np.float32)
        check(lib().rtsim_sensor_download_samples(self.handle, _ptr(out), n, int(sorted), None))
        return out

    def binner_debug(self) -> dict:
        n = self.sample_count()
        H, _ = self.shape()
        keys = np.empty(n, np.uint32)
        seg = np.empty(2 * H, np.uint32)
        win = np.empty(2 * H, np.uint32)
        order = np.empty(n, np.uint32)
        ov = C.c_int()
        check(lib().rtsim_sensor_binner_debug(self.handle, _ptr(keys), _ptr(seg), _ptr(win),
                                              _ptr(order), C.byref(ov), None))
        return {"keys": keys, "segments": seg.reshape(H, 2), "windows": win.reshape(H, 2),
                "order": order, "overlap": ov.value}

    def set_profiling(self, on: bool):
        check(lib().rtsim_sensor_set_profiling(self.handle, int(on)))

    STAGES = ("stream", "sort", "bins", "deposit", "convolution")

    def stage_times(self) -> tuple[dict, int]:
        """Mean ms per stage (HIP events on the compute stream) since profiling on."""
        ms = (C.c_float * 5)()
        n = C.c_int()
        check(lib().rtsim_sensor_stage_times(self.handle, ms, 5, C.byref(n)))
        return dict(zip(self.STAGES, list(ms))), n.value


class SensorInstance1D_Complex(SensorInstance):
    pass


class SensorInstance2D_Complex(SensorInstance):
    pass


def make_sensor(cfg: "SensorConfig") -> SensorInstance:
    return (SensorInstance2D_Complex if cfg.kind == SENSOR_2D else SensorInstance1D_Complex)(cfg)


class PingHeader(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("scalar", C.c_uint32), ("height", C.c_uint32),
                ("width", C.c_uint32), ("range_min", C.c_float), ("range_max", C.c_float),
                ("index", C.c_uint64), ("checksum", C.c_uint64)]


class FileSink:
    """FileSink drop-in (sinks/FileSink.h): appends pings to a ping file
    (record format in rtsim_sink.h)."""

    def __init__(self, filename, overwrite: bool = False):
        self.filename = str(filename)
        h = C.c_void_p()
        check(lib().rtsim_sink_open(self.filename.encode(), int(overwrite), C.byref(h)))
        self.handle = h

    @classmethod
    def Create(cls, filename, overwrite: bool = False) -> "FileSink":
        return cls(filename, overwrite)

    def set_output(self, sensor: "SensorInstance", stream: "Stream | None" = None):
        check(lib().rtsim_sink_write(self.handle, sensor.handle, stream.handle if stream else None))

    def write_host(self, kind: int, ping: np.ndarray, range_min: float, range_max: float,
                   bearings: np.ndarray | None = None):
        p = np.ascontiguousarray(ping, np.complex64)
        H, W = (p.shape[0], 1) if p.ndim == 1 else p.shape
        b = None if bearings is None else np.ascontiguousarray(bearings, np.float32)
        check(lib().rtsim_sink_write_host(self.handle, kind, H, W, range_min, range_max,
                                          None if b is None else _ptr(b), _ptr(p.view(np.float32))))

    def count(self) -> int:
        n = C.c_uint64()
        check(lib().rtsim_sink_count(self.handle, C.byref(n)))
        return n.value

    def close(self):
        if self.handle:
            lib().rtsim_sink_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ping_file_count(path) -> int:
    n = C.c_uint64()
    check(lib().rtsim_ping_file_count(str(path).encode(), C.byref(n)))
    return n.value


def read_ping(path, index: int = 0) -> dict:
    """Record `index` of a ping file: header fields, bearings, data [range][bearing]."""
    h = PingHeader()
    check(lib().rtsim_ping_file_read(str(path).encode(), index, C.byref(h), None, None))
    nb = h.width if h.kind == SENSOR_2D else 0
    b = np.zeros(max(nb, 1), np.float32)
    d = np.zeros((h.height, h.width), np.complex64)
    check(lib().rtsim_ping_file_read(str(path).encode(), index, C.byref(h), _ptr(b),
                                     _ptr(d.view(np.float32))))
    return {"kind": h.kind, "height": h.height, "width": h.width, "range_min": h.range_min,
            "range_max": h.range_max, "index": h.index, "checksum": h.checksum,
            "bearings": b[:nb], "data": d if h.kind == SENSOR_2D else d[:, 0]}


class Batch:
    """Many sensor instances computed from one hipGraph launch (rtsim_batch.h)."""

    def __init__(self, sensors: list, forms: list[int]):
        self.sensors = list(sensors)
        n = len(self.sensors)
        hs = (C.c_void_p * n)(*[s.handle for s in self.sensors])
        fs = (C.c_int * n)(*forms)
        h = C.c_void_p()
        check(lib().rtsim_batch_create(hs, fs, n, C.byref(h)))
        self.handle = h
        self._hits = {}

    def bind_hits(self, index: int, hits: "DeviceHits | None"):
        if hits is None:
            check(lib().rtsim_batch_bind_hits(self.handle, index, None, 0))
            self._hits.pop(index, None)
        else:
            check(lib().rtsim_batch_bind_hits(self.handle, index, C.byref(hits.struct), hits.n))
            self._hits[index] = hits

    def compute(self, stream: "Stream | None" = None):
        check(lib().rtsim_batch_compute(self.handle, stream.handle if stream else None))

    @property
    def persistent(self) -> bool:
        """Grid-form pings in one persistent launch (rtsim_batch_set_persistent)."""
        on = C.c_int()
        check(lib().rtsim_batch_persistent(self.handle, C.byref(on)))
        return bool(on.value)

    @persistent.setter
    def persistent(self, on: bool):
        check(lib().rtsim_batch_set_persistent(self.handle, 1 if on else 0))

    def check(self):
        """Waits for the last launch; raises if its persistent launch was abandoned."""
        check(lib().rtsim_batch_check(self.handle))

    def autotune(self, iterations: int = 5, stream: "Stream | None" = None) -> float:
        """Per class of MFMA-form sensors, the fastest tile shape and wave
        count for the whole batch (set on the sensors); the batch's replay ms."""
        ms = C.c_float()
        check(lib().rtsim_batch_autotune(self.handle, iterations, stream.handle if stream else None,
                                         C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().rtsim_batch_destroy(self.handle)
        except Exception:
            pass


CONFIG_DIR = PKG_DIR / "configs"


def config_from_yaml(path) -> "SensorConfig":
    """SensorInfoFactory::Make(path) through the C ABI, as a SensorConfig
    (host only: no device work)."""
    from .configs import SensorConfig
    info = C.c_void_p()
    check(lib().rtsim_info_from_yaml(str(path).encode(), C.byref(info)))
    try:
        dp = C.POINTER(SensorDesc)()
        check(lib().rtsim_info_desc(info, C.byref(dp)))
        d = dp.contents
        Tr, Tb = d.psf_range_taps, d.psf_bearing_taps
        psf = np.ctypeslib.as_array(d.psf, (2 * Tr * Tb,)).copy().view(np.complex64)
        cfg = SensorConfig(kind=d.kind, range_min=d.range_min, range_max=d.range_max,
                           range_count=d.range_count, psf=psf.reshape(Tr, Tb),
                           psf_range_span=d.psf_range_span, psf_bearing_span=d.psf_bearing_span,
                           margin=d.margin, bearing_axis=d.bearing_axis,
                           frequency=d.frequency, sound_speed=d.sound_speed,
                           spreading=d.spreading, name=Path(str(path)).stem)
        if d.kind == SENSOR_2D:
            cfg.bearings = np.ctypeslib.as_array(d.bearings, (d.bearing_count,)).copy()
        if d.dir_width and d.dir_height:
            cfg.directivity = np.ctypeslib.as_array(
                d.directivity, (d.dir_width * d.dir_height,)).copy().reshape(d.dir_height,
                                                                              d.dir_width)
            cfg.dir_xscale, cfg.dir_yscale = d.dir_xscale, d.dir_yscale
        cfg.rotation = np.array(list(d.rotation), np.float32).reshape(3, 3)
        if d.waveform:
            w = d.waveform.contents
            beam = np.ctypeslib.as_array(w.beam, (w.beam_taps,)).copy() if w.beam else None
            cfg.waveform = {"frequency": w.frequency, "duration": w.duration,
                            "fixed": bool(w.fixed), "oversampling": int(w.oversampling),
                            "beam": beam}
        return cfg
    finally:
        lib().rtsim_info_destroy(info)


def sensor_from_yaml(path) -> SensorInstance:
    """SensorFactory::Make(path): a sensor instance built from a YAML file."""
    return make_sensor(config_from_yaml(path))


from .configs import SensorConfig  # noqa: E402  (re-export)
