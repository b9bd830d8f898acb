"""Sensor and scene configurations of the sample-to-ping workloads.

``SensorConfig`` carries what ``rtsim_sensor_desc`` needs (ranges, bearing
table, explicit PSF taps and spans, Binner margin, receiver directivity
table, acoustics).  ``workload(name)`` returns (sensor, scene, samples, seed)
for the named workloads:

  single_beam_38k   1D, 20 000 samples, 2048 bins over [0, 50] m, 33-tap
                    Hann baseband PSF, 38 kHz piston directivity
  fls_512x1024_sep  2D, 512 sin-uniform bearings (+-65 deg) x 1024 bins over
                    [0.1, 30] m, 1 000 000 seabed samples, 15x9 separable PSF
  fls_512x1024      same ping, non-separable 31x17 complex PSF (Hann LFM x
                    sinc^2 + seeded complex Gaussian perturbation at -30 dB)
  fls_1024x4096     1024 bearings x 4096 bins over [0.1, 120] m at 750 kHz,
                    16 000 000 samples (3 planes + 200 spheres, 60 dB range),
                    63x33 PSF
``batch_uneven()`` returns the 64-ping mixed list (16 x 256x512, 32 x 512x1024,
8 x 1024x2048, 8 x 1D 2048, sample counts log-uniform in [1e5, 4e6]).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

SOUND_SPEED = 1500.0
_AXIS_ANGLE, _AXIS_SINE = 0, 1
_SENSOR_1D, _SENSOR_2D = 1, 2
_SCENE_CONE, _SCENE_SEABED = 1, 2


@dataclass
class SensorConfig:
    kind: int
    range_min: float
    range_max: float
    range_count: int
    psf: np.ndarray                    # complex64 [Tr, Tb]
    psf_range_span: float
    psf_bearing_span: float = 1.0
    margin: float = -1.0
    bearings: np.ndarray | None = None  # float32 radians, ascending
    bearing_axis: int = _AXIS_ANGLE
    directivity: np.ndarray | None = None  # float32 [H, W]
    dir_xscale: float = float(np.float32(1.0 / np.pi))
    dir_yscale: float = float(np.float32(1.0 / np.pi))
    frequency: float = 1.2e6
    sound_speed: float = SOUND_SPEED
    spreading: int = 0
    rotation: np.ndarray = field(default_factory=lambda: np.eye(3, dtype=np.float32))
    name: str = ""
    # sine waveform recipe (rtsim_waveform_sine): {frequency, duration, fixed,
    # oversampling, beam}; None: the taps in `psf` are fixed
    waveform: dict | None = None

    # keep ctypes views alive while a desc is in use
    _keep: list = field(default_factory=list, repr=False)

    @property
    def Tr(self) -> int:
        return int(self.psf.shape[0])

    @property
    def Tb(self) -> int:
        return int(self.psf.shape[1])

    @property
    def H(self) -> int:
        return int(self.range_count)

    @property
    def W(self) -> int:
        return 1 if self.kind == _SENSOR_1D else int(len(self.bearings))

    @property
    def range_step(self) -> float:
        return (np.float64(np.float32(self.range_max)) - np.float64(np.float32(self.range_min))) \
            / (self.range_count - 1)

    def desc(self):
        from . import SensorDesc
        d = SensorDesc()
        d.kind = self.kind
        d.range_min, d.range_max, d.range_count = self.range_min, self.range_max, self.range_count
        psf = np.ascontiguousarray(self.psf.astype(np.complex64)).view(np.float32).ravel()
        keep = [psf]
        d.psf = psf.ctypes.data_as(C.POINTER(C.c_float))
        d.psf_range_taps, d.psf_bearing_taps = self.Tr, self.Tb
        d.psf_range_span, d.psf_bearing_span = self.psf_range_span, self.psf_bearing_span
        d.margin = self.margin
        if self.kind == _SENSOR_2D:
            b = np.ascontiguousarray(self.bearings, dtype=np.float32)
            keep.append(b)
            d.bearings = b.ctypes.data_as(C.POINTER(C.c_float))
            d.bearing_count = len(b)
            d.bearing_axis = self.bearing_axis
        if self.directivity is not None:
            t = np.ascontiguousarray(self.directivity, dtype=np.float32)
            keep.append(t)
            d.directivity = t.ctypes.data_as(C.POINTER(C.c_float))
            d.dir_height, d.dir_width = t.shape
            d.dir_xscale, d.dir_yscale = self.dir_xscale, self.dir_yscale
        d.frequency, d.sound_speed, d.spreading = self.frequency, self.sound_speed, self.spreading
        for i, v in enumerate(np.asarray(self.rotation, np.float32).ravel()):
            d.rotation[i] = float(v)
        if self.waveform is not None:
            from . import WaveformSine
            w = WaveformSine()
            wf = self.waveform
            w.frequency, w.duration = wf["frequency"], wf["duration"]
            w.fixed, w.oversampling = int(bool(wf["fixed"])), int(wf["oversampling"])
            if wf.get("beam") is not None:
                bm = np.ascontiguousarray(wf["beam"], np.float32)
                keep.append(bm)
                w.beam = bm.ctypes.data_as(C.POINTER(C.c_float))
                w.beam_taps = bm.size
            else:
                w.beam_taps = 1
            keep.append(w)
            d.waveform = C.pointer(w)
        self._keep = keep
        return d

    # -- oracle-side descriptions ---------------------------------------------
    def bearing_coords(self) -> np.ndarray:
        b = np.asarray(self.bearings, np.float32).astype(np.float64)
        return np.sin(b) if self.bearing_axis == _AXIS_SINE else b


# ----------------------------------------------------------------------------
# PSF / directivity / bearing builders
# ----------------------------------------------------------------------------
def hann(n: int) -> np.ndarray:
    j = np.arange(n)
    return 0.5 - 0.5 * np.cos(2.0 * np.pi * (j + 1) / (n + 1))


def baseband_pulse(n: int, carrier: float = 0.25) -> np.ndarray:
    """Hann-windowed complex baseband pulse, residual carrier in cycles/2 per tap."""
    j = np.arange(n) - (n - 1) / 2
    return hann(n) * np.exp(1j * np.pi * carrier * j)


def lfm_pulse(n: int, kappa: float) -> np.ndarray:
    """Hann-windowed linear FM (chirp) baseband pulse."""
    j = np.arange(n) - (n - 1) / 2
    return hann(n) * np.exp(1j * np.pi * kappa * j * j)


def sinc2_beam(n: int, width: float) -> np.ndarray:
    k = np.arange(n) - (n - 1) / 2
    return np.sinc(k / width) ** 2


def perturbed(psf: np.ndarray, db: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    noise = (rng.standard_normal(psf.shape) + 1j * rng.standard_normal(psf.shape)) / np.sqrt(2)
    return psf + 10.0 ** (db / 20.0) * np.abs(psf).max() * noise


def sine_bearings(n: int, half_deg: float) -> tuple[np.ndarray, float]:
    """n bearings uniform in sin(bearing) over +-half_deg; returns (table, sine pitch)."""
    s = np.sin(np.deg2rad(half_deg))
    u = np.linspace(-s, s, n)
    return np.arcsin(u).astype(np.float32), 2.0 * s / (n - 1)


def fan_directivity(n: int = 256, bearing_width_deg: float = 130.0,
                    elevation_sigma_deg: float = 6.0) -> np.ndarray:
    """Receiver directivity table: sinc^2 in bearing x Gaussian in elevation,
    single-sided (zero beyond 90 deg).  Texel (w, h) is evaluated at its
    centre angle pi (w + 1/2) / n, where the linear-filter lookup reads it."""
    ang = np.pi * (np.arange(n) + 0.5) / n
    b = np.sinc(ang / np.deg2rad(bearing_width_deg)) ** 2
    e = np.exp(-0.5 * (ang / np.deg2rad(elevation_sigma_deg)) ** 2)
    b[ang > np.pi / 2] = 0.0
    e[ang > np.pi / 2] = 0.0
    return np.outer(e, b).astype(np.float32)


def overlap_margin(ov: int, step: float) -> float:
    """Binner margin giving exactly `ov` overlap bins after the (int) truncation."""
    return float(np.float32(ov * step * (1.0 + 2.0 ** -10)))


def required_overlap(Tr: int) -> int:
    """Window half-width (bins) that holds every sample of the PSF range support."""
    return (Tr + 3) // 2


# ----------------------------------------------------------------------------
# scenes
# ----------------------------------------------------------------------------
def cone_scene(rmin: float, rmax: float, off_axis_deg: float):
    from . import SceneDesc
    s = SceneDesc()
    s.kind = _SCENE_CONE
    s.range_min, s.range_max = rmin, rmax
    s.off_axis_max = np.deg2rad(off_axis_deg)
    s.rayleigh_sigma = 1.0
    return s


def seabed_scene(planes, bearing_deg=65.0, elevation_deg=10.0, noise=0.05, spheres=0,
                 sphere_radius=(0.2, 2.0), sphere_range=(5.0, 30.0), sphere_fraction=0.0,
                 dynamic_range_db=0.0):
    from . import SceneDesc
    s = SceneDesc()
    s.kind = _SCENE_SEABED
    s.bearing_max = np.deg2rad(bearing_deg)
    s.elevation_max = np.deg2rad(elevation_deg)
    s.plane_count = len(planes)
    for i, (n, h) in enumerate(planes):
        n = np.asarray(n, np.float64)
        n = n / np.linalg.norm(n)
        for k in range(3):
            s.planes[i][k] = float(n[k])
        s.planes[i][3] = float(h)
    s.height_noise = noise
    s.sphere_count = spheres
    s.sphere_radius_min, s.sphere_radius_max = sphere_radius
    s.sphere_range_min, s.sphere_range_max = sphere_range
    s.sphere_fraction = sphere_fraction
    s.dynamic_range_db = dynamic_range_db
    s.rayleigh_sigma = 1.0
    s.incidence_power = 2
    return s


def tilted_plane(tilt_deg: float, depth: float, roll_deg: float = 0.0):
    """Seabed plane `depth` m below the sensor rising ahead at `tilt_deg`."""
    t, r = np.deg2rad(tilt_deg), np.deg2rad(roll_deg)
    return ((-np.sin(t), np.sin(r), np.cos(t) * np.cos(r)), depth)


# ----------------------------------------------------------------------------
# workloads
# ----------------------------------------------------------------------------
def single_beam_38k(samples: int = 20_000):
    f = 38e3
    lam = SOUND_SPEED / f
    ka = 1.6163 / np.sin(np.deg2rad(3.5))        # 7 deg -3 dB beamwidth
    diameter = ka * lam / np.pi
    from . import directivity_disk
    N, lo, hi = 2048, 0.0, 50.0
    cfg = SensorConfig(kind=_SENSOR_1D, range_min=lo, range_max=hi, range_count=N,
                       psf=baseband_pulse(33)[:, None].astype(np.complex64),
                       psf_range_span=0.0, frequency=f, name="single_beam_38k")
    step = cfg.range_step
    cfg.psf_range_span = float(np.float32(33 * step))
    cfg.margin = overlap_margin(required_overlap(33), step)
    cfg.directivity = directivity_disk(float(diameter), float(lam), True, 8)
    return cfg, cone_scene(0.5, 50.0, 10.0), samples, 1


def _fls(W, H, lo, hi, f, psf, name):
    bearings, pitch = sine_bearings(W, 65.0)
    cfg = SensorConfig(kind=_SENSOR_2D, range_min=lo, range_max=hi, range_count=H,
                       psf=psf.astype(np.complex64), psf_range_span=0.0,
                       bearings=bearings, bearing_axis=_AXIS_SINE, frequency=f, name=name)
    step = cfg.range_step
    cfg.psf_range_span = float(np.float32(cfg.Tr * step))
    cfg.psf_bearing_span = float(np.float32(cfg.Tb * pitch))
    cfg.margin = overlap_margin(required_overlap(cfg.Tr), step)
    cfg.directivity = fan_directivity()
    return cfg


def psf_31x17(seed: int = 3) -> np.ndarray:
    return perturbed(np.outer(lfm_pulse(31, 0.02), sinc2_beam(17, 5.0)), -30.0, seed)


def psf_63x33(seed: int = 5) -> np.ndarray:
    return perturbed(np.outer(lfm_pulse(63, 0.01), sinc2_beam(33, 10.0)), -30.0, seed)


def fls_scene():
    return seabed_scene([tilted_plane(20.0, 5.0)], noise=0.05)


def fls_512x1024_sep(samples: int = 1_000_000):
    psf = np.outer(baseband_pulse(15), sinc2_beam(9, 3.0))
    return _fls(512, 1024, 0.1, 30.0, 1.2e6, psf, "fls_512x1024_sep"), fls_scene(), samples, 2


def fls_512x1024(samples: int = 1_000_000):
    return _fls(512, 1024, 0.1, 30.0, 1.2e6, psf_31x17(), "fls_512x1024"), fls_scene(), samples, 2


def fls_1024x4096(samples: int = 16_000_000):
    cfg = _fls(1024, 4096, 0.1, 120.0, 750e3, psf_63x33(), "fls_1024x4096")
    scene = seabed_scene([tilted_plane(20.0, 5.0), tilted_plane(3.0, 15.0),
                          tilted_plane(8.0, 30.0, 4.0)], noise=0.05, spheres=200,
                         sphere_radius=(0.2, 2.0), sphere_range=(10.0, 110.0),
                         sphere_fraction=0.25, dynamic_range_db=60.0)
    return cfg, scene, samples, 5


WORKLOADS = {
    "single_beam_38k": single_beam_38k,
    "fls_512x1024_sep": fls_512x1024_sep,
    "fls_512x1024": fls_512x1024,
    "fls_1024x4096": fls_1024x4096,
}


def workload(name: str, samples: int | None = None):
    fn = WORKLOADS[name]
    return fn() if samples is None else fn(samples)


def batch_uneven(seed: int = 4):
    """64 mixed pings: (cfg, scene, samples, seed) each."""
    rng = np.random.default_rng(seed)
    shapes = [(256, 512)] * 16 + [(512, 1024)] * 32 + [(1024, 2048)] * 8 + [(1, 2048)] * 8
    counts = np.exp(rng.uniform(np.log(1e5), np.log(4e6), len(shapes))).astype(np.int64)
    out = []
    for i, ((W, H), S) in enumerate(zip(shapes, counts)):
        if W == 1:
            cfg = SensorConfig(kind=_SENSOR_1D, range_min=0.0, range_max=50.0, range_count=H,
                               psf=lfm_pulse(31, 0.02)[:, None].astype(np.complex64),
                               psf_range_span=0.0, frequency=38e3, name=f"batch_{i}_1d")
            step = cfg.range_step
            cfg.psf_range_span = float(np.float32(31 * step))
            cfg.margin = overlap_margin(required_overlap(31), step)
            out.append((cfg, cone_scene(0.5, 50.0, 10.0), int(S), 100 + i))
        else:
            cfg = _fls(W, H, 0.1, 30.0, 1.2e6, psf_31x17(), f"batch_{i}_{W}x{H}")
            out.append((cfg, fls_scene(), int(S), 100 + i))
    return out
