"""GPU parity against the FP64 oracle for the reference's own sensors and for
the raygen branches the synthetic workloads leave off:

  * pings of the reference's example receivers read from their YAML
    (SensorInfoFactory.cpp:7-235): the Oculus M1200d (256 CSV bearings on the
    angle axis, adaptive sine waveform x sinc beam, binOverlap 1) and the
    EA400 38 kHz echo sounder (1D, fixed 128 us waveform, binOverlap 4);
  * SensorInstance::set_ranges (SensorInstance.cu:23-34) on those sensors:
    the adaptive waveform's duration, PSF table, span and margin follow the
    new ranges (checked with taps restated in numpy, not read back);
  * the 1/r^2 spreading of RayCaster.cu:69-71 and a non-identity receiver
    pose (Receiver.h:56: local = R^T dir), in the records and in the pings.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from rtac_simulation_amd import configs
from sine_waveform import after_set_ranges

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "reference_configs"
PLAIN_TOL = 1e-5
MFMA_TOL = 1e-3


def _rotation(yaw, pitch, roll):
    cy, sy = np.cos(yaw), np.sin(yaw)
    cp, sp = np.cos(pitch), np.sin(pitch)
    cr, sr = np.cos(roll), np.sin(roll)
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    return (Rz @ Ry @ Rx).astype(np.float32)


def _check_records(rec, ref):
    assert np.array_equal(rec[:, 2], ref[:, 2].astype(np.float32)), "travel must pass through"
    v = rec[:, 0] + 1j * rec[:, 1]
    vr = ref[:, 0] + 1j * ref[:, 1]
    assert O.rel_l2(v, vr) < 2e-6
    assert np.max(np.abs(rec[:, 3] - ref[:, 3])) < 2e-6


def _check_plain(s, cfg, rows=None):
    b = O.binning(cfg, s.samples())
    d = s.binner_debug()
    assert d["overlap"] == b["overlap"]
    for key in ("order", "keys", "segments", "windows"):
        assert np.array_equal(d[key], b[key]), key
    if cfg.kind == 2 and rows is not None:
        ref = O.conv_rows(cfg, b["sorted"], b["windows"], rows)
        got = s.output()[rows]
    else:
        ref = O.conv(cfg, b["sorted"], b["windows"])
        got = s.output()
    assert np.abs(ref).max() > 0
    err = O.rel_l2(got, ref)
    assert err <= PLAIN_TOL, err
    return b


# ---------------------------------------------------------------------------
# spreading and pose
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("via", ["desc", "set_pose"])
def test_spreading_and_pose_vs_oracle(gpu, via):
    R = gpu
    cfg = configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "pose")
    cfg.spreading = R.SPREAD_INV_R2
    rot = _rotation(np.deg2rad(12.0), np.deg2rad(-4.0), np.deg2rad(7.0))
    hits = R.generate_hits(configs.fls_scene(), 31, 60_000)
    if via == "desc":
        cfg.rotation = rot
        s = R.make_sensor(cfg)
    else:
        s = R.make_sensor(cfg)
        cfg.rotation = rot
        s.set_pose(rot)
    s.set_sample_count(hits.n)
    dh = hits.to_device()
    s.write_samples(dh)
    ref = O.stream(cfg, hits)
    _check_records(s.samples(), ref)
    # the pose really moves the bearings (the oracle's R^T convention is
    # pinned on the CPU, test_stream_pose_rotation)
    still = O.stream(configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "t"), hits)
    ok = hits.range >= 1e-4
    assert np.abs(still[ok, 3] - ref[ok, 3]).max() > 1e-2
    s.compute_output(0)
    b = _check_plain(s, cfg)
    for form in (1, 2):
        s.write_samples(dh)
        s.compute_output(form)
        full = O.conv(cfg, b["sorted"], b["windows"])
        err = O.rel_l2(s.output(), full)
        assert err <= MFMA_TOL, (form, err)


# ---------------------------------------------------------------------------
# the reference's example receivers
# ---------------------------------------------------------------------------
def _oculus_scene():
    # ranges uniform over the sensor's span and bearings over +-60 deg: every
    # range row holds samples, so every window of binOverlap 1 is non-empty
    return configs.cone_scene(0.2, 9.5, 60.0)


def _oculus(R):
    cfg = R.config_from_yaml(GOLDEN / "oculus_M1200d_1_receiver.yaml")
    assert cfg.waveform is not None and not cfg.waveform["fixed"]
    return cfg


def test_oculus_yaml_ping_vs_oracle(gpu):
    R = gpu
    cfg = _oculus(R)
    assert cfg.bearing_axis == R.AXIS_ANGLE and cfg.W == 256
    hits = R.generate_hits(_oculus_scene(), 41, 400_000)
    s = R.sensor_from_yaml(GOLDEN / "oculus_M1200d_1_receiver.yaml")
    s.set_sample_count(hits.n)
    s.write_samples(hits.to_device())
    _check_records(s.samples(), O.stream(cfg, hits))
    s.compute_output(0)
    b = _check_plain(s, cfg, rows=np.arange(0, cfg.H, 3))
    assert b["overlap"] == 1
    # non-uniform bearings and a PSF pitch finer than the grid: the MFMA form
    # cannot restate this sensor exactly and refuses it
    with pytest.raises(R.RtsimError) as e:
        s.compute_output(2)
    assert e.value.code == R.ERR_UNSUPPORTED


def test_oculus_set_ranges_rebuilds_the_waveform(gpu):
    R = gpu
    cfg = _oculus(R)
    s = R.make_sensor(cfg)
    hits = R.generate_hits(_oculus_scene(), 42, 300_000)
    s.set_sample_count(hits.n)
    dh = hits.to_device()
    for lo, hi, n in [(0.1, 4.0, 300), (1.0, 9.0, 720)]:
        s.set_ranges(lo, hi, n)
        want = after_set_ranges(cfg, lo, hi, n)
        assert want.Tr != cfg.Tr               # the table really changes size
        s.write_samples(dh)
        s.compute_output(0)
        b = _check_plain(s, want, rows=np.arange(1, n, 4))
        assert b["overlap"] == O.overlap(want.psf_range_span, lo, hi, n)


@pytest.mark.parametrize("fname,new_count", [("EA400_38KHz_receiver.yaml", 800),
                                             ("EA400_200KHz_receiver.yaml", 2900)])
def test_echo_sounder_yaml_ping_vs_oracle(gpu, fname, new_count):
    R = gpu
    cfg = R.config_from_yaml(GOLDEN / fname)
    assert cfg.kind == 1 and cfg.waveform["fixed"]
    hits = R.generate_hits(configs.cone_scene(0.5, 50.0, 10.0), 43, 50_000)
    s = R.sensor_from_yaml(GOLDEN / fname)
    s.set_sample_count(hits.n)
    s.write_samples(hits.to_device())
    _check_records(s.samples(), O.stream(cfg, hits))
    s.compute_output(0)
    _check_plain(s, cfg)
    # a fixed waveform keeps its duration: same taps, margin c * tau
    # (new_count keeps binOverlap = (int)(c tau / resolution) >= 1)
    s.set_ranges(1.0, 30.0, new_count)
    want = after_set_ranges(cfg, 1.0, 30.0, new_count)
    assert O.overlap(want.psf_range_span, 1.0, 30.0, new_count) >= 1
    assert want.psf.shape == cfg.psf.shape
    assert np.allclose(want.psf, cfg.psf, rtol=0, atol=2e-6)
    s.write_samples(hits.to_device())
    s.compute_output(0)
    _check_plain(s, want)
    # ranges too coarse for the fixed pulse: binOverlap 0 empties every
    # window (Binner.cu:96-97), as in the reference: a zero ping
    s.set_ranges(1.0, 30.0, 100)
    zero = after_set_ranges(cfg, 1.0, 30.0, 100)
    s.write_samples(hits.to_device())
    s.compute_output(0)
    b = O.binning(zero, s.samples())
    assert b["overlap"] == 0 and s.binner_debug()["overlap"] == 0
    assert not b["windows"].any() and not np.abs(s.output()).any()
