"""rtsim_sensor_compute_noalloc (the entry point for callers that cannot
allocate device memory) against rtsim_sensor_compute, through the C ABI: two
calls in a row on one GPU's scratch, the second with fewer samples, each ping
equal to the allocating entry point's bit for bit (every form sums in a fixed
order or exactly), the caller's ping fully overwritten, the sensor's own
state untouched."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from rtac_simulation_amd import configs

pytestmark = pytest.mark.gpu


def _noalloc(R, s, form, rec_buf, n, ping_buf):
    R.check(R.lib().rtsim_sensor_compute_noalloc(s.handle, form, rec_buf.ptr, n, ping_buf.ptr, None))
    R.check(R.lib().rtsim_stream_sync(None))


def _poisoned_ping(R, H, W):
    return R.DeviceBuffer.from_array(np.full((H, W), np.nan + 1j * np.nan, np.complex64))


def _equal(got, want):
    return np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("kind", ["2d", "1d"])
def test_noalloc_equals_compute_twice(gpu, kind):
    R = gpu
    if kind == "2d":
        cfg = configs._fls(96, 200, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "noalloc")
        scene, forms = configs.fls_scene(), (0, 1, 2)
    else:
        cfg, scene, _, _ = configs.workload("single_beam_38k")
        forms = (0, 1, 2)
    hits = R.generate_hits(scene, 81, 120_000)
    s = R.make_sensor(cfg)
    s.set_sample_count(hits.n)
    s.write_samples(hits.to_device())
    rec = s.samples()
    H, W = s.shape()
    rec_buf = R.DeviceBuffer.from_array(rec)
    for form in forms:
        for n in (hits.n, hits.n // 3):          # second call: a smaller input
            s.set_sample_count(n)
            s.upload_samples(rec[:n])
            s.compute_output(form)
            want = s.output()
            want_sorted = s.samples(sorted=True)
            ping = _poisoned_ping(R, H, W)
            _noalloc(R, s, form, rec_buf, n, ping)
            got = ping.download(np.complex64, H * W).reshape(H, W)
            if kind == "1d":
                got = got[:, 0]
            assert _equal(got, want), (form, n)
            # the sensor's own ping and sorted records are untouched
            assert np.array_equal(s.output().view(np.uint32), want.view(np.uint32)), (form, n)
            assert np.array_equal(s.samples(sorted=True).view(np.uint32), want_sorted.view(np.uint32))


def test_noalloc_matches_oracle_and_rejects_what_it_cannot_do(gpu):
    R = gpu
    cfg = configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "noalloc2")
    hits = R.generate_hits(configs.fls_scene(), 82, 50_000)
    s = R.make_sensor(cfg)
    s.set_sample_count(hits.n)
    s.write_samples(hits.to_device())
    rec = s.samples()
    rec_buf = R.DeviceBuffer.from_array(rec)
    ping = _poisoned_ping(R, cfg.H, cfg.W)
    # the MFMA form is not prepared on this sensor yet: refused, nothing queued
    rc = R.lib().rtsim_sensor_compute_noalloc(s.handle, 2, rec_buf.ptr, hits.n, ping.ptr, None)
    assert rc == R.ERR_INVALID and b"not prepared" in R.lib().rtsim_last_error()
    # beyond the scratch: refused
    rc = R.lib().rtsim_sensor_compute_noalloc(s.handle, 0, rec_buf.ptr, (1 << 20) + 1, ping.ptr, None)
    assert rc == R.ERR_INVALID
    rc = R.lib().rtsim_sensor_compute_noalloc(s.handle, 0, None, 10, ping.ptr, None)
    assert rc == R.ERR_INVALID
    _noalloc(R, s, 0, rec_buf, hits.n, ping)
    got = ping.download(np.complex64, cfg.H * cfg.W).reshape(cfg.H, cfg.W)
    ref, _ = O.reference_ping(cfg, rec)
    assert O.rel_l2(got, ref) <= 1e-5
    # zero samples: a zero ping
    _noalloc(R, s, 0, rec_buf, 0, ping)
    assert not np.abs(ping.download(np.complex64, cfg.H * cfg.W)).any()


def test_noalloc_plain_full_size_equals_compute(gpu):
    """The 512 x 1024 ping from 1M samples in the plain form: the allocating
    entry stages each tile's samples from its strip's sample list (32 strips of
    16 bearings, a sample reaches at most 4), the no-allocation entry scans the
    windows; every lane adds the same terms in the same order, so the pings
    are equal bit for bit."""
    R = gpu
    cfg, scene, S, seed = configs.workload("fls_512x1024")
    hits = R.generate_hits(scene, seed, S)
    s = R.make_sensor(cfg)
    s.set_sample_count(hits.n)
    s.write_samples(hits.to_device())
    s.compute_output(0)
    want = s.output()
    rec_buf = R.DeviceBuffer.from_array(s.samples())
    H, W = s.shape()
    ping = _poisoned_ping(R, H, W)
    _noalloc(R, s, 0, rec_buf, hits.n, ping)
    got = ping.download(np.complex64, H * W).reshape(H, W)
    assert np.abs(want).max() > 0
    assert _equal(got, want)
