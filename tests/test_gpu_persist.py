"""Persistent launch of the grid-form pings (rtsim_batch_set_persistent): every
stage of every MFMA-form ping of a batch in one kernel, tasks taken from a
ticket in level-major order.  Every replay must reproduce each sensor's own
compute bit for bit (same tile shape, 8-wave correlation tiles), with eager
work that overwrites the outputs between replays, and stay within the MFMA
form's bound of the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
from rtac_simulation_amd import configs

pytestmark = pytest.mark.gpu

MFMA_TOL = 1e-3


def _uneven_specs():
    """(config, scene, samples, form, tile shape): 2D pings of uneven sizes
    in both MFMA dtypes and every tile shape, a 1D ping, an empty sensor and
    a plain ping sharing the batch."""
    specs = []
    for k, (W, H, n, form, shape) in enumerate([
            (64, 128, 30_000, 1, 0), (40, 96, 12_000, 2, 1), (96, 200, 50_000, 2, 2),
            (128, 256, 70_000, 1, 3), (32, 64, 0, 2, 1), (80, 160, 25_000, 2, 1),
            (48, 96, 8_000, 0, None), (56, 112, 15_000, 1, 2)]):
        cfg = configs._fls(W, H, 0.1, 30.0, 1.2e6, configs.psf_31x17(), f"p{k}")
        specs.append((cfg, configs.fls_scene(), n, form, shape))
    c1, sc1, _, _ = configs.workload("single_beam_38k")
    specs.append((c1, sc1, 9_000, 2, None))
    return specs


def _make(R, specs, seed0):
    sensors, hits, forms = [], [], []
    for k, (cfg, scene, n, form, shape) in enumerate(specs):
        s = R.make_sensor(cfg)
        s.set_sample_count(n)
        if shape is not None and form != 0:
            s.set_tile_shape(shape)
        sensors.append(s)
        hits.append(R.generate_hits(scene, seed0 + k, max(n, 1)).to_device().prefix(n))
        forms.append(form)
    return sensors, hits, forms


def test_persistent_batch_equals_eager_and_oracle(gpu):
    R = gpu
    specs = _uneven_specs()
    sensors, hits, forms = _make(R, specs, 300)
    b = R.Batch(sensors, forms)
    b.persistent = True
    assert b.persistent
    for s, f in zip(sensors, forms):
        if f != 0:
            assert s.tile_waves() == 8
    # ping 2 keeps its records from an eager write: its bucket counts come
    # from the records (no streaming level in the persistent launch)
    unbound = 2
    for k, h in enumerate(hits):
        if h.n and k != unbound:
            b.bind_hits(k, h)
    sensors[unbound].write_samples(hits[unbound])
    eager = []
    for k, (s, h, f) in enumerate(zip(sensors, hits, forms)):
        if h.n:
            s.write_samples(h)
        s.compute_output(f)
        eager.append((s.output().copy(), s.samples().copy()))
    stream = R.Stream()
    for rep in range(3):
        # overwrite every output (and the bound pings' records) eagerly
        for k, (s, h, f) in enumerate(zip(sensors, hits, forms)):
            if k == unbound:
                s.compute_output(0)
            elif h.n:
                s.upload_samples(np.zeros((h.n, 4), np.float32))
                s.compute_output(f)
        b.compute(stream)
        stream.sync()
        b.check()
        for k, ((ping, rec), s) in enumerate(zip(eager, sensors)):
            assert np.array_equal(s.samples().view(np.uint32), rec.view(np.uint32)), (rep, k)
            assert np.array_equal(s.output().view(np.uint32), ping.view(np.uint32)), (rep, k)
    for k, ((cfg, _, n, form, _), (ping, rec)) in enumerate(zip(specs, eager)):
        if not n:
            assert not np.any(ping), k
            continue
        ref, _ = O.reference_ping(cfg, rec)
        tol = 1e-5 if form == 0 else MFMA_TOL
        assert O.rel_l2(ping, ref) <= tol, (k, O.rel_l2(ping, ref))


def test_persistent_headline_ping_vs_oracle(gpu):
    """The 512 x 1024, 1M-sample ping with the 31 x 17 PSF as a persistent
    batch of one: equal to the sensor's own compute, and every 32nd row within
    the form's bound of the oracle, for both dtypes."""
    R = gpu
    cfg, scene, S, seed = configs.workload("fls_512x1024")
    dh = R.generate_hits(scene, seed, S).to_device()
    for form in (R.FORM_MFMA_BF16, R.FORM_MFMA_FP16):
        s = R.make_sensor(cfg)
        s.set_sample_count(S)
        b = R.Batch([s], [form])
        b.persistent = True
        b.bind_hits(0, dh)
        stream = R.Stream()
        for _ in range(2):
            b.compute(stream)
        stream.sync()
        b.check()
        got = s.output().copy()
        rec = s.samples()
        s.write_samples(dh)
        s.compute_output(form)
        assert np.array_equal(got.view(np.uint32), s.output().view(np.uint32)), form
        bo = O.binning(cfg, rec)
        rows = np.arange(7, cfg.H, 32)
        ref = O.conv_rows(cfg, bo["sorted"], bo["windows"], rows)
        err = O.rel_l2(got[rows], ref)
        print(f"persistent headline form {form}: rel_l2 {err:.3e}", flush=True)
        assert err <= MFMA_TOL, err


def test_persistent_pose_update(gpu):
    """set_pose between persistent replays rewrites the ping's streaming
    arguments in place; the next replay equals an eager compute with that pose."""
    R = gpu
    c0 = configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "q0")
    hits = R.generate_hits(configs.fls_scene(), 77, 30_000)
    dh = hits.to_device()
    s = R.make_sensor(c0)
    s.set_sample_count(hits.n)
    b = R.Batch([s], [2])
    b.persistent = True
    b.bind_hits(0, dh)
    stream = R.Stream()
    c, sn = np.cos(0.3), np.sin(0.3)
    rot = np.array([[c, -sn, 0], [sn, c, 0], [0, 0, 1]], np.float32)
    for rep, pose in enumerate((np.eye(3, dtype=np.float32), rot, np.eye(3, dtype=np.float32))):
        s.set_pose(pose)
        b.bind_hits(0, dh)
        b.compute(stream)
        stream.sync()
        b.check()
        got = s.output().copy()
        cr = configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "qr")
        cr.rotation = pose
        e = R.make_sensor(cr)
        e.set_sample_count(hits.n)
        e.set_tile_shape(s.tile_shape()[0])
        e.set_tile_waves(8)
        e.write_samples(dh)
        e.compute_output(2)
        assert np.array_equal(got.view(np.uint32), e.output().view(np.uint32)), rep
        if rep == 1:
            rec = O.stream(cr, hits).astype(np.float32)
            ref, _ = O.reference_ping(cr, rec)
            assert O.rel_l2(got, ref) <= MFMA_TOL


def test_persistent_and_graph_modes_switch(gpu):
    """Turning persistent mode off and on re-captures; both modes give the
    same pings (same tiles)."""
    R = gpu
    specs = _uneven_specs()[:4]
    sensors, hits, forms = _make(R, specs, 400)
    b = R.Batch(sensors, forms)
    b.persistent = True
    for k, h in enumerate(hits):
        if h.n:
            b.bind_hits(k, h)
    stream = R.Stream()
    b.compute(stream)
    stream.sync()
    b.check()
    first = [s.output().copy() for s in sensors]
    b.persistent = False
    b.compute(stream)
    stream.sync()
    for k, (p, s) in enumerate(zip(first, sensors)):
        assert np.array_equal(p.view(np.uint32), s.output().view(np.uint32)), k
    b.persistent = True
    b.compute(stream)
    stream.sync()
    b.check()
    for k, (p, s) in enumerate(zip(first, sensors)):
        assert np.array_equal(p.view(np.uint32), s.output().view(np.uint32)), k
