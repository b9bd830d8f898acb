"""Batched pings from one hipGraph launch: mixed 1D / 2D sensors of uneven
sizes, plain and MFMA forms.  The graph is replayed several times with eager
work and synchronisations between replays; every replay must reproduce the
eager per-sensor results (binning bit-exact, pings to FP32 summation order)."""
import numpy as np
import pytest

import oracle as O
from rtac_simulation_amd import configs

pytestmark = pytest.mark.gpu


def _small_batch(R):
    specs = [
        (configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "b0"), 30_000, 1),
        (configs._fls(40, 96, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "b1"), 12_000, 2),
        (configs._fls(96, 200, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "b2"), 50_000, 0),
    ]
    c1, sc1, _, _ = configs.workload("single_beam_38k")
    out = []
    for cfg, n, form in specs:
        out.append((cfg, configs.fls_scene(), n, form))
    out.append((c1, sc1, 9_000, 0))
    return out


def test_batch_graph_matches_eager(gpu):
    R = gpu
    items = _small_batch(R)
    sensors, hits, forms = [], [], []
    for k, (cfg, scene, n, form) in enumerate(items):
        s = R.make_sensor(cfg)
        s.set_sample_count(n)
        h = R.generate_hits(scene, 40 + k, n).to_device()
        sensors.append(s)
        hits.append(h)
        forms.append(form)
    # eager reference results
    eager = []
    for s, h, f in zip(sensors, hits, forms):
        s.write_samples(h)
        s.compute_output(f)
        eager.append((s.output(), s.binner_debug()["order"]))
    b = R.Batch(sensors, forms)
    for k, h in enumerate(hits):
        b.bind_hits(k, h)
    stream = R.Stream()
    for rep in range(3):
        # scramble outputs and samples eagerly between replays
        for s, h, f in zip(sensors, hits, forms):
            s.upload_samples(np.zeros((s.sample_count(), 4), np.float32))
        b.compute(stream)
        stream.sync()
        for (ping, order), s in zip(eager, sensors):
            assert np.array_equal(s.binner_debug()["order"], order)
            assert np.array_equal(s.output().view(np.uint32), ping.view(np.uint32)), rep


def test_batch_autotune_keeps_batched_equal_to_eager(gpu):
    """The batch autotuner sets a tile shape per sensor class; the batched
    pings it leaves must still equal each sensor's own compute with that shape
    (binning bit-exact, pings to FP32 summation order), and the MFMA pings stay
    within the form's bound of the oracle."""
    R = gpu
    items = _small_batch(R)
    sensors, hits, forms = [], [], []
    for k, (cfg, scene, n, form) in enumerate(items):
        s = R.make_sensor(cfg)
        s.set_sample_count(n)
        sensors.append(s)
        hits.append(R.generate_hits(scene, 60 + k, n).to_device())
        forms.append(form)
    b = R.Batch(sensors, forms)
    for k, h in enumerate(hits):
        b.bind_hits(k, h)
    stream = R.Stream()
    ms = b.autotune(2, stream)
    assert ms > 0.0
    stream.sync()
    batched = [(s.output(), s.binner_debug()["order"]) for s in sensors]
    for (ping, order), s, h, f, (cfg, _, n, _) in zip(batched, sensors, hits, forms, items):
        s.write_samples(h)
        s.compute_output(f)
        assert np.array_equal(s.binner_debug()["order"], order)
        assert np.array_equal(s.output().view(np.uint32), ping.view(np.uint32))
        if f != R.FORM_PLAIN_FP32:
            ref, _ = O.reference_ping(cfg, s.samples())
            assert O.rel_l2(ping, ref) < 1e-3


def test_batch_recaptures_after_resize(gpu):
    R = gpu
    cfg = configs._fls(48, 96, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "r")
    s = R.make_sensor(cfg)
    s.set_sample_count(10_000)
    s.write_samples(R.generate_hits(configs.fls_scene(), 50, 10_000).to_device())
    b = R.Batch([s], [1])
    b.compute()
    s.set_sample_count(20_000)
    h = R.generate_hits(configs.fls_scene(), 51, 20_000).to_device()
    s.write_samples(h)
    b.compute()
    R.Stream().sync()
    ref, _ = O.reference_ping(cfg, s.samples())
    assert O.rel_l2(s.output(), ref) <= 1e-3


def test_batch_rejects_unsupported_form(gpu):
    R = gpu
    c1, sc1, _, _ = configs.workload("single_beam_38k")
    c1.psf_range_span *= 2.0        # PSF pitch of two range steps: no grid form
    s = R.make_sensor(c1)
    s.set_sample_count(100)
    with pytest.raises(R.RtsimError) as e:
        R.Batch([s], [1])
    assert e.value.code == R.ERR_UNSUPPORTED


def test_batched_launches_many_uneven_pings(gpu):
    """Twelve pings of uneven sizes, both MFMA forms on different tile shapes
    (so the correlation runs as several batched launches), plain 2D and 1D,
    and an empty sensor: every stage runs once for the whole batch and must
    reproduce the per-sensor computes bit for bit."""
    R = gpu
    specs = []
    for k, (W, H, n, form, shape) in enumerate([
            (64, 128, 30_000, 1, 0), (40, 96, 12_000, 2, 1), (96, 200, 50_000, 2, 2),
            (48, 96, 8_000, 0, None), (128, 256, 70_000, 1, 3), (32, 64, 0, 2, 1),
            (80, 160, 25_000, 2, 1), (56, 112, 15_000, 1, 1), (72, 150, 40_000, 0, None),
            (64, 128, 5_000, 2, 0)]):
        cfg = configs._fls(W, H, 0.1, 30.0, 1.2e6, configs.psf_31x17(), f"u{k}")
        specs.append((cfg, configs.fls_scene(), n, form, shape))
    c1, sc1, _, _ = configs.workload("single_beam_38k")
    specs.append((c1, sc1, 9_000, 0, None))
    specs.append((c1, sc1, 3_000, 0, None))
    sensors, hits, forms = [], [], []
    for k, (cfg, scene, n, form, shape) in enumerate(specs):
        s = R.make_sensor(cfg)
        s.set_sample_count(n)
        if shape is not None:
            s.set_tile_shape(shape)
        h = R.generate_hits(scene, 60 + k, max(n, 1)).to_device().prefix(n)
        sensors.append(s)
        hits.append(h)
        forms.append(form)
    eager = []
    for s, h, f in zip(sensors, hits, forms):
        if h.n:
            s.write_samples(h)
        s.compute_output(f)
        eager.append((s.output().copy(), s.binner_debug()))
    b = R.Batch(sensors, forms)
    for k, h in enumerate(hits):
        if h.n:
            b.bind_hits(k, h)
    stream = R.Stream()
    for rep in range(2):
        for s in sensors:
            if s.sample_count():
                s.upload_samples(np.zeros((s.sample_count(), 4), np.float32))
        b.compute(stream)
        stream.sync()
        for k, ((ping, dbg), s) in enumerate(zip(eager, sensors)):
            d = s.binner_debug()
            for key in ("order", "keys", "segments", "windows"):
                assert np.array_equal(d[key], dbg[key]), (rep, k, key)
            assert np.array_equal(s.output().view(np.uint32), ping.view(np.uint32)), (rep, k)


def test_batch_pose_update_is_applied_in_place(gpu):
    """set_pose between batched computes: the next replay streams the hits
    with the new rotation (rewritten in the batch's descriptors, not
    re-captured) and equals an eager compute with that pose bit for bit; the
    MFMA ping is within its bound of the oracle's ping of the rotated records."""
    R = gpu
    c0 = configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "p0")
    c1 = configs._fls(48, 96, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "p1")
    hits = [R.generate_hits(configs.fls_scene(), 70 + k, n) for k, n in enumerate((30_000, 20_000))]
    dh = [h.to_device() for h in hits]
    sensors = [R.make_sensor(c) for c in (c0, c1)]
    for s, h in zip(sensors, hits):
        s.set_sample_count(h.n)
    b = R.Batch(sensors, [2, 0])
    for k, h in enumerate(dh):
        b.bind_hits(k, h)
    stream = R.Stream()
    b.compute(stream)
    stream.sync()
    before = sensors[0].output().copy()
    c = np.cos(0.2)
    sn = np.sin(0.2)
    rot = np.array([[c, -sn, 0], [sn, c, 0], [0, 0, 1]], np.float32)
    for rep in range(2):
        sensors[0].set_pose(rot if rep == 0 else np.eye(3, dtype=np.float32))
        b.bind_hits(0, dh[0])                 # same arrays: the graph is kept
        b.compute(stream)
        stream.sync()
        got = sensors[0].output()
        cr = configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "pr")
        cr.rotation = rot if rep == 0 else np.eye(3, dtype=np.float32)
        e = R.make_sensor(cr)
        e.set_sample_count(hits[0].n)
        e.write_samples(dh[0])
        e.compute_output(2)
        assert np.array_equal(got.view(np.uint32), e.output().view(np.uint32)), rep
        if rep == 0:
            assert O.rel_l2(got, before) > 1e-2       # the rotation changed the ping
            rec = O.stream(cr, hits[0]).astype(np.float32)
            ref, _ = O.reference_ping(cr, rec)
            assert O.rel_l2(got, ref) <= 1e-3
        else:
            assert np.array_equal(got.view(np.uint32), before.view(np.uint32))
