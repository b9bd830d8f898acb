"""GPU parity at the full sizes of the large and batched workloads, through the
C ABI, against the FP64 oracle:

  * the 1024 x 4096 ping from 16 000 000 samples with the 63 x 33 PSF
    (SensorInstance2D.cu:44-83), plain FP32 and both MFMA forms: Binner keys,
    segments, windows and the travel order bit-exact (after the MFMA forms,
    which bin without sorting, Binner introspection sorts by travel), every
    64th range row within the form's bound of the oracle;
  * the 64 uneven pings (16 x 256x512, 32 x 512x1024, 8 x 1024x2048 and
    8 x 1D 2048, 1e5..4e6 samples each; plain and both MFMA forms, 1D pings
    included) computed by one batched graph launch, each ping checked the
    same way.

At these sizes the sort's tile-count scan runs its LDS-staged branch (more
than 2048 tiles per pass) and pings of up to 16M samples go through every
kernel.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

import oracle as O
from rtac_simulation_amd import configs

pytestmark = pytest.mark.gpu

PLAIN_TOL = 1e-5
MFMA_TOL = 1e-3


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _check_binning(name, dbg, b, order):
    assert dbg["overlap"] == b["overlap"], name
    for key in ("keys", "segments", "windows"):
        assert np.array_equal(dbg[key], b[key]), (name, key)
    assert np.array_equal(dbg["order"], order), (name, "order")


def _check_rows(name, cfg, ping, b, rows, tol):
    if cfg.kind == 2:
        ref = O.conv_rows(cfg, b["sorted"], b["windows"], rows)
        got = ping[rows]
    else:
        ref = O.conv(cfg, b["sorted"], b["windows"])
        got = ping
    assert np.abs(ref).max() > 0, name
    err = O.rel_l2(got, ref)
    assert err <= tol, (name, err)
    return err


@pytest.fixture(scope="module")
def large(gpu):
    R = gpu
    cfg, scene, S, seed = configs.workload("fls_1024x4096")
    print(f"large ping: generating {S} hits", flush=True)
    dh = R.generate_hits(scene, seed, S, nthreads=_threads()).to_device()
    s = R.make_sensor(cfg)
    s.set_sample_count(S)
    s.write_samples(dh)
    rec = s.samples()
    b = O.binning(cfg, rec)
    print("large ping: oracle binning done", flush=True)
    return R, cfg, s, dh, rec, b


@pytest.mark.parametrize("form", [0, 1, 2])
def test_large_ping_vs_oracle(large, form):
    R, cfg, s, dh, rec, b = large
    s.write_samples(dh)
    s.compute_output(form)
    ping = s.output()
    dbg = s.binner_debug()
    _check_binning(f"large form {form}", dbg, b, b["order"])
    rows = np.arange(5, cfg.H, 64)
    err = _check_rows(f"large form {form}", cfg, ping, b, rows,
                      PLAIN_TOL if form == 0 else MFMA_TOL)
    print(f"large ping form {form}: rel_l2 {err:.3e} on {rows.size} rows", flush=True)


def _batch_forms(items):
    """Plain FP32 for every fourth ping, the MFMA forms (BF16 and FP16 in
    turn) for the others, 1D pings included."""
    return [0 if i % 4 == 3 else (1 if i % 2 else 2) for i in range(len(items))]


def test_uneven_batch_vs_oracle(gpu):
    R = gpu
    items = configs.batch_uneven()
    big = {}
    sensors = []
    forms = _batch_forms(items)
    for cfg, scene, S, seed in items:
        if cfg.kind not in big:
            big[cfg.kind] = R.generate_hits(scene, seed, 4_000_000, nthreads=_threads()).to_device()
        s = R.make_sensor(cfg)
        s.set_sample_count(S)
        sensors.append(s)
    b = R.Batch(sensors, forms)
    for i, (cfg, _, S, _) in enumerate(items):
        b.bind_hits(i, big[cfg.kind].prefix(S))
    stream = R.Stream()
    b.compute(stream)
    stream.sync()
    total = sum(S for _, _, S, _ in items)
    print(f"batch: 64 pings, {total} samples computed", flush=True)

    def check(i):
        cfg, _, S, _ = items[i]
        s = sensors[i]
        rec = s.samples()
        bo = O.binning(cfg, rec)
        return bo, bo["order"]

    pings = [s.output() for s in sensors]
    dbgs = [s.binner_debug() for s in sensors]
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        oracles = list(ex.map(check, range(len(items))))
    worst = {0: 0.0, 1: 0.0, 2: 0.0}
    for i, ((cfg, _, S, _), (bo, order)) in enumerate(zip(items, oracles)):
        name = f"batch ping {i} ({cfg.W}x{cfg.H}, {S} samples, form {forms[i]})"
        _check_binning(name, dbgs[i], bo, order)
        rows = np.arange(3 + i % 7, cfg.H, 64)
        err = _check_rows(name, cfg, pings[i], bo, rows, PLAIN_TOL if forms[i] == 0 else MFMA_TOL)
        worst[forms[i]] = max(worst[forms[i]], err)
    print(f"batch: worst rel_l2 per form {worst}", flush=True)
