"""GPU parity of the MFMA grid form (BF16 / FP16 hi+lo split, FP32
accumulation) against the FP64 oracle: relative L2 <= 1e-3 (tolerance of the
MFMA form), through the C ABI."""
import numpy as np
import pytest

import oracle as O
from rtac_simulation_amd import configs

pytestmark = pytest.mark.gpu

MFMA_TOL = 1e-3
FORMS = [1, 2]          # RTSIM_FORM_MFMA_BF16, RTSIM_FORM_MFMA_FP16


def _sensor(R, cfg, hits):
    s = R.make_sensor(cfg)
    s.set_sample_count(hits.n)
    s.write_samples(hits.to_device())
    return s


def _psf(kind):
    if kind == "15x9":
        return np.outer(configs.baseband_pulse(15), configs.sinc2_beam(9, 3.0))
    if kind == "31x17":
        return configs.psf_31x17()
    if kind == "21x33":
        return configs.perturbed(np.outer(configs.lfm_pulse(21, 0.03),
                                          configs.sinc2_beam(33, 8.0)), -30.0, 9)
    if kind == "8x6":
        return configs.perturbed(np.outer(configs.baseband_pulse(8), configs.sinc2_beam(6, 2.0)),
                                 -30.0, 10)
    raise KeyError(kind)


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("W,H,psf", [(48, 96, "31x17"), (64, 128, "15x9"), (40, 100, "21x33"),
                                     (37, 70, "8x6")])
def test_mfma_small_vs_oracle(gpu, form, W, H, psf):
    R = gpu
    cfg = configs._fls(W, H, 0.1, 30.0, 1.2e6, _psf(psf), f"m{W}x{H}")
    hits = R.generate_hits(configs.fls_scene(), 21, 40_000)
    s = _sensor(R, cfg, hits)
    s.compute_output(form)
    got = s.output()
    ref, b = O.reference_ping(cfg, s.samples())
    err = O.rel_l2(got, ref)
    assert err <= MFMA_TOL, err
    # masked rows are exactly zero
    empty = b["windows"][:, 1] == 0
    assert not np.abs(got[empty]).any()


@pytest.mark.parametrize("form", FORMS)
def test_mfma_full_size_vs_plain_and_oracle(gpu, form):
    R = gpu
    cfg, scene, S, seed = configs.workload("fls_512x1024")
    hits = R.generate_hits(scene, seed, S)
    s = _sensor(R, cfg, hits)
    s.compute_output(0)
    plain = s.output()
    s.compute_output(form)
    mf = s.output()
    assert O.rel_l2(mf, plain) <= MFMA_TOL
    b = O.binning(cfg, s.samples())
    rows = np.arange(7, cfg.H, 64)
    win = np.zeros_like(b["windows"])
    win[rows] = b["windows"][rows]
    ref = O.conv(cfg, b["sorted"], win)
    assert O.rel_l2(mf[rows], ref[rows]) <= MFMA_TOL


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("S", [20_000, 3_000_000])
def test_mfma_1d_vs_oracle(gpu, form, S):
    """A 1D ping in the grid form (one column, one tap column): the echo
    sounder's 33-tap ping (SensorInstance1D.cu:35-72) against the FP64 oracle,
    its masked bins exactly zero, its binning bit-exact."""
    R = gpu
    cfg, scene, _, seed = configs.workload("single_beam_38k")
    hits = R.generate_hits(scene, seed, S)
    s = _sensor(R, cfg, hits)
    s.compute_output(form)
    got = s.output()
    rec = s.samples()
    ref, b = O.reference_ping(cfg, rec)
    err = O.rel_l2(got, ref)
    assert err <= MFMA_TOL, err
    assert not np.abs(got[b["windows"][:, 1] == 0]).any()
    d = s.binner_debug()
    for key in ("order", "keys", "segments", "windows"):
        assert np.array_equal(d[key], b[key]), key


def test_mfma_rejects_what_it_cannot_restate(gpu):
    R = gpu
    # 1D sensor whose PSF pitch (two range steps) differs from the range step
    cfg1, scene1, S1, seed1 = configs.workload("single_beam_38k")
    cfg1.psf_range_span *= 2.0
    s = _sensor(R, cfg1, R.generate_hits(scene1, seed1, 1000))
    with pytest.raises(R.RtsimError) as e:
        s.compute_output(1)
    assert e.value.code == R.ERR_UNSUPPORTED
    # bearing table not uniform in the PSF axis (reference angle axis on a sine grid)
    cfg = configs._fls(32, 64, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "angle")
    cfg.bearing_axis = 0
    s = _sensor(R, cfg, R.generate_hits(configs.fls_scene(), 3, 1000))
    with pytest.raises(R.RtsimError) as e:
        s.compute_output(2)
    assert e.value.code == R.ERR_UNSUPPORTED
    # overlap below the PSF support: the windowed gather truncates the PSF
    cfg = configs._fls(32, 64, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "ov")
    cfg.margin = configs.overlap_margin(3, cfg.range_step)
    s = _sensor(R, cfg, R.generate_hits(configs.fls_scene(), 3, 1000))
    with pytest.raises(R.RtsimError) as e:
        s.compute_output(1)
    assert e.value.code == R.ERR_UNSUPPORTED
    # the plain form still serves it
    s.compute_output(0)


@pytest.mark.parametrize("form", FORMS)
def test_mfma_dynamic_range(gpu, form):
    """Samples spanning 60 dB: per-tile power-of-two scaling keeps FP16 in range."""
    R = gpu
    cfg = configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "dr")
    hits = R.generate_hits(configs.fls_scene(), 22, 40_000)
    rng = np.random.default_rng(0)
    g = 10.0 ** (-3.0 * rng.uniform(0, 1, hits.n)).astype(np.float32)
    hits.amplitude[0::2] *= g
    hits.amplitude[1::2] *= g
    s = _sensor(R, cfg, hits)
    s.compute_output(form)
    ref, _ = O.reference_ping(cfg, s.samples())
    assert O.rel_l2(s.output(), ref) <= MFMA_TOL


@pytest.mark.parametrize("waves", [4, 8])
@pytest.mark.parametrize("shape", [0, 1, 2, 3])
@pytest.mark.parametrize("W,H,psf", [(48, 96, "31x17"), (37, 70, "8x6"), (40, 100, "21x33")])
def test_mfma_every_tile_shape_vs_oracle(gpu, shape, waves, W, H, psf):
    """Every correlation tile shape and wave count the autotuner may pick
    restates the gather."""
    R = gpu
    cfg = configs._fls(W, H, 0.1, 30.0, 1.2e6, _psf(psf), f"t{W}x{H}")
    hits = R.generate_hits(configs.fls_scene(), 23, 30_000)
    s = _sensor(R, cfg, hits)
    s.set_tile_shape(shape)
    s.set_tile_waves(waves)
    assert s.tile_shape()[0] == shape and s.tile_waves() == waves
    ref = None
    for form in FORMS:
        s.compute_output(form)
        got = s.output()
        if ref is None:
            ref, b = O.reference_ping(cfg, s.samples())
        assert O.rel_l2(got, ref) <= MFMA_TOL, (form, shape)
        empty = b["windows"][:, 1] == 0
        assert not np.abs(got[empty]).any()


def test_mfma_autotune_keeps_results(gpu):
    R = gpu
    cfg = configs._fls(96, 160, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "tune")
    hits = R.generate_hits(configs.fls_scene(), 24, 60_000)
    s = _sensor(R, cfg, hits)
    s.compute_output(2)
    before = s.output()
    shape, ms = s.autotune(2, iterations=3)
    assert 0 <= shape < 4 and ms > 0
    assert s.tile_shape()[0] == shape
    s.compute_output(2)
    after = s.output()
    assert O.rel_l2(after, before) <= MFMA_TOL
    # a second sensor of the same class starts on the cached shape
    s2 = _sensor(R, cfg, hits)
    s2.compute_output(2)
    assert s2.tile_shape()[0] == shape


@pytest.mark.parametrize("form", FORMS)
def test_mfma_bitwise_reproducible(gpu, form):
    """The deposit sums in exact fixed point and the correlation has a fixed
    reduction order, so two computes of the same samples agree bit for bit."""
    R = gpu
    cfg = configs._fls(96, 192, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "rep")
    hits = R.generate_hits(configs.fls_scene(), 25, 80_000)
    s = _sensor(R, cfg, hits)
    s.compute_output(form)
    a = s.output().view(np.uint32).copy()
    s.write_samples(hits.to_device())
    s.compute_output(form)
    b = s.output().view(np.uint32)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("form", FORMS)
def test_mfma_binning_full_size(gpu, form):
    """The grid form bins by grid row without sorting; its window mask comes
    from the Binner-key occupancy of the bucketing: every row whose oracle
    window is empty is exactly zero, the other rows hold the ping.  Binner
    introspection after it (a sort by travel) is bit-exact against the
    oracle: order, keys, segments, windows and the sorted records."""
    R = gpu
    cfg, scene, S, seed = configs.workload("fls_512x1024")
    hits = R.generate_hits(scene, seed, S)
    s = _sensor(R, cfg, hits)
    s.compute_output(form)
    ping = s.output()
    rec = s.samples()
    b = O.binning(cfg, rec)
    empty = b["windows"][:, 1] == 0
    assert empty.any() and (~empty).any()
    assert not np.abs(ping[empty]).any()
    assert (np.abs(ping[~empty]).max(axis=1) > 0).mean() > 0.99
    d = s.binner_debug()
    for key in ("order", "keys", "segments", "windows"):
        assert np.array_equal(d[key], b[key]), key
    assert np.array_equal(s.samples(sorted=True).view(np.uint32), b["sorted"].view(np.uint32))


@pytest.mark.parametrize("form", FORMS)
def test_mfma_ping_independent_of_sample_order(gpu, form):
    """Permuting the input records changes the order of the items inside
    every bucket but not the ping: the deposit sums in exact fixed point, so
    the two pings agree bit for bit."""
    R = gpu
    cfg = configs._fls(128, 256, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "perm")
    hits = R.generate_hits(configs.fls_scene(), 27, 150_000)
    s = _sensor(R, cfg, hits)
    s.compute_output(form)
    a = s.output().view(np.uint32).copy()
    rec = s.samples()
    perm = np.random.default_rng(3).permutation(rec.shape[0])
    s.upload_samples(np.ascontiguousarray(rec[perm]))
    s.compute_output(form)
    assert np.array_equal(s.output().view(np.uint32), a)


def test_mfma_tall_grid(gpu):
    """33000 range bins: 2066 buckets of 16 grid rows and 1032 words of key
    occupancy per tile; binning stays exact and the ping matches the oracle."""
    R = gpu
    cfg = configs._fls(64, 33000, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "tall")
    hits = R.generate_hits(configs.fls_scene(), 31, 200_000)
    s = _sensor(R, cfg, hits)
    s.compute_output(2)
    rec = s.samples()
    b = O.binning(cfg, rec)
    rows = np.arange(3, cfg.H, 97)
    ref = O.conv_rows(cfg, b["sorted"], b["windows"], rows)
    assert O.rel_l2(s.output()[rows], ref) <= MFMA_TOL
    d = s.binner_debug()
    for key in ("order", "keys", "segments", "windows"):
        assert np.array_equal(d[key], b[key]), key


def _key_interval(k, inside):
    """Ends of the float interval of travels with a Binner key, by bisection
    over the ordered float32 images with the oracle's key function."""
    def img(f):
        u = int(np.float32(f).view(np.uint32))
        return (~u & 0xFFFFFFFF) if u & 0x80000000 else (u | 0x80000000)

    def flt(u):
        b = (u & 0x7FFFFFFF) if u & 0x80000000 else (~u & 0xFFFFFFFF)
        return float(np.uint32(b).view(np.float32))

    def search(out, inn):
        while abs(inn - out) > 1:
            mid = (inn + out) // 2
            if O.key_eval(k, flt(mid)) == 0xFFFFFFFF:
                out = mid
            else:
                inn = mid
        return flt(inn)
    return search(img(-np.inf), img(inside)), search(img(np.inf), img(inside))


@pytest.mark.gpu
@pytest.mark.parametrize("form", [1, 2])
def test_mfma_key_interval_edges(gpu, form):
    """Samples exactly at both ends of the in-range travel interval are
    deposited and their outside neighbours (one float away, 1000x louder,
    inside the grid's PSF halo) are not: the MFMA ping matches the oracle."""
    R = gpu
    cfg = configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "edges")
    k = O.key_make(cfg.range_count, cfg.range_min, cfg.range_max, O.effective_margin(cfg))
    lo, hi = _key_interval(k, 0.5 * (cfg.range_min + cfg.range_max))
    assert O.key_eval(k, lo) == 0 and O.key_eval(k, hi) == cfg.range_count - 1
    below = float(np.nextafter(np.float32(lo), np.float32(-np.inf)))
    above = float(np.nextafter(np.float32(hi), np.float32(np.inf)))
    assert O.key_eval(k, below) == 0xFFFFFFFF and O.key_eval(k, above) == 0xFFFFFFFF
    step = (cfg.range_max - cfg.range_min) / (cfg.range_count - 1)
    halo = (cfg.Tr - 1) // 2 * step                      # grid rows beyond both range ends
    assert 0 < cfg.range_min - lo < halo and 0 < hi - cfg.range_max < halo
    rng = np.random.default_rng(5)
    n = 4000
    rec = np.zeros((n + 4, 4), np.float32)
    rec[:n, 0] = rng.normal(size=n)
    rec[:n, 1] = rng.normal(size=n)
    rec[:n, 2] = rng.uniform(cfg.range_min, cfg.range_max, n)
    rec[:n, 3] = rng.uniform(-1.0, 1.0, n)
    rec[n:, 2] = [lo, hi, below, above]
    rec[n:, 0] = [1.0, 1.0, 1000.0, 1000.0]
    rec[n:, 3] = [0.1, -0.2, 0.1, -0.2]
    s = R.make_sensor(cfg)
    s.set_sample_count(rec.shape[0])
    s.upload_samples(rec)
    s.compute_output(form)
    ref, _ = O.reference_ping(cfg, rec)
    assert O.rel_l2(s.output(), ref) <= MFMA_TOL


@pytest.mark.gpu
def test_mfma_eight_wave_correlation_large_psf(gpu):
    """A 63 x 33 PSF needs so much LDS per workgroup that one fits per CU:
    the correlation then runs 8 waves per workgroup (JS doubles, so every
    wave takes half as many tap steps and the partial sums of 2 to 8 waves
    meet in LDS).  Every shape matches the oracle on sampled rows."""
    R = gpu
    cfg = configs._fls(512, 1024, 0.1, 120.0, 750e3, configs.psf_63x33(), "wide_psf")
    _, scene, _, _ = configs.fls_1024x4096()
    hits = R.generate_hits(scene, 7, 400_000)
    s = _sensor(R, cfg, hits)
    rec = s.samples()
    b = O.binning(cfg, rec)
    rows = np.arange(5, cfg.H, 61)
    win = np.zeros_like(b["windows"])
    win[rows] = b["windows"][rows]
    ref = O.conv(cfg, b["sorted"], win)
    tested = 0
    for sh in range(4):
        try:
            s.set_tile_shape(sh)
        except R.RtsimError:
            continue
        s.compute_output(2)
        assert O.rel_l2(s.output()[rows], ref[rows]) <= MFMA_TOL, sh
        tested += 1
    assert tested >= 2


@pytest.mark.gpu
def test_mfma_bucket_counts_fused_into_write_samples(gpu):
    """Once the grid form is prepared, a write of every sample also makes the
    bucket counts; the computes that use them give the same ping (bit for
    bit) and binning as computes that count themselves, and writes by other
    means, partial writes and other forms discard them."""
    R = gpu
    cfg = configs._fls(128, 256, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "fused")
    hits = R.generate_hits(configs.fls_scene(), 29, 120_000)
    dh = hits.to_device()
    s = R.make_sensor(cfg)
    s.set_sample_count(hits.n)
    s.write_samples(dh)                      # grid form not prepared yet: plain stream kernel
    s.compute_output(2)
    ref = s.output().view(np.uint32).copy()
    b = O.binning(cfg, s.samples())
    s.compute_output(2)                      # bucket counts made by the compute itself
    assert np.array_equal(s.output().view(np.uint32), ref)
    for form in (2, 1, 2):
        s.write_samples(dh)                  # fused
        s.compute_output(form)
        if form == 2:
            assert np.array_equal(s.output().view(np.uint32), ref)
        d = s.binner_debug()
        for key in ("keys", "segments", "windows"):
            assert np.array_equal(d[key], b[key]), (form, key)
    # a fused write followed by the plain form, then the grid form (unfused)
    s.write_samples(dh)
    s.compute_output(0)
    ref_plain, _ = O.reference_ping(cfg, s.samples())
    assert O.rel_l2(s.output(), ref_plain) <= 1e-5
    s.compute_output(2)
    assert np.array_equal(s.output().view(np.uint32), ref)
    # a partial write (the same hits of the second half) is not fused and
    # discards the fused histogram of the full write before it
    s.write_samples(dh)
    n1 = hits.n // 2
    tail = R.generate_hits(configs.fls_scene(), 29, hits.n - n1, first=n1).to_device()
    s.write_samples(tail, offset=n1)
    s.compute_output(2)
    assert np.array_equal(s.output().view(np.uint32), ref)
