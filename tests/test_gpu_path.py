"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Bars: binning (sort order, keys, segments, windows, overlap) bit-exact;
records from the streaming kernel within 1e-6 relative (values) and exact
travel; plain FP32 pings within 1e-5 relative L2 of the FP64 oracle.
"""
import numpy as np
import pytest

import oracle as O
from rtac_simulation_amd import configs

pytestmark = pytest.mark.gpu

PLAIN_TOL = 1e-5


def _run(R, cfg, hits, form=0):
    s = R.make_sensor(cfg)
    s.set_sample_count(hits.n)
    s.write_samples(hits.to_device())
    s.compute_output(form)
    return s


def _check_binning(s, cfg, rec=None):
    if rec is None:
        rec = s.samples()
    b = O.binning(cfg, rec)
    d = s.binner_debug()
    assert d["overlap"] == b["overlap"]
    assert np.array_equal(d["order"], b["order"]), "sort order"
    assert np.array_equal(d["keys"], b["keys"]), "keys"
    assert np.array_equal(d["segments"], b["segments"]), "segments"
    assert np.array_equal(d["windows"], b["windows"]), "windows"
    # bitwise: NaN travels must land in the same places
    assert np.array_equal(s.samples(sorted=True).view(np.uint32),
                          b["sorted"].view(np.uint32)), "sorted records"
    return b


@pytest.mark.parametrize("name", ["single_beam_38k", "fls_512x1024_sep"])
def test_stream_kernel_matches_oracle(gpu, name):
    R = gpu
    cfg, scene, S, seed = configs.workload(name, samples=200_000 if "fls" in name else None)
    hits = R.generate_hits(scene, seed, S)
    s = R.make_sensor(cfg)
    s.set_sample_count(hits.n)
    s.write_samples(hits.to_device())
    rec = s.samples()
    ref = O.stream(cfg, hits)
    assert np.array_equal(rec[:, 2], ref[:, 2].astype(np.float32)), "travel must pass through"
    v = rec[:, 0] + 1j * rec[:, 1]
    vr = ref[:, 0] + 1j * ref[:, 1]
    assert O.rel_l2(v, vr) < 2e-6
    assert np.max(np.abs(rec[:, 3] - ref[:, 3])) < 2e-6


@pytest.mark.parametrize("name", ["single_beam_38k", "fls_512x1024_sep", "fls_512x1024"])
def test_binning_bit_exact_full_size(gpu, name):
    R = gpu
    cfg, scene, S, seed = configs.workload(name)
    hits = R.generate_hits(scene, seed, S)
    s = _run(R, cfg, hits)
    _check_binning(s, cfg)


def test_plain_1d_full_size(gpu):
    R = gpu
    cfg, scene, S, seed = configs.workload("single_beam_38k")
    hits = R.generate_hits(scene, seed, S)
    s = _run(R, cfg, hits)
    rec = s.samples()
    ref, _ = O.reference_ping(cfg, rec)
    assert O.rel_l2(s.output(), ref) <= PLAIN_TOL


@pytest.mark.parametrize("W,H,psf", [(64, 128, "sep"), (48, 96, "31x17"), (33, 70, "31x17")])
def test_plain_2d_small(gpu, W, H, psf):
    R = gpu
    taps = np.outer(configs.baseband_pulse(15), configs.sinc2_beam(9, 3.0)) if psf == "sep" \
        else configs.psf_31x17()
    cfg = configs._fls(W, H, 0.1, 30.0, 1.2e6, taps, f"t{W}x{H}")
    hits = R.generate_hits(configs.fls_scene(), 11, 30_000)
    s = _run(R, cfg, hits)
    b = _check_binning(s, cfg)
    ref = O.conv(cfg, b["sorted"], b["windows"])
    assert O.rel_l2(s.output(), ref) <= PLAIN_TOL


def test_plain_2d_full_size_row_subset(gpu):
    """512x1024, 1M samples, 31x17 PSF: every 32nd range row against the FP64
    oracle (the oracle's full gather is O(W sum|view|))."""
    R = gpu
    cfg, scene, S, seed = configs.workload("fls_512x1024")
    hits = R.generate_hits(scene, seed, S)
    s = _run(R, cfg, hits)
    b = _check_binning(s, cfg)
    rows = np.arange(5, cfg.H, 32)
    win = np.zeros_like(b["windows"])
    win[rows] = b["windows"][rows]
    ref = O.conv(cfg, b["sorted"], win)
    assert O.rel_l2(s.output()[rows], ref[rows]) <= PLAIN_TOL


def test_angle_axis_reference_semantics(gpu):
    """Bearing PSF axis in radians (KernelView2D), non-uniform bearing table."""
    R = gpu
    cfg = configs._fls(40, 80, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "angle")
    cfg.bearing_axis = 0
    cfg.psf_bearing_span = float(np.float32(17 * np.deg2rad(130.0) / 39))
    hits = R.generate_hits(configs.fls_scene(), 12, 25_000)
    s = _run(R, cfg, hits)
    b = _check_binning(s, cfg)
    ref = O.conv(cfg, b["sorted"], b["windows"])
    assert O.rel_l2(s.output(), ref) <= PLAIN_TOL


# ---------------------------------------------------------------------------
# edge cases the reference's Binner meets (Binner.cu:51-136)
# ---------------------------------------------------------------------------
def _records_case(case, n=5000, rng=None):
    rng = rng or np.random.default_rng(0)
    rec = np.zeros((n, 4), np.float32)
    rec[:, 0] = rng.standard_normal(n)
    rec[:, 1] = rng.standard_normal(n)
    rec[:, 3] = rng.uniform(-1.0, 1.0, n)
    if case == "ties":
        rec[:, 2] = np.round(rng.uniform(0.1, 30.0, n), 1)       # many equal travels
    elif case == "all_oor":
        rec[:, 2] = rng.uniform(31.0, 40.0, n)
    elif case == "all_null":
        rec[:, 2] = -1.0
    elif case == "nan_inf_zero":
        rec[:, 2] = rng.uniform(0.0, 32.0, n)
        rec[::97, 2] = np.nan
        rec[1::89, 2] = np.inf
        rec[2::83, 2] = -np.inf
        rec[3::79, 2] = -0.0
        rec[4::73, 2] = 0.0
    elif case == "single_bin":
        rec[:, 2] = 15.0
    return rec


@pytest.mark.parametrize("case", ["ties", "all_oor", "all_null", "nan_inf_zero", "single_bin"])
def test_binner_edge_cases(gpu, case):
    R = gpu
    cfg = configs._fls(32, 64, 0.1, 30.0, 1.2e6, configs.psf_31x17(), case)
    rec = _records_case(case)
    s = R.make_sensor(cfg)
    s.set_sample_count(rec.shape[0])
    s.upload_samples(rec)
    s.compute_output(0)
    b = _check_binning(s, cfg, rec)
    ref = O.conv(cfg, b["sorted"], b["windows"])
    out = s.output()
    if np.abs(ref).max() == 0:
        assert np.abs(out).max() == 0
    else:
        assert O.rel_l2(out, ref) <= PLAIN_TOL


def test_zero_samples_gives_zero_ping(gpu):
    R = gpu
    cfg = configs._fls(32, 64, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "empty")
    s = R.make_sensor(cfg)
    s.set_sample_count(0)
    s.compute_output(0)
    assert np.abs(s.output()).max() == 0
    d = s.binner_debug()
    assert not d["segments"].any() and not d["windows"].any()


@pytest.mark.parametrize("ov", [0, 1, 3])
def test_small_overlap_windows(gpu, ov):
    """Reference default overlaps (1 for the adaptive waveform) and ov = 0
    (every window empty by the istart >= iend rule)."""
    R = gpu
    cfg = configs._fls(32, 64, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "ov")
    cfg.margin = configs.overlap_margin(ov, cfg.range_step) if ov else 0.0
    hits = R.generate_hits(configs.fls_scene(), 13, 20_000)
    s = _run(R, cfg, hits)
    b = _check_binning(s, cfg)
    assert b["overlap"] == ov
    ref = O.conv(cfg, b["sorted"], b["windows"])
    out = s.output()
    if ov == 0:
        assert np.abs(out).max() == 0 and np.abs(ref).max() == 0
    else:
        assert O.rel_l2(out, ref) <= PLAIN_TOL


def test_set_ranges_reconfigures(gpu):
    R = gpu
    cfg = configs._fls(32, 64, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "ranges")
    hits = R.generate_hits(configs.fls_scene(), 14, 20_000)
    s = _run(R, cfg, hits)
    cfg2 = configs._fls(32, 100, 5.0, 25.0, 1.2e6, configs.psf_31x17(), "ranges2")
    cfg2.psf_range_span = cfg.psf_range_span
    s.set_ranges(5.0, 25.0, 100, cfg2.margin)
    s.compute_output(0)
    b = _check_binning(s, cfg2)
    ref = O.conv(cfg2, b["sorted"], b["windows"])
    assert O.rel_l2(s.output(), ref) <= PLAIN_TOL


def test_repeated_compute_is_stable(gpu):
    R = gpu
    cfg = configs._fls(32, 64, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "repeat")
    hits = R.generate_hits(configs.fls_scene(), 15, 20_000)
    s = _run(R, cfg, hits)
    first = s.output()
    o1 = s.binner_debug()["order"]
    for _ in range(3):
        s.compute_output(0)
    assert np.array_equal(o1, s.binner_debug()["order"])
    # the gather sums every cell in a fixed order: bit-reproducible
    assert np.array_equal(s.output().view(np.uint32), first.view(np.uint32))


@pytest.mark.parametrize("case", ["ties", "all_oor", "all_null", "nan_inf_zero", "single_bin", "empty"])
def test_mfma_binner_edge_cases(gpu, case):
    """The grid form's bucketing on the edge cases: the ping within the MFMA
    tolerance (or zero), Binner introspection after it bit-exact."""
    R = gpu
    cfg = configs._fls(32, 64, 0.1, 30.0, 1.2e6, configs.psf_31x17(), case)
    rec = _records_case(case) if case != "empty" else np.zeros((0, 4), np.float32)
    s = R.make_sensor(cfg)
    s.set_sample_count(rec.shape[0])
    if rec.shape[0]:
        s.upload_samples(rec)
    for form in (1, 2):
        s.compute_output(form)
        b = O.binning(cfg, rec)
        d = s.binner_debug()
        for key in ("order", "keys", "segments", "windows"):
            assert np.array_equal(d[key], b[key]), (form, key)
        ref = O.conv(cfg, b["sorted"], b["windows"])
        out = s.output()
        if np.abs(ref).max() == 0:
            assert np.abs(out).max() == 0, form
        else:
            assert O.rel_l2(out, ref) <= 1e-3, form


def test_fused_stream_kernel_writes_identical_records(gpu):
    """The streaming kernel fused with the grid form's bucket counts (used
    once the grid form is prepared) writes the same records, bit for bit, as
    the plain streaming kernel."""
    R = gpu
    cfg = configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "fusedrec")
    h = R.generate_hits(configs.fls_scene(), 60, 30_000).to_device()
    a = R.make_sensor(cfg)
    a.set_sample_count(h.n)
    a.write_samples(h)
    b = R.make_sensor(cfg)
    b.set_sample_count(h.n)
    b.set_tile_shape(0)                 # prepares the grid form: the next write is fused
    b.write_samples(h)
    assert np.array_equal(a.samples().view(np.uint32), b.samples().view(np.uint32))


def test_measure_peaks_in_hardware_range(gpu):
    """rtsim_measure_peaks reports figures an MI355X can reach: copy
    bandwidth above 2 TB/s and at most the 8 TB/s HBM3E peak; dense MFMA
    rates above 250 TFLOP/s and at most the 2.5 PFLOP/s FP16/BF16 peak."""
    p = gpu.measure_peaks()
    assert 2000.0 < p["hbm_copy_gbs"] <= 8000.0, p
    for k in ("mfma_fp16_tflops", "mfma_bf16_tflops"):
        assert 250.0 < p[k] <= 2500.0, p
