"""Pins the oracle's floating-point restatements: directivity lookup
(Directivity.h:30-34, Directivity.cpp:109-166), sample weighting
(RayCaster.cu:62-74, Receiver.h:51-60), PSF (ReductionKernel.h:44-73) and the
sparse gathers (SensorInstance{1D,2D}.cu), each against closed forms or a
brute force written independently with numpy/scipy."""
import numpy as np
import pytest
from scipy.special import j1

import oracle as O
import rtac_simulation_amd as R
from rtac_simulation_amd import configs


# ---------------------------------------------------------------------------
# directivity
# ---------------------------------------------------------------------------
def _dir_for_angles(u_ang, v_ang):
    """Unit vector whose atan2(y, x) = u_ang and atan2(z, x) = v_ang (x > 0)."""
    x = 1.0
    return np.array([x, np.tan(u_ang), np.tan(v_ang)]) / np.sqrt(1 + np.tan(u_ang) ** 2
                                                                  + np.tan(v_ang) ** 2)


def test_directivity_texel_centres_and_bilinear():
    rng = np.random.default_rng(0)
    W, H = 40, 30
    tab = rng.uniform(0, 1, (H, W)).astype(np.float32)
    s = np.float32(1 / np.pi)
    for (w, h) in [(3, 4), (10, 2), (19, 14)]:          # texels in the front quarter
        ua, va = np.pi * (w + 0.5) / W, np.pi * (h + 0.5) / H
        d = _dir_for_angles(ua / (np.pi * s) * np.pi * s, va)
        got = O.directivity(tab, s, s, d)
        assert abs(got - tab[h, w]) < 1e-5
        # halfway between two texels along u = the mean
        ua2 = np.pi * (w + 1.0) / W
        got2 = O.directivity(tab, s, s, _dir_for_angles(ua2, va))
        assert abs(got2 - 0.5 * (tab[h, w] + tab[h, w + 1])) < 1e-5


def test_directivity_mirror_symmetry():
    rng = np.random.default_rng(1)
    tab = rng.uniform(0, 1, (16, 16)).astype(np.float32)
    s = np.float32(1 / np.pi)
    for _ in range(50):
        ua, va = rng.uniform(-1.2, 1.2, 2)
        d = _dir_for_angles(ua, va)
        dm = _dir_for_angles(-ua, -va)
        assert abs(O.directivity(tab, s, s, d) - O.directivity(tab, s, s, dm)) < 1e-12


def test_disk_table_matches_scipy_piston():
    lam, D = 0.03947, 0.3327
    tab = R.directivity_disk(D, lam, True, 8)
    N = tab.shape[0]
    assert N == max(64, 8 * int(8 * np.pi / (2 * lam / D)))
    ka = np.pi * D / lam
    # along h = 0 (elevation 0) the off-axis angle equals the bearing pi w / N
    for w in (1, 5, 20, 60, N // 4):
        th = np.pi * w / N
        a = ka * np.sin(th)
        want = 2 * j1(a) / a
        assert abs(tab[0, w] - want) < 1e-5
    assert tab[0, 0] == 1.0
    assert not tab[N // 2 + 1:, :].any() and not tab[:, N // 2 + 1:].any()   # single-sided


# ---------------------------------------------------------------------------
# sample weighting
# ---------------------------------------------------------------------------
def test_stream_closed_form():
    cfg, scene, S, seed = configs.workload("fls_512x1024", samples=3000)
    hits = R.generate_hits(scene, seed, S)
    out = O.stream(cfg, hits)
    r = hits.range.astype(np.float64)
    null = hits.range < 1e-4
    assert (out[null, 2] == -1).all() and (out[null, :2] == 0).all()
    ok = ~null
    A = hits.amplitude[0::2].astype(np.float64) + 1j * hits.amplitude[1::2]
    v = out[:, 0] + 1j * out[:, 1]
    d = np.stack([hits.dir_x, hits.dir_y, hits.dir_z], 1).astype(np.float64)
    s = np.float32(1 / np.pi)
    D = np.array([O.directivity(cfg.directivity, s, s, d[i]) for i in np.nonzero(ok)[0]])
    ph = 2 * np.pi * np.mod(2 * r[ok] * cfg.frequency / cfg.sound_speed, 1.0)
    want = A[ok] * np.exp(1j * ph) * D
    assert np.allclose(v[ok], want, rtol=1e-9, atol=1e-12)
    assert np.allclose(out[ok, 3], np.arctan2(-d[ok, 1], d[ok, 0]), atol=1e-12)
    assert np.array_equal(out[ok, 2], r[ok])


def test_stream_pose_rotation():
    cfg, scene, S, seed = configs.workload("fls_512x1024", samples=500)
    hits = R.generate_hits(scene, seed, S)
    c = np.cos(0.3); s_ = np.sin(0.3)
    Rz = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]], np.float32)
    cfg.rotation = Rz
    out = O.stream(cfg, hits)
    d = np.stack([hits.dir_x, hits.dir_y, hits.dir_z], 1).astype(np.float64)
    loc = d @ Rz.astype(np.float64)            # R^T d for row vectors
    ok = hits.range >= 1e-4
    assert np.allclose(out[ok, 3], np.arctan2(-loc[ok, 1], loc[ok, 0]), atol=1e-12)


# ---------------------------------------------------------------------------
# PSF
# ---------------------------------------------------------------------------
def test_psf_taps_support_and_separability():
    rng = np.random.default_rng(2)
    pr = rng.standard_normal(7) + 1j * rng.standard_normal(7)
    pb = rng.standard_normal(5)
    taps = np.outer(pr, pb).astype(np.complex64)
    span_r, span_b = 7 * 0.03, 5 * 0.004
    pitch_r, pitch_b = span_r / 7, span_b / 5
    for i in range(7):
        for j in range(5):
            dr, dx = (i - 3) * pitch_r, (j - 2) * pitch_b
            assert abs(O.psf(taps, span_r, span_b, dr, dx) - taps[i, j]) < 1e-6
    # zero beyond (T + 1)/2 pitches
    assert O.psf(taps, span_r, span_b, 4.0 * pitch_r + 1e-9, 0.0) == 0
    assert O.psf(taps, span_r, span_b, 0.0, -3.0 * pitch_b - 1e-9) == 0
    # separable bilinear == product of 1-D hat interpolations (numpy interp)
    xr = (np.arange(7) - 3) * pitch_r
    xb = (np.arange(5) - 2) * pitch_b
    ext_r = np.concatenate([[xr[0] - pitch_r], xr, [xr[-1] + pitch_r]])
    ext_b = np.concatenate([[xb[0] - pitch_b], xb, [xb[-1] + pitch_b]])
    pr64 = pr.astype(np.complex64).astype(np.complex128)
    for _ in range(200):
        dr, dx = rng.uniform(-4.5 * pitch_r, 4.5 * pitch_r), rng.uniform(-3.5 * pitch_b, 3.5 * pitch_b)
        kr = np.interp(dr, ext_r, np.r_[0, pr64.real, 0], left=0, right=0) + \
            1j * np.interp(dr, ext_r, np.r_[0, pr64.imag, 0], left=0, right=0)
        kb = np.interp(dx, ext_b, np.r_[0, pb, 0], left=0, right=0)
        want = np.outer(kr, kb).item() if False else kr * kb
        got = O.psf(taps, span_r, span_b, dr, dx)
        # the table holds the float32 products; compare to that precision
        assert abs(got - want) < 1e-5 * (1 + abs(want))


# ---------------------------------------------------------------------------
# gathers
# ---------------------------------------------------------------------------
def _tiny_cfg(W=12, H=20, Tr=5, Tb=3, seed=3):
    rng = np.random.default_rng(seed)
    taps = (rng.standard_normal((Tr, Tb)) + 1j * rng.standard_normal((Tr, Tb))).astype(np.complex64)
    return configs._fls(W, H, 1.0, 3.0, 1.2e6, taps, "tiny")


def _brute_conv(cfg, records, win):
    """out(r,b) = sum_{s in view(r)} v_s K(x_s - x_b, t_s - r_r) in numpy, with
    K from a direct bilinear formula."""
    taps = cfg.psf.astype(np.complex64).astype(np.complex128)
    Tr, Tb = taps.shape
    rr = np.float64(np.float32(cfg.range_min)) + np.arange(cfg.H) * cfg.range_step
    xb = cfg.bearing_coords()
    out = np.zeros((cfg.H, cfg.W), np.complex128)
    for r in range(cfg.H):
        st, sz = win[r]
        for s in records[st:st + sz]:
            v = complex(s[0], s[1])
            yB = ((float(s[2]) - rr[r]) / cfg.psf_range_span + 0.5) * Tr - 0.5
            xs = np.sin(np.float64(s[3])) if cfg.bearing_axis == 1 else np.float64(s[3])
            for b in range(cfg.W):
                xB = ((xs - xb[b]) / cfg.psf_bearing_span + 0.5) * Tb - 0.5
                acc = 0j
                for j in range(Tr):
                    wy = max(0.0, 1 - abs(yB - j))
                    if wy == 0:
                        continue
                    for i in range(Tb):
                        wx = max(0.0, 1 - abs(xB - i))
                        acc += wy * wx * taps[j, i]
                out[r, b] += v * acc
    return out


def _tiny_records(cfg, n, seed):
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, 4), np.float32)
    rec[:, 0], rec[:, 1] = rng.standard_normal(n), rng.standard_normal(n)
    rec[:, 2] = rng.uniform(0.9, 3.1, n)
    rec[:, 3] = np.arcsin(rng.uniform(-0.95, 0.95, n))
    return rec


def test_conv2d_brute_force():
    cfg = _tiny_cfg()
    rec = _tiny_records(cfg, 150, 4)
    ref, b = O.reference_ping(cfg, rec, threads=2)
    want = _brute_conv(cfg, b["sorted"], b["windows"])
    assert np.allclose(ref, want, rtol=1e-12, atol=1e-12)


def test_conv2d_single_sample_is_shifted_psf():
    cfg = _tiny_cfg(W=16, H=24, Tr=5, Tb=3)
    cfg.margin = configs.overlap_margin(8, cfg.range_step)
    rec = np.zeros((3, 4), np.float32)
    # three samples in one bin so the window holds >= 2 non-empty bins? no: a
    # single non-empty bin gives an empty view, so place 2 samples in
    # neighbouring bins and subtract the second's contribution
    rr = np.float32(cfg.range_min) + np.float32(cfg.range_step) * np.arange(cfg.H)
    rec[0] = (1.0, 0.0, rr[10] + 0.3 * cfg.range_step, np.arcsin(0.05))
    rec[1] = (0.0, 0.0, rr[12], 0.0)               # zero-valued companion
    rec[2] = (0.0, 0.0, rr[5], 0.0)
    ref, b = O.reference_ping(cfg, rec, threads=1)
    xb = cfg.bearing_coords()
    rows = [r for r in range(cfg.H) if b["windows"][r, 1] > 0]
    assert rows
    for r in rows:
        for bb in range(cfg.W):
            want = O.psf(cfg.psf, cfg.psf_range_span, cfg.psf_bearing_span,
                         float(rec[0, 2]) - float(np.float64(np.float32(cfg.range_min))
                                                  + r * cfg.range_step),
                         np.sin(np.float64(rec[0, 3])) - xb[bb])
            assert abs(ref[r, bb] - want) < 1e-9


def test_conv2d_linearity():
    cfg = _tiny_cfg(W=20, H=30, Tr=7, Tb=5, seed=8)
    rec = _tiny_records(cfg, 400, 9)
    b = O.binning(cfg, rec)
    y1 = O.conv(cfg, b["sorted"], b["windows"], threads=2)
    s2 = b["sorted"].copy()
    s2[:, :2] *= np.float32(2.0)
    y2 = O.conv(cfg, s2, b["windows"], threads=2)
    assert np.allclose(y2, 2 * y1, rtol=1e-12, atol=1e-12)


def test_conv1d_matches_2d_with_one_bearing():
    cfg1, scene, S, seed = configs.workload("single_beam_38k", samples=2000)
    hits = R.generate_hits(scene, seed, S)
    rec = O.stream(cfg1, hits).astype(np.float32)
    y1, b = O.reference_ping(cfg1, rec, threads=2)
    cfg2 = configs.SensorConfig(kind=2, range_min=cfg1.range_min, range_max=cfg1.range_max,
                                range_count=cfg1.range_count, psf=cfg1.psf,
                                psf_range_span=cfg1.psf_range_span, psf_bearing_span=1.0,
                                margin=cfg1.margin, bearings=np.float32([0.0]))
    y2 = O.conv(cfg2, b["sorted"], b["windows"], threads=2)
    assert np.allclose(y1, y2[:, 0], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("Tr,Tb", [(5, 3), (7, 5), (6, 4)])
def test_grid_form_equals_plain_form(Tr, Tb):
    """Deposit + correlation + mask == the gather when the axes are uniform at
    the tap pitch and the overlap is >= (Tr + 3) / 2 (odd and even taps)."""
    cfg = _tiny_cfg(W=24, H=40, Tr=Tr, Tb=Tb, seed=Tr * 10 + Tb)
    cfg.margin = configs.overlap_margin(configs.required_overlap(Tr), cfg.range_step)
    rec = _tiny_records(cfg, 2000, 11)
    rec[::7, 2] = rec[::7, 2] + 5.0            # some out of range
    plain, b = O.reference_ping(cfg, rec, threads=2)
    u = cfg.bearing_coords()
    pitch = (u[-1] - u[0]) / (cfg.W - 1)
    grid = O.grid_form(cfg, rec, b["windows"], u[0], pitch)
    assert O.rel_l2(grid, plain) < 1e-6


def test_grid_form_matches_scipy_correlate():
    from scipy.signal import correlate2d
    cfg = _tiny_cfg(W=16, H=24, Tr=5, Tb=3, seed=21)
    cfg.margin = configs.overlap_margin(24, cfg.range_step)
    rng = np.random.default_rng(22)
    # samples exactly on grid nodes: the deposit is a plain delta grid
    n = 60
    rows = rng.integers(3, cfg.H - 3, n)
    cols = rng.integers(2, cfg.W - 2, n)
    u = cfg.bearing_coords()
    pitch = (u[-1] - u[0]) / (cfg.W - 1)
    rec = np.zeros((n, 4), np.float32)
    rec[:, 0], rec[:, 1] = rng.standard_normal(n), rng.standard_normal(n)
    rec[:, 2] = (np.float64(np.float32(cfg.range_min)) + rows * cfg.range_step).astype(np.float32)
    rec[:, 3] = np.arcsin(u[0] + cols * pitch).astype(np.float32)
    b = O.binning(cfg, rec)
    got = O.grid_form(cfg, rec, b["windows"], u[0], pitch)
    G = np.zeros((cfg.H, cfg.W), np.complex128)
    # exact node positions after float32 rounding of travel / bearing
    y = (rec[:, 2].astype(np.float64) - np.float64(np.float32(cfg.range_min))) / cfg.range_step
    x = (np.sin(rec[:, 3].astype(np.float64)) - u[0]) / pitch
    for k in range(n):
        iy, ix = int(np.floor(y[k] + 0.5)), int(np.floor(x[k] + 0.5))
        G[iy, ix] += rec[k, 0] + 1j * rec[k, 1]
    taps = cfg.psf.astype(np.complex64).astype(np.complex128)
    want = correlate2d(G, np.conj(taps), mode="same")       # sum_k G[n+k] T[k]
    mask = b["windows"][:, 1] == 0
    want[mask] = 0
    assert O.rel_l2(got, want) < 1e-4


def test_stream_inverse_square_spreading_closed_form():
    """RayCaster.cu:69-71 divides the value by the squared travel: at travels
    whose round trip holds a whole number of wavelengths (phase 0), a unit
    amplitude along the axis of a uniform receiver reads exactly 1 / r^2."""
    n = 6
    f, c = 768e3, 1500.0
    lam2 = c / (2 * f)                       # one cycle of 2 r f / c: 2^-10 m, exact in FP32
    r = (np.array([1, 3, 10, 40, 200, 1000], np.float64) * lam2).astype(np.float32)
    hits = R.HostHits(n)
    hits.range[:] = r
    hits.dir_x[:], hits.dir_y[:], hits.dir_z[:] = 1.0, 0.0, 0.0
    hits.amplitude[0::2], hits.amplitude[1::2] = 1.0, 0.0
    cfg = configs._fls(8, 16, 0.1, 2.0, f, configs.psf_31x17(), "spread")
    cfg.directivity = None
    cfg.spreading = R.SPREAD_INV_R2
    out = O.stream(cfg, hits)
    q = 2.0 * r.astype(np.float64) * f / c
    ph = 2 * np.pi * (q - np.floor(q))
    assert not ph.any()                     # the travels sit on whole cycles
    want = 1.0 / r.astype(np.float64) ** 2
    assert np.allclose(out[:, 0], want, rtol=1e-12, atol=0)
    assert np.abs(out[:, 1]).max() < 1e-15 * want.max()
    cfg.spreading = R.SPREAD_NONE
    flat = O.stream(cfg, hits)
    assert np.allclose(np.abs(flat[:, 0] + 1j * flat[:, 1]), 1.0, rtol=1e-12)


@pytest.mark.parametrize("axis", [0, 1])
def test_conv2d_rows_equals_full_gather(axis):
    """The row-subset gather (each sample visits only the bearings in the PSF
    support) gives the full gather's rows: the skipped terms are exact 0."""
    rng = np.random.default_rng(11)
    taps = (rng.standard_normal((7, 5)) + 1j * rng.standard_normal((7, 5))).astype(np.complex64)
    cfg = configs._fls(40, 30, 1.0, 3.0, 1.2e6, taps, "rows")
    if axis == 0:
        cfg.bearing_axis = 0
        cfg.bearings = np.sort(rng.uniform(-1.1, 1.1, 40)).astype(np.float32)
        cfg.psf_bearing_span = float(np.float32(5 * 2.2 / 39))
    rec = _tiny_records(cfg, 4000, 12)
    full, b = O.reference_ping(cfg, rec, threads=2)
    rows = np.array([0, 3, 7, 15, 29])
    got = O.conv_rows(cfg, b["sorted"], b["windows"], rows, threads=2)
    assert np.abs(full[rows]).max() > 0
    assert O.rel_l2(got, full[rows]) < 1e-14


def test_cpu_baseline_gather_computes_the_ping():
    """The CPU baseline (FP32 port of sort -> Binner -> per-cell gather) must
    compute the ping it is timed on: its binning equals the oracle's bit for
    bit and its rows are within the plain form's 1e-5 of the FP64 gather."""
    cfg, scene, _, seed = configs.workload("fls_512x1024", samples=10)
    cfg = configs._fls(96, 160, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "cpu_base")
    hits = R.generate_hits(scene, seed, 60_000)
    rec = O.stream(cfg, hits).astype(np.float32)
    srt, win = O.cpu_binning(cfg, rec, 4)
    b = O.binning(cfg, rec)
    assert np.array_equal(srt.view(np.uint32), b["sorted"].view(np.uint32))
    assert np.array_equal(win, b["windows"].reshape(-1))
    rows = np.arange(2, cfg.H, 9)
    base = np.zeros((rows.size, cfg.W), np.complex64)
    for i, r in enumerate(rows):
        base[i] = O.cpu_gather2d(cfg, srt, win, int(r), int(r) + 1, 2)[r]
    ref = O.conv_rows(cfg, b["sorted"], b["windows"], rows, threads=4)
    assert np.abs(ref).max() > 0
    assert O.rel_l2(base, ref) <= 1e-5


def test_cpu_baseline_1d_gather_computes_the_ping():
    """The 1D CPU baseline (FP32 port of SensorInstance1D.cu:35-72) against the
    FP64 oracle gather on the echo-sounder workload.  Like the reference it
    forms t - r_r in FP32: at 50 m one ulp of travel is 1.6e-4 of the 24 mm
    tap pitch, so the port is held to 1e-4 (the HIP plain form, with
    double-float range coordinates, meets 1e-5)."""
    cfg, scene, S, seed = configs.workload("single_beam_38k")
    hits = R.generate_hits(scene, seed, S)
    rec = O.stream(cfg, hits).astype(np.float32)
    srt, win = O.cpu_binning(cfg, rec, 2)
    base = O.cpu_gather1d(cfg, srt, win, 2)
    ref, _ = O.reference_ping(cfg, rec, threads=2)
    assert np.abs(ref).max() > 0
    assert O.rel_l2(base, ref) <= 1e-4
