"""The emitter's ray fan, initial values and ray hits (Emitter.cu:38-80,
Emitter.h:54-63), host side: pinned against an independent numpy float32
restatement of generate_polar_directions, the oracle's FP64 directivity
lookup, and ray / plane geometry in closed form."""
import math

import numpy as np
import pytest

import oracle as O
import rtac_simulation_amd as R


def _polar_ref(res_deg, bearing_deg, elevation_deg):
    """Emitter.cu:59-80 restated with numpy float32 scalars: `x *= M_PI / 180.0f`
    widens to double and rounds back; every later operation is FP32."""
    f = np.float32
    k = math.pi / float(np.float32(180.0))
    res, bap, eap = (f(float(f(v)) * k) for v in (res_deg, bearing_deg, elevation_deg))
    with np.errstate(invalid="ignore", divide="ignore"):
        H = int(eap / res) + 1
        out = []
        for h in range(H):
            phi = eap * (f(h) / f(H - 1) - f(0.5))
            wr = f(abs(np.cos(phi))) * bap / res
            if math.isnan(wr):
                continue                      # (int) NaN is INT_MIN on x86: no rays
            W = int(wr) + 1
            for w in range(W):
                theta = bap * (f(w) / f(W - 1) - f(0.5))
                out.append((np.cos(theta) * np.cos(phi), np.sin(theta) * np.cos(phi), np.sin(phi)))
    return np.array(out, np.float32).reshape(-1, 3)


@pytest.mark.parametrize("res,bap,eap", [(0.5, 130.0, 20.0), (1.0, 90.0, 10.0), (0.25, 30.0, 3.0),
                                         (2.0, 60.0, 60.0)])
def test_polar_directions_match_restatement(res, bap, eap):
    got = R.emitter_polar_directions(res, bap, eap)
    ref = _polar_ref(res, bap, eap)
    assert got.shape == ref.shape
    finite = np.isfinite(ref).all(axis=1)
    assert np.array_equal(np.isfinite(got).all(axis=1), finite)
    np.testing.assert_allclose(got[finite], ref[finite], rtol=0, atol=2e-7)
    # unit vectors inside the apertures
    g = got[finite].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(g, axis=1), 1.0, atol=1e-6)
    assert np.all(np.abs(np.degrees(np.arctan2(g[:, 1], g[:, 0]))) <= bap / 2 + 1e-3)
    assert np.all(np.abs(np.degrees(np.arcsin(g[:, 2]))) <= eap / 2 + 1e-3)


def test_polar_fan_degenerate_rows():
    # elevation below the resolution: one row, phi = 0/0 -> no rays (reference x86 behaviour)
    assert R.emitter_polar_directions(1.0, 90.0, 0.5).shape == (0, 3)
    # rows whose |cos phi| bearing < resolution hold one ray with a NaN direction
    d = R.emitter_polar_directions(10.0, 15.0, 180.0)
    assert np.isnan(d[:, 0]).any()
    assert np.isnan(d[:, 0]).sum() == np.isnan(d[:, 1]).sum()


def test_initial_values_match_oracle_directivity():
    table = R.directivity_disk(0.05, 1500.0 / 1.2e6, True, 4)
    d = R.emitter_polar_directions(1.0, 120.0, 30.0)
    d = d[np.isfinite(d).all(axis=1)]
    v = R.emitter_initial_values(d, table)
    assert np.all(v.imag == 0)
    ref = np.array([O.directivity(table, 1 / math.pi, 1 / math.pi, x) for x in d.astype(np.float64)])
    np.testing.assert_allclose(v.real, ref, rtol=0, atol=2e-5 * np.abs(ref).max())
    assert np.all(R.emitter_initial_values(d) == 1)


def _flat_seabed(depth=5.0, power=1):
    sc = R.SceneDesc()
    sc.kind = R.SCENE_SEABED
    sc.plane_count = 1
    sc.planes[0][0], sc.planes[0][1], sc.planes[0][2], sc.planes[0][3] = 0.0, 0.0, 1.0, depth
    sc.rayleigh_sigma = 1.0
    sc.incidence_power = power
    return sc


def test_emitter_hits_on_a_plane_closed_form():
    sc = _flat_seabed(5.0, 2)
    d = R.emitter_polar_directions(1.0, 60.0, 40.0)
    d = d[np.isfinite(d).all(axis=1)]
    tilt = np.radians(30.0)   # pitch the fan down (x -> x cos t - z sin t) onto the seabed z = -5
    rot = np.array([[np.cos(tilt), 0, np.sin(tilt)], [0, 1, 0], [-np.sin(tilt), 0, np.cos(tilt)]],
                   np.float32)
    vals = R.emitter_initial_values(d)
    n = 3 * d.shape[0]
    h = R.generate_emitter_hits(sc, 9, d, vals, n, rotation=rot)
    world = (rot.astype(np.float64) @ np.tile(d, (3, 1)).astype(np.float64).T).T
    np.testing.assert_allclose(np.stack([h.dir_x, h.dir_y, h.dir_z], 1), world, atol=1e-6)
    nz = world[:, 2]
    hit = nz < -1e-9
    r_ref = np.where(hit, 5.0 / -np.where(hit, nz, -1.0), 0.0)
    np.testing.assert_allclose(h.range, r_ref.astype(np.float32), rtol=1e-6)
    assert hit.mean() > 0.5
    # amplitude = value * cos(incidence)^p * Rayleigh draw: doubling the values doubles it
    h2 = R.generate_emitter_hits(sc, 9, d, 2 * vals, n, rotation=rot)
    np.testing.assert_allclose(h2.amplitude, 2 * h.amplitude, rtol=1e-6, atol=0)
    assert np.all(h.amplitude.reshape(-1, 2)[~hit] == 0)


def test_emitter_hits_split_invariant():
    sc = _flat_seabed(4.0, 1)
    sc.height_noise = 0.05
    d = R.emitter_polar_directions(2.0, 90.0, 30.0)
    vals = R.emitter_initial_values(d)
    whole = R.generate_emitter_hits(sc, 3, d, vals, 5000, nthreads=4)
    parts = [R.generate_emitter_hits(sc, 3, d, vals, 1000, first=k * 1000, nthreads=1)
             for k in range(5)]
    for a in ("range", "dir_x", "dir_y", "dir_z"):
        assert np.array_equal(getattr(whole, a), np.concatenate([getattr(p, a) for p in parts]))
    assert np.array_equal(whole.amplitude, np.concatenate([p.amplitude for p in parts]))


def test_emitter_hits_reject_bad_arguments():
    d = R.emitter_polar_directions(2.0, 90.0, 30.0)
    vals = R.emitter_initial_values(d)
    cone = R.SceneDesc()
    cone.kind = R.SCENE_CONE
    with pytest.raises(R.RtsimError):
        R.generate_emitter_hits(cone, 1, d, vals, 10)
    with pytest.raises(ValueError):
        R.generate_emitter_hits(_flat_seabed(), 1, d, vals[:-1], 10)


@pytest.mark.gpu
@pytest.mark.parametrize("form", [0, 2])
def test_emitter_ping_vs_oracle(gpu, form):
    """Hits of an emitter fan through the streaming kernel and a form, against
    the FP64 oracle on the same records."""
    from rtac_simulation_amd import configs
    R_ = gpu
    cfg = configs._fls(64, 128, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "em")
    dirs = R_.emitter_polar_directions(0.25, 130.0, 20.0)
    vals = R_.emitter_initial_values(dirs, R_.directivity_disk(0.02, 1500.0 / 1.2e6, True, 8))
    hits = R_.generate_emitter_hits(configs.fls_scene(), 4, dirs, vals, dirs.shape[0])
    s = R_.make_sensor(cfg)
    s.set_sample_count(hits.n)
    s.write_samples(hits.to_device())
    s.compute_output(form)
    ref, _ = O.reference_ping(cfg, s.samples())
    assert O.rel_l2(s.output(), ref) <= (1e-5 if form == 0 else 1e-3)
