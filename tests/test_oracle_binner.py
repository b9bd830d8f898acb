"""Pins the oracle's Binner restatement (Binner.h:25-54, Binner.cu:51-136,
Receiver.cu:9-42) against golden constants, closed forms and brute force
written independently of oracle/oracle.cpp."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

OOR = O.OOR
GRIDS = [(2048, 0.0, 50.0), (1024, 0.1, 30.0), (4096, 0.1, 120.0)]


def test_key_constants_golden():
    # evaluated FP32 constants of the 512x1024 [0.1, 30] m grid
    k = O.key_make(1024, 0.1, 30.0, 0.0)
    assert np.float32(k.a) == np.float32(34.180668)
    assert np.float32(k.b) == np.float32(-2.4190426)


def _key_constants_numpy(N, lo, hi, margin):
    """Binner.h:37-44 restated with numpy float32 scalars (no contraction)."""
    f = np.float32
    lo, hi, margin = f(lo), f(hi), f(margin)
    res = (hi - lo) / f(N)
    m = lo - f(0.5) * res
    M = hi + f(0.5) * res
    a = f(N - 1) / (M - m)
    b = (-m) * a + f(0.5)
    return a, b, m - margin, M + margin


@pytest.mark.parametrize("N,lo,hi", GRIDS)
@pytest.mark.parametrize("margin", [0.0, 0.5, 3.0])
def test_key_constants_match_numpy_float32(N, lo, hi, margin):
    k = O.key_make(N, lo, hi, margin)
    a, b, lower, upper = _key_constants_numpy(N, lo, hi, margin)
    assert (np.float32(k.a), np.float32(k.b)) == (a, b)
    assert (np.float32(k.lower), np.float32(k.upper)) == (lower, upper)


def _round_f32(q: Fraction) -> float:
    """Round an exact rational to the nearest float32 (ties to even)."""
    if q == 0:
        return 0.0
    sign = -1 if q < 0 else 1
    q = abs(q)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    if Fraction(2) ** e > q:
        e -= 1
    scale = Fraction(2) ** (23 - e)
    m = q * scale
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return float(np.float32(float(sign * fl / scale)))


def _fmaf(a, t, b) -> float:
    if math.isinf(t):
        return float(np.float32(a) * np.float32(t))
    return _round_f32(Fraction(float(a)) * Fraction(float(t)) + Fraction(float(b)))


def _key_ref(a, b, lower, upper, N, t):
    """Binner.h:48-53 with an exactly rounded float32 fma (rational arithmetic)
    and truncation toward zero."""
    t = np.float32(t)
    if math.isnan(t):
        return OOR
    v = np.float32(_fmaf(np.float32(a), t, np.float32(b)))
    if math.isnan(v):
        return OOR
    if v >= 2.0 ** 31:
        res = 2 ** 31 - 1
    elif v <= -2.0 ** 31:
        res = -2 ** 31
    else:
        res = int(v)                                  # trunc toward zero
    if res < 0:
        return 0 if np.float32(res) >= lower else OOR
    if res >= N:
        return N - 1 if np.float32(res) <= upper else OOR
    return res


@pytest.mark.parametrize("N,lo,hi", GRIDS)
def test_key_matches_independent_restatement(N, lo, hi):
    rng = np.random.default_rng(N)
    k = O.key_make(N, lo, hi, 0.7)
    t = np.concatenate([rng.uniform(lo - 3, hi + 3, 4000).astype(np.float32),
                        np.float32([-1.0, 0.0, -0.0, np.inf, -np.inf, np.nan, 1e30, -1e30])])
    got = O.keys(k, t)
    want = [_key_ref(k.a, k.b, k.lower, k.upper, N, x) for x in t]
    assert list(got) == want


@pytest.mark.parametrize("N,lo,hi", GRIDS)
def test_key_of_output_centres_is_index(N, lo, hi):
    k = O.key_make(N, lo, hi, 0.0)
    r = (np.float32(lo) + np.arange(N) * ((np.float32(hi) - np.float32(lo)) / np.float32(N - 1))
         ).astype(np.float32)
    assert np.array_equal(O.keys(k, r), np.arange(N, dtype=np.uint32))


@pytest.mark.parametrize("N,lo,hi", GRIDS)
def test_in_range_keys_are_monotone(N, lo, hi):
    k = O.key_make(N, lo, hi, 0.0)
    t = np.sort(np.random.default_rng(1).uniform(lo - 1, hi + 1, 200_000).astype(np.float32))
    kk = O.keys(k, t).astype(np.int64)
    inr = kk != OOR
    assert np.all(np.diff(kk[inr]) >= 0)
    # out-of-range samples only at both ends of the sorted array
    idx = np.nonzero(inr)[0]
    assert inr[idx[0]:idx[-1] + 1].all()


def test_key_clamp_branches():
    # margin large enough that lower <= -1 and upper >= N: both clamps live
    N, lo, hi = 16, 0.0, 4.0
    k = O.key_make(N, lo, hi, 40.0)
    a, b = k.a, k.b
    # a travel whose fma lands in (-2, -1] -> res = -1 -> clamped to 0
    t = np.float32((-1.5 - b) / a)
    assert O.key_eval(k, t) == 0
    # a travel whose fma lands above N -> clamped to N-1
    t = np.float32((N + 3.5 - b) / a)
    assert O.key_eval(k, t) == N - 1
    # far beyond the margins -> out of range
    assert O.key_eval(k, np.float32(-1000.0)) == OOR
    assert O.key_eval(k, np.float32(1000.0)) == OOR
    # without margin the same travels are out of range (res compared with metres)
    k0 = O.key_make(N, lo, hi, 0.0)
    assert O.key_eval(k0, np.float32((-1.5 - b) / a)) == OOR


def test_key_zero_covers_fma_in_open_unit_interval():
    k = O.key_make(1024, 0.1, 30.0, 0.0)
    for v in (-0.999, -0.5, 0.0, 0.5, 0.999):
        t = np.float32((v - k.b) / k.a)
        assert O.key_eval(k, t) == 0


# ---------------------------------------------------------------------------
# sort order
# ---------------------------------------------------------------------------
def test_sort_order_brute_force():
    rng = np.random.default_rng(5)
    t = np.round(rng.uniform(-2, 5, 3000), 1).astype(np.float32)
    t[::50] = np.nan
    t[1::40] = -0.0
    t[2::45] = 0.0
    t[3::60] = np.inf
    t[4::70] = -np.inf
    t[5::33] = -1.0

    def key(i):
        x = float(t[i])
        if math.isnan(x):
            return (2, 0.0, i)
        return (1, x, i)          # -0.0 == 0.0 compares equal in python
    want = sorted(range(t.size), key=key)
    assert list(O.sort_order(t)) == want


def test_travel_image_order():
    vals = np.float32([-np.inf, -1e30, -1.0, -1e-30, -0.0, 0.0, 1e-30, 1.0, 1e30, np.inf, np.nan])
    img = [O.travel_image(v) for v in vals]
    assert img[4] == img[5]                    # -0 == +0
    assert all(img[i] < img[i + 1] for i in range(len(img) - 1) if i != 4)


# ---------------------------------------------------------------------------
# segments and windows: independent formulations
# ---------------------------------------------------------------------------
def _segments_ref(keys, N):
    seg = np.zeros((N, 2), np.uint32)
    for k in set(int(x) for x in keys if x != OOR):
        pos = np.nonzero(keys == k)[0]
        seg[k] = (pos[0], pos[-1] + 1)
    return seg


def _windows_ref(seg, ov):
    """View of bin r: the non-empty bins of [r-ov, r+ov] clipped to the grid;
    fewer than two of them -> empty, else [start(first), end(last))."""
    N = seg.shape[0]
    out = np.zeros((N, 2), np.uint32)
    ov = max(0, min(ov, N))
    for r in range(N):
        nz = [i for i in range(max(0, r - ov), min(N - 1, r + ov) + 1) if seg[i, 1] != seg[i, 0]]
        if len(nz) >= 2:
            out[r] = (seg[nz[0], 0], seg[nz[-1], 1] - seg[nz[0], 0])
    return out


@pytest.mark.parametrize("seed", range(6))
def test_segments_and_windows_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    N = int(rng.integers(1, 64))
    n = int(rng.integers(0, 400))
    # sorted keys: OOR prefix, monotone in-range keys with gaps, OOR suffix
    pre, suf = int(rng.integers(0, 20)), int(rng.integers(0, 20))
    body = np.sort(rng.choice(N, size=n, replace=True)).astype(np.uint32) if N else []
    keys = np.concatenate([np.full(pre, OOR, np.uint32), body,
                           np.full(suf, OOR, np.uint32)]).astype(np.uint32)
    seg = O.segments(keys, N)
    assert np.array_equal(seg, _segments_ref(keys, N))
    # partition invariant
    assert int((seg[:, 1] - seg[:, 0]).sum()) == int((keys != OOR).sum())
    for ov in (0, 1, 2, 5, N + 3):
        assert np.array_equal(O.windows(seg, ov), _windows_ref(seg, ov)), (N, ov)


@pytest.mark.parametrize("N,lo,hi", GRIDS)
def test_overlap_closed_form(N, lo, hi):
    step = (np.float32(hi) - np.float32(lo)) / np.float32(N - 1)
    for ov in (0, 1, 9, 17, 33):
        margin = np.float32(ov * float(step) * (1 + 2 ** -10))
        assert O.overlap(margin, lo, hi, N) == int(np.float32(margin) / step)
        assert O.overlap(margin, lo, hi, N) == ov


def test_single_nonempty_bin_window_is_empty():
    # Binner.cu:96-97: a window holding one non-empty bin yields an empty view
    seg = np.zeros((8, 2), np.uint32)
    seg[3] = (0, 10)
    assert not O.windows(seg, 2).any()
    seg[5] = (10, 12)
    w = O.windows(seg, 2)
    assert tuple(w[4]) == (0, 12) and tuple(w[3]) == (0, 12) and tuple(w[1]) == (0, 0)


def _group_ref(cfg, t):
    """Grouped order of the grid form restated with exact rational fmas: bin
    slot from the (pinned) key, out-of-range samples split at the range
    midpoint (every OOR travel lies beyond one end), G row from the rounded
    fma, then a stable lexicographic sort."""
    range0, inv_dr, off_r, fr, GH = O.group_params(cfg)
    k = O.key_make(cfg.range_count, cfg.range_min, cfg.range_max, O.effective_margin(cfg))
    keys = O.keys(k, t)
    mid = 0.5 * (cfg.range_min + cfg.range_max)
    slot, row = [], []
    for x, key in zip(t, keys):
        x = np.float32(x)
        if key != OOR:
            slot.append(int(key) + 1)
        else:
            slot.append(0 if (not math.isnan(x) and x < mid) else cfg.range_count + 1)
        if math.isnan(x):
            row.append(GH + 2)
            continue
        d = np.float32(x - np.float32(range0))
        y = d if math.isinf(d) else np.float32(_fmaf(d, np.float32(inv_dr), np.float32(off_r)))
        if math.isnan(y):
            row.append(GH + 2)
        elif math.isinf(y):
            row.append(GH + 2 if y > 0 else -2)
        else:
            row.append(int(min(max(math.floor(y) + fr, -2), GH + 2)))
    return np.lexsort((np.arange(t.size), np.array(row), np.array(slot))).astype(np.uint32)


def test_group_order_matches_independent_restatement():
    from rtac_simulation_amd import configs
    cfg = configs._fls(48, 96, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "g")
    rng = np.random.default_rng(11)
    t = rng.uniform(-1.0, 31.0, 3000).astype(np.float32)
    t[::7] = t[::7][::-1]                        # repeated travels (ties)
    t[::97] = -1.0                               # null samples
    t[1::101] = np.nan
    t[2::103] = -0.0
    t[3::107] = np.inf
    r = O.group_params(cfg)
    edges = ((np.arange(-3, 40) - r[2] - r[3]) / r[1] + r[0]).astype(np.float32)
    t[4::11][:edges.size] = edges[:t[4::11].size]   # travels on G-row boundaries
    got = O.group_order(cfg, t)
    assert np.array_equal(got, _group_ref(cfg, t))


def test_group_order_is_travel_order_within_runs():
    """Sorted by the grouped order, the keys and hence segments and windows
    equal those of the travel order; inside each equal-key run the travels
    are a permutation of the travel-sorted ones."""
    from rtac_simulation_amd import configs
    cfg = configs._fls(64, 256, 0.1, 30.0, 1.2e6, configs.psf_31x17(), "g")
    rng = np.random.default_rng(12)
    t = np.round(rng.uniform(-0.5, 30.5, 50_000), 3).astype(np.float32)
    t[::500] = -1.0
    t[1::700] = np.nan
    og, ot = O.group_order(cfg, t), O.sort_order(t)
    assert np.array_equal(np.sort(og), np.arange(t.size))
    k = O.key_make(cfg.range_count, cfg.range_min, cfg.range_max, O.effective_margin(cfg))
    kg, kt = O.keys(k, t[og]), O.keys(k, t[ot])
    assert np.array_equal(kg, kt)
    seg = O.segments(kg, cfg.range_count)
    for b in range(cfg.range_count):
        s0, s1 = seg[b]
        assert np.array_equal(np.sort(t[og][s0:s1]), t[ot][s0:s1])
    assert not np.array_equal(og, ot)            # the orders do differ inside runs
