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


def test_window_emptiness_is_two_occupied_bins():
    """The grid form keeps only which windows hold samples and finds them
    from key occupancy: a window is non-empty exactly when at least two bins
    of [r - ov, r + ov] (clipped to [0, N)) hold samples -- the reading of
    Binner.cu:90-106 (skip empty bins inward, istart >= iend is empty) this
    checks against the oracle's windows on random occupancies."""
    rng = np.random.default_rng(21)
    for trial in range(200):
        N = int(rng.integers(1, 80))
        ov = int(rng.integers(0, 12))
        occ = rng.random(N) < rng.uniform(0.02, 0.6)
        seg = np.zeros((N, 2), np.uint32)
        pos = 0
        for k in range(N):
            if occ[k]:
                n = int(rng.integers(1, 5))
                seg[k] = (pos, pos + n)
                pos += n
        win = O.windows(seg, ov)
        for r in range(N):
            lo, hi = max(0, r - ov), min(N - 1, r + ov)
            assert (win[r, 1] > 0) == (occ[lo:hi + 1].sum() >= 2), (trial, r)
