"""ctypes binding of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
OOR = 0xFFFFFFFF


class Key(C.Structure):
    _fields_ = [("n", C.c_uint32), ("a", C.c_float), ("b", C.c_float),
                ("lower", C.c_float), ("upper", C.c_float)]


class StreamParams(C.Structure):
    _fields_ = [("rotation", C.c_float * 9), ("frequency", C.c_float),
                ("sound_speed", C.c_float), ("spreading", C.c_int),
                ("directivity", C.POINTER(C.c_float)), ("dir_width", C.c_uint32),
                ("dir_height", C.c_uint32), ("dir_xscale", C.c_float),
                ("dir_yscale", C.c_float)]


class Grid(C.Structure):
    _fields_ = [("range_lo", C.c_float), ("range_hi", C.c_float), ("range_count", C.c_uint32),
                ("bearing_count", C.c_uint32), ("axis", C.c_int),
                ("psf_range_taps", C.c_int), ("psf_bearing_taps", C.c_int),
                ("psf_range_span", C.c_float), ("psf_bearing_span", C.c_float)]


_P = C.c_void_p
_SIGS = {
    "orc_key_make": (None, [C.c_uint32, C.c_float, C.c_float, C.c_float, C.POINTER(Key)]),
    "orc_key_eval": (C.c_uint32, [C.POINTER(Key), C.c_float]),
    "orc_keys": (None, [C.POINTER(Key), _P, C.c_size_t, C.c_size_t, _P]),
    "orc_travel_image": (C.c_uint32, [C.c_float]),
    "orc_sort_order": (None, [_P, C.c_size_t, C.c_size_t, _P]),
    "orc_segments": (None, [_P, C.c_size_t, C.c_uint32, _P]),
    "orc_overlap": (C.c_int, [C.c_float, C.c_float, C.c_float, C.c_uint32]),
    "orc_windows": (None, [_P, C.c_uint32, C.c_int, _P]),
    "orc_directivity": (C.c_double, [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                     C.c_double, C.c_double, C.c_double]),
    "orc_stream": (None, [C.c_size_t, _P, _P, _P, _P, _P, C.POINTER(StreamParams), _P]),
    "orc_psf": (None, [_P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                       C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "orc_conv2d": (None, [_P, _P, C.POINTER(Grid), _P, _P, _P, C.c_int]),
    "orc_conv1d": (None, [_P, _P, C.POINTER(Grid), _P, _P, C.c_int]),
    "orc_conv2d_rows": (None, [_P, _P, C.POINTER(Grid), _P, _P, _P, C.c_size_t, _P, C.c_int]),
    "orc_grid2d": (None, [_P, C.c_size_t, C.POINTER(Key), _P, C.POINTER(Grid), C.c_double,
                          C.c_double, _P, _P]),
    "cpu_binning": (None, [_P, C.c_size_t, C.POINTER(Grid), C.c_float, _P, _P, C.c_int]),
    "cpu_gather1d_f32": (None, [_P, _P, C.POINTER(Grid), _P, _P, C.c_int]),
    "cpu_gather2d_f32": (None, [_P, _P, C.POINTER(Grid), _P, _P, _P, C.c_long, C.c_long,
                                C.c_int]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags.c_contiguous, "array must be contiguous"
    return C.c_void_p(a.ctypes.data)


def nthreads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# ---------------------------------------------------------------------------
# Binner
# ---------------------------------------------------------------------------
def key_make(N, lo, hi, margin) -> Key:
    k = Key()
    lib().orc_key_make(N, lo, hi, margin, C.byref(k))
    return k


def key_eval(k: Key, t: float) -> int:
    return int(lib().orc_key_eval(C.byref(k), t))


def keys(k: Key, travel: np.ndarray) -> np.ndarray:
    t = np.ascontiguousarray(travel, np.float32)
    out = np.empty(t.size, np.uint32)
    lib().orc_keys(C.byref(k), _p(t), t.size, 1, _p(out))
    return out


def travel_image(t: float) -> int:
    return int(lib().orc_travel_image(t))


def sort_order(travel: np.ndarray) -> np.ndarray:
    t = np.ascontiguousarray(travel, np.float32)
    out = np.empty(t.size, np.uint32)
    lib().orc_sort_order(_p(t), t.size, 1, _p(out))
    return out


def segments(sorted_keys: np.ndarray, N: int) -> np.ndarray:
    k = np.ascontiguousarray(sorted_keys, np.uint32)
    out = np.empty(2 * N, np.uint32)
    lib().orc_segments(_p(k), k.size, N, _p(out))
    return out.reshape(N, 2)


def overlap(margin, lo, hi, N) -> int:
    return int(lib().orc_overlap(margin, lo, hi, N))


def windows(seg: np.ndarray, ov: int) -> np.ndarray:
    s = np.ascontiguousarray(seg, np.uint32).reshape(-1)
    N = s.size // 2
    out = np.empty(2 * N, np.uint32)
    lib().orc_windows(_p(s), N, ov, _p(out))
    return out.reshape(N, 2)


def binning(cfg, records: np.ndarray) -> dict:
    """Whole Binner chain of the oracle: order, sorted records, keys, segments,
    overlap, windows."""
    rec = np.ascontiguousarray(records, np.float32).reshape(-1, 4)
    order = sort_order(rec[:, 2])
    srt = np.ascontiguousarray(rec[order])
    k = key_make(cfg.range_count, cfg.range_min, cfg.range_max, effective_margin(cfg))
    skeys = keys(k, srt[:, 2])
    seg = segments(skeys, cfg.range_count)
    ov = overlap(effective_margin(cfg), cfg.range_min, cfg.range_max, cfg.range_count)
    win = windows(seg, ov)
    return {"order": order, "sorted": srt, "keys": skeys, "segments": seg, "overlap": ov,
            "windows": win, "key": k}


def effective_margin(cfg) -> float:
    return float(cfg.margin) if cfg.margin >= 0 else float(cfg.psf_range_span)


# ---------------------------------------------------------------------------
# streaming (directivity x phase)
# ---------------------------------------------------------------------------
def stream(cfg, hits) -> np.ndarray:
    """FP64 records (n, 4) for host hits (rtac_simulation_amd.HostHits)."""
    p = StreamParams()
    for i, v in enumerate(np.asarray(cfg.rotation, np.float32).ravel()):
        p.rotation[i] = float(v)
    p.frequency, p.sound_speed, p.spreading = cfg.frequency, cfg.sound_speed, cfg.spreading
    keep = None
    if cfg.directivity is not None:
        keep = np.ascontiguousarray(cfg.directivity, np.float32)
        p.directivity = keep.ctypes.data_as(C.POINTER(C.c_float))
        p.dir_height, p.dir_width = keep.shape
        p.dir_xscale, p.dir_yscale = cfg.dir_xscale, cfg.dir_yscale
    out = np.empty((hits.n, 4), np.float64)
    lib().orc_stream(hits.n, _p(hits.range), _p(hits.dir_x), _p(hits.dir_y), _p(hits.dir_z),
                     _p(hits.amplitude), C.byref(p), _p(out))
    del keep
    if cfg.kind == 1:
        out[:, 3] = 0.0
    return out


def directivity(table: np.ndarray, xs: float, ys: float, l) -> float:
    t = np.ascontiguousarray(table, np.float32)
    return float(lib().orc_directivity(_p(t), t.shape[1], t.shape[0], xs, ys,
                                       float(l[0]), float(l[1]), float(l[2])))


# ---------------------------------------------------------------------------
# PSF and convolution
# ---------------------------------------------------------------------------
def grid(cfg) -> Grid:
    g = Grid()
    g.range_lo, g.range_hi, g.range_count = cfg.range_min, cfg.range_max, cfg.range_count
    g.bearing_count = cfg.W
    g.axis = cfg.bearing_axis if cfg.kind == 2 else 0
    g.psf_range_taps, g.psf_bearing_taps = cfg.Tr, cfg.Tb
    g.psf_range_span, g.psf_bearing_span = cfg.psf_range_span, cfg.psf_bearing_span
    return g


def psf(taps: np.ndarray, span_r: float, span_b: float, dr: float, dx: float) -> complex:
    t = np.ascontiguousarray(taps.astype(np.complex64)).view(np.float32)
    re, im = C.c_double(), C.c_double()
    lib().orc_psf(_p(t), taps.shape[0], taps.shape[1], span_r, span_b, dr, dx,
                  C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def _taps(cfg) -> np.ndarray:
    return np.ascontiguousarray(cfg.psf.astype(np.complex64)).view(np.float32).ravel()


def conv(cfg, sorted_records: np.ndarray, win: np.ndarray, threads: int | None = None):
    """FP64 ping (complex128) of the plain form from sorted FP32 records."""
    srt = np.ascontiguousarray(sorted_records, np.float32).reshape(-1, 4)
    w = np.ascontiguousarray(win, np.uint32).reshape(-1)
    taps = _taps(cfg)
    g = grid(cfg)
    out = np.empty(2 * cfg.H * cfg.W, np.float64)
    if cfg.kind == 2:
        b = np.ascontiguousarray(cfg.bearings, np.float32)
        lib().orc_conv2d(_p(srt), _p(w), C.byref(g), _p(b), _p(taps), _p(out),
                         threads or nthreads())
        return out.view(np.complex128).reshape(cfg.H, cfg.W)
    lib().orc_conv1d(_p(srt), _p(w), C.byref(g), _p(taps), _p(out), threads or nthreads())
    return out.view(np.complex128)


def conv_rows(cfg, sorted_records: np.ndarray, win: np.ndarray, rows,
              threads: int | None = None) -> np.ndarray:
    """FP64 rows (len(rows) x W, complex128) of the 2D plain-form ping: the
    rows of `conv`, each sample visiting only the bearings in the PSF support."""
    assert cfg.kind == 2
    srt = np.ascontiguousarray(sorted_records, np.float32).reshape(-1, 4)
    w = np.ascontiguousarray(win, np.uint32).reshape(-1)
    r = np.ascontiguousarray(rows, np.uint32)
    assert r.size == 0 or int(r.max()) < cfg.H
    out = np.empty(2 * r.size * cfg.W, np.float64)
    b = np.ascontiguousarray(cfg.bearings, np.float32)
    taps = _taps(cfg)          # held: the call reads it through a raw pointer
    lib().orc_conv2d_rows(_p(srt), _p(w), C.byref(grid(cfg)), _p(b), _p(taps), _p(r), r.size,
                          _p(out), threads or nthreads())
    return out.view(np.complex128).reshape(r.size, cfg.W)


def grid_form(cfg, records: np.ndarray, win: np.ndarray, bearing0: float, pitch: float):
    """FP64 grid form (deposit + correlation + empty-window mask)."""
    rec = np.ascontiguousarray(records, np.float32).reshape(-1, 4)
    w = np.ascontiguousarray(win, np.uint32).reshape(-1)
    k = key_make(cfg.range_count, cfg.range_min, cfg.range_max, effective_margin(cfg))
    taps = _taps(cfg)
    g = grid(cfg)
    out = np.empty(2 * cfg.H * cfg.W, np.float64)
    lib().orc_grid2d(_p(rec), rec.shape[0], C.byref(k), _p(w), C.byref(g), bearing0, pitch,
                     _p(taps), _p(out))
    return out.view(np.complex128).reshape(cfg.H, cfg.W)


def reference_ping(cfg, records: np.ndarray, threads: int | None = None):
    """Oracle ping of unsorted FP32 records: Binner chain + FP64 gather."""
    b = binning(cfg, records)
    return conv(cfg, b["sorted"], b["windows"], threads), b


# ---------------------------------------------------------------------------
# CPU baseline (FP32 restatement of the reference pipeline)
# ---------------------------------------------------------------------------
def cpu_binning(cfg, records: np.ndarray, threads: int | None = None):
    rec = np.ascontiguousarray(records, np.float32).reshape(-1, 4)
    srt = np.empty_like(rec)
    win = np.empty(2 * cfg.H, np.uint32)
    g = grid(cfg)
    lib().cpu_binning(_p(rec), rec.shape[0], C.byref(g), effective_margin(cfg), _p(srt),
                      _p(win), threads or nthreads())
    return srt, win


def cpu_gather1d(cfg, srt, win, threads: int | None = None):
    out = np.zeros(2 * cfg.H, np.float32)
    g = grid(cfg)
    taps = _taps(cfg)          # held: the call reads these through raw pointers
    s = np.ascontiguousarray(srt, np.float32)
    w = np.ascontiguousarray(win, np.uint32)
    lib().cpu_gather1d_f32(_p(s), _p(w), C.byref(g), _p(taps), _p(out), threads or nthreads())
    return out.view(np.complex64)


def cpu_gather2d(cfg, srt, win, row_begin: int, row_end: int, threads: int | None = None):
    out = np.zeros(2 * cfg.H * cfg.W, np.float32)
    g = grid(cfg)
    b = np.ascontiguousarray(cfg.bearings, np.float32)
    taps = _taps(cfg)          # held: the call reads it through a raw pointer
    lib().cpu_gather2d_f32(_p(srt), _p(win), C.byref(g), _p(b), _p(taps), _p(out),
                           row_begin, row_end, threads or nthreads())
    return out.view(np.complex64).reshape(cfg.H, cfg.W)


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
