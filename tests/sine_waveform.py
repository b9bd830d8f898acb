"""Test-side restatement of the sine-waveform PSF a sensor rebuilds when its
ranges change (Waveform.h:115-134, SensorInstance.cu:23-34,
SensorInstance2D.cu:20-42), written with numpy FP32 arithmetic independently
of the library's C++ so that a ping computed after set_ranges can be checked
against the oracle with taps the library did not produce."""
from __future__ import annotations

import copy

import numpy as np

F32 = np.float32


def linspace_resolution(lo: float, hi: float, n: int) -> np.float32:
    """Linspace<float>::resolution(): (hi - lo) / (n - 1) in FP32."""
    return F32(F32(F32(hi) - F32(lo)) / F32(n - 1))


def sine_taps(frequency: float, duration: float, oversampling: int) -> np.ndarray:
    """Waveform_Sine samples: size = (unsigned)(2 f os * tau) with the rate in
    FP32, sample n = exp(i n 2 pi f / rate) (phase in double, then FP32)."""
    rate = F32(F32(F32(2.0) * F32(frequency)) * F32(oversampling))
    size = int(F32(rate * F32(duration)))
    ph = (np.arange(size, dtype=np.float64) * (2.0 * np.pi * float(F32(frequency)) / float(rate)))
    ph = ph.astype(np.float32)
    return (np.cos(ph) + 1j * np.sin(ph)).astype(np.complex64)


def after_set_ranges(cfg, lo: float, hi: float, n: int, margin: float = -1.0):
    """The sensor description SensorInstance::set_ranges leaves: an adaptive
    duration becomes resolution / c, the PSF table wave[h] * beam[w], its range
    span c tau and (margin < 0) the Binner margin c tau."""
    out = copy.copy(cfg)
    out.range_min, out.range_max, out.range_count = lo, hi, n
    out.margin = margin
    wf = dict(cfg.waveform)
    c = F32(cfg.sound_speed)
    tau = F32(wf["duration"]) if wf["fixed"] else F32(linspace_resolution(lo, hi, n) / c)
    wf["duration"] = float(tau)
    out.waveform = wf
    wave = sine_taps(wf["frequency"], tau, wf["oversampling"])
    beam = np.ones(1, np.float32) if wf.get("beam") is None else np.asarray(wf["beam"], np.float32)
    out.psf = (wave[:, None] * beam[None, :]).astype(np.complex64)
    out.psf_range_span = float(F32(tau * c))
    out._keep = []
    return out
