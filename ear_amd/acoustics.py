"""Room-acoustic figures computed from rendered histograms and meshes.

* ``edc`` / ``t60_iso``: Schroeder backward integration of the unsigned energy
  histogram and T60 = 60 dB / slope of the -5..-35 dB range (T30, ISO 3382-1).
* ``t60_reference``: RecorderTrack::T60 (src/Recorder.cpp:303-340) on a track,
  the figure `EAR calc T60` prints (src/EAR.cpp:259-262).
* ``sabine`` / ``eyring``: the statistical predictions printed by
  `EAR calc T60` (src/EAR.cpp:264-285).
"""
from __future__ import annotations

import math

import numpy as np

FS = 44100.0


def edc(energy: np.ndarray) -> np.ndarray:
    """Energy decay curve EDC(t) = sum_{tau >= t} E(tau) (float64)."""
    e = np.asarray(energy, np.float64)
    return np.cumsum(e[::-1])[::-1]


def edc_db(energy: np.ndarray) -> np.ndarray:
    c = edc(energy)
    with np.errstate(divide="ignore"):
        return 10.0 * np.log10(c / c[0])


def t60_iso(energy: np.ndarray, lo_db: float = -5.0, hi_db: float = -35.0) -> float:
    """T60 from the linear fit of the EDC (dB) between lo_db and hi_db."""
    L = edc_db(energy)
    idx = np.nonzero((L <= lo_db) & (L >= hi_db))[0]
    if len(idx) < 2:
        raise ValueError("decay range not covered by the histogram")
    t = idx / FS
    m, _ = np.polyfit(t, L[idx], 1)
    return float(-60.0 / m)


def edc_error(e_a: np.ndarray, e_b: np.ndarray, floor_db: float = -60.0) -> float:
    """max |EDC_a / EDC_b - 1| over t up to where EDC_b falls below floor_db."""
    ca, cb = edc(e_a), edc(e_b)
    Lb = 10.0 * np.log10(np.maximum(cb, 1e-300) / cb[0])
    keep = Lb >= floor_db
    return float(np.max(np.abs(ca[keep] / cb[keep] - 1.0)))


def power(track: np.ndarray, a: float = 0.335) -> np.ndarray:
    """FloatBuffer::Power (src/Recorder.cpp:91-96): sign-preserving |x|^a."""
    x = np.asarray(track, np.float32)
    f = np.power(np.abs(x), np.float32(a)).astype(np.float32)
    return np.where(x < 0, -f, f).astype(np.float32)


def t60_reference(track: np.ndarray, first: int | None = None, last: int | None = None) -> float:
    """RecorderTrack::T60 (src/Recorder.cpp:303-340): from the first sample that
    is lower than its predecessor (end of the 'direct lobe'), the last sample
    above direct/1000 (60 dB in amplitude), in seconds.  `first`/`last` are the
    track's first_sample / real_length (default: first / last non-zero bin)."""
    x = np.asarray(track, np.float32)
    nz = np.nonzero(x)[0]
    if first is None:
        first = int(nz[0]) if len(nz) else 0
    if last is None:
        last = int(nz[-1]) if len(nz) else 0
    gain = np.float32(10.0) ** np.float32(60.0 / 20.0)
    prev = np.float32(-1.0)
    inside = False
    min_gain = np.float32(0.0)
    last_sig = None
    direct_off = None
    for j in range(first, last):
        v = x[j]
        if inside:
            if v > min_gain:
                last_sig = j
        elif v < prev:
            inside = True
            min_gain = prev / gain
            direct_off = j
        prev = v
    if last_sig is None or direct_off is None:
        return float("nan")  # the reference reads uninitialised offsets here
    return float(last_sig - direct_off) / 44100.0


def sabine(V: float, A: float, m: float) -> float:
    """0.1611 V / (A + 4 m V), src/EAR.cpp:281-282."""
    return 0.1611 * V / (A + 4.0 * m * V)


def eyring(V: float, S: float, a: float, m: float) -> float:
    """0.1611 V / (-S ln(1 - a) + 4 m V), src/EAR.cpp:283-284."""
    return 0.1611 * V / (-S * math.log(1.0 - a) + 4.0 * m * V)
