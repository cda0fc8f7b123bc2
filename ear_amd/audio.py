"""Python binding of include/ear_audio.h (libear_host.so): WAV files and the
three-band equalizer of the auralization stage.  Host code; no device.

    x = read_wav("dry.wav")                     # float32, channels averaged
    lo, mid, hi = equalizer_split(x)            # FREQ defaults 0.2 / 1 / 2 kHz
    write_wav("out.wav", y)                     # mono 16-bit, 44.1 kHz
    write_wav("out.wav", left, right, norm=True)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
HOST_LIB_PATH = os.path.join(PKG, "libear_host.so")
EXPORTS = ["ear_wave_read", "ear_wave_write", "ear_equalizer_split"]

_lib = None
_fp = C.POINTER(C.c_float)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(HOST_LIB_PATH):
        raise ImportError(f"{HOST_LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(HOST_LIB_PATH)
    L.ear_wave_read.argtypes = [C.c_char_p, _fp, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_int),
                                C.POINTER(C.c_int)]
    L.ear_wave_read.restype = C.c_int
    L.ear_wave_write.argtypes = [C.c_char_p, _fp, C.c_uint64, _fp, C.c_uint64, C.c_int, C.c_float]
    L.ear_wave_write.restype = C.c_int
    L.ear_equalizer_split.argtypes = [_fp, C.c_uint64, _fp, _fp, _fp, C.c_float, C.c_float, C.c_float]
    L.ear_equalizer_split.restype = C.c_int
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


def read_wav(path: str, with_format: bool = False):
    """float32 samples (channels averaged, scaled by 2^(bits-1)); with
    with_format also (channels, bits).  Raises OSError if the file cannot be
    read as 8/16/24-bit integer PCM."""
    L = load()
    n, ch, bits = C.c_uint64(0), C.c_int(0), C.c_int(0)
    p = os.fsencode(path)
    rc = L.ear_wave_read(p, None, 0, C.byref(n), C.byref(ch), C.byref(bits))
    if rc != 0:
        raise OSError(f"cannot read {path} as integer PCM WAV (code {rc})")
    out = np.zeros(n.value, np.float32)
    rc = L.ear_wave_read(p, _ptr(out), n.value, C.byref(n), None, None)
    if rc != 0:
        raise OSError(f"cannot read {path} (code {rc})")
    return (out, (ch.value, bits.value)) if with_format else out


def write_wav(path: str, left, right=None, norm: bool = False, norm_max: float = -1.0) -> None:
    """16-bit PCM at 44.1 kHz; mono unless `right` is given (WaveFile::FromFloat
    normalisation rules, see include/ear_audio.h)."""
    L = load()
    l = np.ascontiguousarray(left, np.float32)
    r = None if right is None else np.ascontiguousarray(right, np.float32)
    rc = L.ear_wave_write(os.fsencode(path), _ptr(l), l.size, None if r is None else _ptr(r),
                          0 if r is None else r.size, int(bool(norm)), float(norm_max))
    if rc != 0:
        raise OSError(f"cannot write {path} (code {rc})")


def equalizer_split(x, f1: float = 200.0, f2: float = 1000.0, f3: float = 2000.0):
    """(low, mid, high) float32 bands of x at crossovers (f1+f2)/2 and
    (f2+f3)/2 Hz (fourth-order Linkwitz-Riley, 44.1 kHz)."""
    L = load()
    x = np.ascontiguousarray(x, np.float32)
    lo, mi, hi = (np.empty_like(x) for _ in range(3))
    rc = L.ear_equalizer_split(_ptr(x), x.size, _ptr(lo), _ptr(mi), _ptr(hi), f1, f2, f3)
    if rc != 0:
        raise ValueError(f"ear_equalizer_split failed (code {rc})")
    return lo, mi, hi
