"""CPU restatement of the auralization stage of `EAR render` in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
never imports it.

Everything is computed in float64 from the same inputs, so the checks against
the product (single precision, device FFTs) use tolerances stated in the tests.
Each function cites the reference code it follows and the independent check
that pins it (tests/test_audio.py):

- lr4_coefficients / equalizer_split: lib/equalizer/Equalizer.cpp:27-95,
  exact (float64).  Pinned against scipy.signal.butter (a second-order
  Butterworth by the bilinear transform with prewarping) squared, and by the
  crossover identities (LP + HP is all-pass, both -6 dB at fc).
- lr4_coefficients_f32 / equalizer_split_f32: the single-precision
  restatement in oracle/ear_oracle.c (the reference's float operation order).
  Pinned by agreement with the exact forms (coefficients to a few ulp, band
  signals to the float recursion's error).
- wav_quantize: lib/wave/WaveFile.cpp:191-256.  Pinned by Python's `wave`
  module reading the product's files and by hand-computed samples.
- convolve / process_track: src/Recorder.cpp:145-243 (FFT form), the fades
  of :172-181, and the FloatBuffer growth rules of src/Recorder.cpp:33-59.
  The convolution is pinned against np.convolve's direct sum; the buffer
  rules against hand-worked lengths.
- Track.add / listener_output: src/Recorder.cpp:294-301,
  :433-461, src/EAR.cpp:358-386.
"""
from __future__ import annotations

import numpy as np

FS = 44100
INITIAL = 3 * FS  # INITIAL_BUFFER_SIZE, src/Recorder.h:36
STEP = FS         # INCREMENTAL_BUFFER_SIZE, src/Recorder.h:37
PI_REF = 3.14285714285714  # the reference's pi (22/7)


# ---------------------------------------------------------------- equalizer
def lr4_coefficients(fc: float, high_pass: bool):
    """(numerator b, denominator a) of the fourth-order Linkwitz-Riley section
    at fc Hz, a[0] = 1.

    Analog prototype (w^2 / (s^2 + sqrt2 w s + w^2))^2 (low-pass) or
    (s^2 / (...))^2 (high-pass), w = 2 pi fs, bilinear s = k (1-z)/(1+z) with
    k = w / tan(pi fc / fs), pi = 22/7 as the reference."""
    w = 2.0 * PI_REF * FS
    k = w / np.tan(PI_REF * fc / FS)
    r2 = np.sqrt(2.0)
    q = np.array([k * k + r2 * w * k + w * w, 2.0 * (w * w - k * k), k * k - r2 * w * k + w * w])
    den = np.convolve(q, q)
    num = (k ** 4 * np.array([1.0, -4.0, 6.0, -4.0, 1.0])) if high_pass else (w ** 4 * np.array([1.0, 4, 6, 4, 1]))
    return num / den[0], den / den[0]


def lr4_filter(x, fc: float, high_pass: bool):
    from scipy.signal import lfilter

    b, a = lr4_coefficients(fc, high_pass)
    return lfilter(b, a, np.asarray(x, np.float64))


def lr4_coefficients_f32(fc: float, high_pass: bool):
    """The single-precision coefficients of the oracle's C restatement
    (Equalizer::Pass in the reference's order of float operations):
    (numerator, denominator)."""
    import ctypes as C

    from . import oracle as O

    L = O.lib()
    a = (C.c_float * 5)()
    b = (C.c_float * 5)()
    L.or_lr4_coefficients(C.c_float(fc), int(high_pass), a, b)
    return np.array(a[:], np.float32), np.array(b[:], np.float32)


def equalizer_split_f32(x, f1=200.0, f2=1000.0, f3=2000.0):
    """Equalizer::Split in single precision (oracle C restatement)."""
    import ctypes as C

    from . import oracle as O

    L = O.lib()
    x = np.ascontiguousarray(x, np.float32)
    lo, mi, hi = (np.empty_like(x) for _ in range(3))
    fp = C.POINTER(C.c_float)
    L.or_equalizer_split(x.ctypes.data_as(fp), C.c_uint32(x.size), lo.ctypes.data_as(fp), mi.ctypes.data_as(fp),
                         hi.ctypes.data_as(fp), C.c_float(f1), C.c_float(f2), C.c_float(f3))
    return lo, mi, hi


def equalizer_split(x, f1=200.0, f2=1000.0, f3=2000.0):
    """Equalizer::Split (lib/equalizer/Equalizer.cpp:81-95), exact (float64)
    filters."""
    fc1 = (np.float32(f1) + np.float32(f2)) / np.float32(2.0)
    fc2 = (np.float32(f2) + np.float32(f3)) / np.float32(2.0)
    low = lr4_filter(x, float(fc1), False)
    mid = lr4_filter(lr4_filter(x, float(fc2), False), float(fc1), True)
    high = lr4_filter(x, float(fc2), True)
    return low, mid, high


# ---------------------------------------------------------------- WAV samples
def wav_quantize(x, norm=False, norm_max=-1.0):
    """Mono 16-bit samples of WaveFile::FromFloat: trunc(x / max * 32768),
    max = 1, or max|x| / 0.8 with norm, or norm_max / 0.95 when norm_max >= 0
    (out-of-range values saturate)."""
    x = np.asarray(x, np.float64)
    if norm:
        mx = (np.abs(x).max() / 0.8) if norm_max < 0 else norm_max / 0.95
    else:
        mx = 1.0
    return np.clip(np.trunc(x / mx * 32768.0), -32768, 32767).astype(np.int16)


# ---------------------------------------------------------------- convolution
def convolve(ir, sig):
    """Linear convolution, float64 (the M + N + 1 outputs of the reference's
    FFT form: the last two are zero)."""
    ir = np.asarray(ir, np.float64)
    sig = np.asarray(sig, np.float64)
    n = len(ir) + len(sig) + 1
    out = np.zeros(n)
    if len(ir) and len(sig):
        L = 1 << int(np.ceil(np.log2(n)))
        c = np.fft.irfft(np.fft.rfft(ir, L) * np.fft.rfft(sig, L), L)
        out[: len(ir) + len(sig) - 1] = c[: len(ir) + len(sig) - 1]
    return out


def fades(n: int):
    """The running float factors of src/Recorder.cpp:172-181: fade-out starts
    at 1 and fade-in at 0, each stepping by 1/n in single precision."""
    fo = np.empty(n, np.float32)
    fi = np.empty(n, np.float32)
    df = np.float32(1.0) / np.float32(max(n, 1))
    a, b = np.float32(1.0), np.float32(0.0)
    for i in range(n):
        fo[i], fi[i] = a, b
        a = np.float32(a - df)
        b = np.float32(b + df)
    return fo, fi


def fresh_buffer_length(first: int, last: int) -> int:
    """Length of a new FloatBuffer after writes at first..last in order
    (src/Recorder.cpp:43-59: starts at INITIAL, and a write at index i >=
    length grows it to i + STEP)."""
    L = INITIAL
    if first >= L:
        L = first + STEP
    while last >= L:
        L += STEP
    return L


def grow_to(L: int, last: int) -> int:
    """Buffer length after sequential writes 0..last into a buffer of length L."""
    while last >= L:
        L += STEP
    return L


class Track:
    """A FloatBuffer: data (float64) over the allocated length, real_length =
    largest index written, first_sample."""

    def __init__(self):
        self.data = np.zeros(INITIAL)
        self.real_length = 0
        self.first_sample = INITIAL - 1

    def write_span(self, start: int, values):
        last = start + len(values) - 1
        if len(values) == 0:
            return
        L = len(self.data)
        if start >= L:
            L = start + STEP
        L = grow_to(L, last)
        if L > len(self.data):
            self.data = np.concatenate([self.data, np.zeros(L - len(self.data))])
        self.data[start:last + 1] = values
        self.real_length = max(self.real_length, last)
        self.first_sample = min(self.first_sample, start)

    def add(self, other: "Track"):
        """FloatBuffer add over other's whole buffer (getLength(0) = its length)."""
        n = len(other.data)
        L = grow_to(len(self.data), n - 1)
        if L > len(self.data):
            self.data = np.concatenate([self.data, np.zeros(L - len(self.data))])
        self.data[:n] += other.data
        self.real_length = max(self.real_length, n - 1)
        self.first_sample = 0


def process_track(ir, sig, offset: int, ir2=None) -> Track:
    """RecorderTrack::Process of the first len(ir) response samples with a
    section starting `offset` samples into the output (src/Recorder.cpp:145-243);
    with ir2 the cross-fade: ir with the faded-out section plus ir2 with the
    faded-in one, the second result a buffer of its own added over its length."""
    sig = np.asarray(sig, np.float32)
    t = Track()
    if ir2 is None:
        t.write_span(offset, convolve(ir, sig))
        return t
    fo, fi = fades(len(sig))
    a = convolve(ir, (sig * fo).astype(np.float32))
    b = convolve(ir2, (sig * fi).astype(np.float32))
    t.write_span(offset, a)
    u = Track()
    u.write_span(offset, b)
    t.add(u)
    return t


def section(sig, start_s: float, length_s: float = -1.0):
    """SoundFile::Section (src/SoundFile.cpp:164-173) of a band signal whose
    offset is 0: (samples, offset)."""
    sig = np.asarray(sig)
    s = int(np.float32(start_s) * np.float32(FS))
    n = len(sig)
    ln = n - s if length_s < 0 else int(np.float32(length_s) * np.float32(FS))
    if s >= n:
        return np.zeros(0, np.float32), 0
    return sig[s:s + min(ln, n - s)], s


def listener_output(tracks):
    """Sum of the processed tracks of every context (Recorder::Add), then
    Normalize(0.8) over [0, real_length) and the samples a Save() writes:
    [0, real_length) of the total as 16-bit (norm off)."""
    total = Track()
    for t in tracks:
        total.add(t)
    n = total.real_length
    x = total.data[:n]
    mx = np.abs(x).max() if n else 0.0
    if mx > 0:
        x = x * (0.8 / mx)
    return wav_quantize(x)
