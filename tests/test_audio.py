"""The audio side of auralization on the CPU: WAV files, the three-band
equalizer (include/ear_audio.h, libear_host.so) and the oracle's
restatement of the convolution stage's buffer rules.

Pins: the exact equalizer against scipy's Butterworth design and the
Linkwitz-Riley identities; the WAV writer against Python's `wave` module and
hand-computed samples; the oracle's FloatBuffer rules against hand-worked
lengths; its convolution against the direct sum."""
import os
import wave

import numpy as np
import pytest

from ear_amd import audio
from oracle import audio as OA

FS = 44100


# ---------------------------------------------------------------- equalizer
@pytest.mark.parametrize("fc", [100.0, 600.0, 1500.0, 4000.0])
@pytest.mark.parametrize("hp", [False, True])
def test_lr4_exact_is_squared_butterworth(fc, hp):
    from scipy.signal import butter

    b, a = OA.lr4_coefficients(fc, hp)
    # the reference's pi = 22/7 shifts the prewarped cutoff by pi'/pi
    fc_eff = fc * OA.PI_REF / np.pi
    bb, aa = butter(2, fc_eff / (FS / 2), "high" if hp else "low")
    np.testing.assert_allclose(b, np.convolve(bb, bb), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(a, np.convolve(aa, aa), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("fc", [600.0, 1500.0])
def test_lr4_crossover_identities(fc):
    from scipy.signal import freqz

    bl, al = OA.lr4_coefficients(fc, False)
    bh, ah = OA.lr4_coefficients(fc, True)
    w = np.linspace(1e-3, np.pi - 1e-3, 2000)
    _, hl = freqz(bl, al, worN=w)
    _, hh = freqz(bh, ah, worN=w)
    np.testing.assert_allclose(np.abs(hl + hh), 1.0, atol=1e-9)  # all-pass sum
    wc = 2 * np.pi * fc * OA.PI_REF / np.pi / FS
    _, h = freqz(bl, al, worN=[wc])
    _, g = freqz(bh, ah, worN=[wc])
    assert abs(abs(h[0]) - 0.5) < 1e-9 and abs(abs(g[0]) - 0.5) < 1e-9  # -6 dB at the crossover
    assert abs(np.sum(bl) / np.sum(al) - 1.0) < 1e-9  # DC gain of the low-pass
    sgn = (-1.0) ** np.arange(5)
    assert abs(np.dot(bh, sgn) / np.dot(ah, sgn) - 1.0) < 1e-9  # Nyquist gain of the high-pass


@pytest.mark.parametrize("fc", [100.0, 600.0, 1500.0, 4000.0])
@pytest.mark.parametrize("hp", [False, True])
def test_lr4_f32_coefficients_match_exact(fc, hp):
    # (numerator, denominator) in both forms
    b32, a32 = OA.lr4_coefficients_f32(fc, hp)
    b, a = OA.lr4_coefficients(fc, hp)
    assert a32[0] == 1.0
    # a few ulp of the largest coefficient: the float form rounds each step
    np.testing.assert_allclose(a32, a, rtol=0, atol=4e-6 * np.abs(a).max())
    np.testing.assert_allclose(b32, b, rtol=0, atol=4e-6 * np.abs(b).max())


def _signals():
    rng = np.random.default_rng(7)
    t = np.arange(FS) / FS
    yield "noise", (0.3 * rng.standard_normal(FS)).astype(np.float32)
    yield "tones", (0.3 * np.sin(2 * np.pi * 150 * t) + 0.2 * np.sin(2 * np.pi * 1800 * t)
                    + 0.1 * np.sin(2 * np.pi * 5000 * t)).astype(np.float32)
    imp = np.zeros(4096, np.float32)
    imp[0] = 1.0
    yield "impulse", imp
    yield "ragged", rng.standard_normal(5).astype(np.float32)
    yield "single", np.array([0.25], np.float32)


@pytest.mark.parametrize("name,x", list(_signals()))
def test_equalizer_bit_exact_with_oracle(name, x):
    got = audio.equalizer_split(x)
    want = OA.equalizer_split_f32(x)
    for g, w in zip(got, want):
        assert np.array_equal(g, w), name


@pytest.mark.parametrize("freq", [(200.0, 1000.0, 2000.0), (300.0, 2000.0, 6000.0)])
def test_equalizer_f32_close_to_exact(freq):
    x = next(_signals())[1]
    got = OA.equalizer_split_f32(x, *freq)
    want = OA.equalizer_split(x, *freq)
    # the float recursion with float coefficients: measured 0.3 % of the band's
    # peak for the low band at these crossovers (double poles amplify the
    # coefficient rounding); a wrong sign or term is O(1)
    for g, w in zip(got, want):
        assert np.abs(g - w).max() <= 0.01 * np.abs(w).max()


def test_equalizer_empty_input():
    lo, mi, hi = audio.equalizer_split(np.zeros(0, np.float32))
    assert lo.size == mi.size == hi.size == 0


def test_equalizer_impulse_bands_sum_near_allpass():
    x = np.zeros(1 << 15, np.float32)
    x[0] = 1.0
    lo, mi, hi = audio.equalizer_split(x)
    s = lo.astype(np.float64) + mi + hi
    # |LP1 + HP1 LP2 + HP2| is not exactly all-pass (two crossovers), but the
    # energy stays within a few percent of the impulse's
    e = np.sum(np.abs(np.fft.rfft(s)) ** 2) / (len(s) / 2)
    assert 0.8 < e < 1.2


# ---------------------------------------------------------------- WAV files
def _read_wave(path):
    with wave.open(str(path), "rb") as w:
        n, ch, sw, fr = w.getnframes(), w.getnchannels(), w.getsampwidth(), w.getframerate()
        raw = w.readframes(n)
    return np.frombuffer(raw, "<i2").reshape(-1, ch), ch, sw, fr


def test_write_mono_matches_hand_computed(tmp_path):
    p = tmp_path / "m.wav"
    x = np.array([0.0, 0.5, -0.5, 0.25, -0.999, 1.0, -1.0, 2.0, 1e-6], np.float32)
    audio.write_wav(str(p), x)
    s, ch, sw, fr = _read_wave(p)
    assert (ch, sw, fr) == (1, 2, 44100)
    # trunc(x * 32768); 1.0 and 2.0 exceed the 16-bit range and saturate
    assert s[:, 0].tolist() == [0, 16384, -16384, 8192, -32735, 32767, -32768, 32767, 0]
    raw = p.read_bytes()
    assert raw[:4] == b"RIFF" and int.from_bytes(raw[4:8], "little") == 2 * len(x) + 36 and len(raw) == 44 + 2 * len(x)


@pytest.mark.parametrize("norm_max", [-1.0, 0.5])
def test_write_mono_normalised(tmp_path, norm_max):
    p = tmp_path / "n.wav"
    x = np.array([0.1, -0.4, 0.2, 0.3], np.float32)
    audio.write_wav(str(p), x, norm=True, norm_max=norm_max)
    s, *_ = _read_wave(p)
    # the reference's float steps: max / 0.8 (or norm_max / 0.95), then x / max * 32768
    mx = np.float32(0.4) / np.float32(0.8) if norm_max < 0 else np.float32(norm_max) / np.float32(0.95)
    want = [int(np.trunc(np.float32(np.float32(v / mx) * np.float32(32768.0)))) for v in x]
    assert s[:, 0].tolist() == want
    assert s[1, 0] == -26214 if norm_max < 0 else True


def test_write_stereo_pads_and_normalises_by_signed_max(tmp_path):
    p = tmp_path / "s.wav"
    left = np.array([0.1, -0.9, 0.2], np.float32)
    right = np.array([0.4], np.float32)
    audio.write_wav(str(p), left, right, norm=True)
    s, ch, sw, fr = _read_wave(p)
    assert (ch, sw, fr) == (2, 2, 44100) and s.shape == (3, 2)
    mx = np.float32(0.4) / np.float32(0.8)  # largest signed sample, not |-0.9|
    f = lambda v: int(np.clip(np.trunc(np.float32(np.float32(v / mx) * np.float32(32768.0))), -32768, 32767))
    assert s[:, 0].tolist() == [f(0.1), f(-0.9), f(0.2)]
    assert s[:, 1].tolist() == [f(0.4), 0, 0]


@pytest.mark.parametrize("bits", [8, 16, 24])
@pytest.mark.parametrize("ch", [1, 2, 3])
def test_read_pcm(tmp_path, bits, ch):
    rng = np.random.default_rng(bits * 10 + ch)
    n = 1001
    full = 1 << (bits - 1)
    v = rng.integers(-full, full, size=(n, ch))
    if bits == 8:
        raw = (v + 128).astype(np.uint8).tobytes()
    elif bits == 16:
        raw = v.astype("<i2").tobytes()
    else:
        u = (v & 0xFFFFFF).astype("<u4")
        raw = np.stack([(u >> 0) & 255, (u >> 8) & 255, (u >> 16) & 255], -1).astype(np.uint8).tobytes()
    p = tmp_path / "r.wav"
    with wave.open(str(p), "wb") as w:
        w.setnchannels(ch)
        w.setsampwidth(bits // 8)
        w.setframerate(22050)
        w.writeframes(raw)
    x, (c, b) = audio.read_wav(str(p), with_format=True)
    assert (c, b) == (ch, bits) and x.shape == (n,)
    # float steps of WaveFile::ToFloat: sum of v / 2^(bits-1) per channel, / channels
    want = np.zeros(n, np.float32)
    for k in range(ch):
        want = (want + (v[:, k].astype(np.float32) / np.float32(full))).astype(np.float32)
    want = (want / np.float32(ch)).astype(np.float32)
    assert np.array_equal(x, want)


def _riff(chunks):
    body = b"WAVE" + b"".join(cid + len(data).to_bytes(4, "little") + data + (b"\0" if len(data) % 2 else b"")
                              for cid, data in chunks)
    return b"RIFF" + len(body).to_bytes(4, "little") + body


def _fmt(tag=1, ch=1, rate=44100, bits=16, extra=b""):
    ba = ch * bits // 8
    return (tag.to_bytes(2, "little") + ch.to_bytes(2, "little") + rate.to_bytes(4, "little")
            + (rate * ba).to_bytes(4, "little") + ba.to_bytes(2, "little") + bits.to_bytes(2, "little") + extra)


def test_read_skips_other_chunks_and_joins_data_chunks(tmp_path):
    a = np.array([100, -200, 300], "<i2").tobytes()
    b = np.array([-400], "<i2").tobytes()
    p = tmp_path / "c.wav"
    p.write_bytes(_riff([(b"fmt ", _fmt(extra=b"\0\0")), (b"LIST", b"odd"), (b"data", a), (b"junk", b"xy"),
                         (b"data", b)]))
    x = audio.read_wav(str(p))
    assert np.array_equal(x, np.array([100, -200, 300, -400], np.float32) / np.float32(32768))


@pytest.mark.parametrize("case", ["missing", "float", "no_fmt", "not_riff", "empty_data", "bits12"])
def test_read_rejects(tmp_path, case):
    p = tmp_path / "bad.wav"
    data = np.array([1, 2], "<i2").tobytes()
    if case == "float":
        p.write_bytes(_riff([(b"fmt ", _fmt(tag=3, bits=32)), (b"data", b"\0" * 8)]))
    elif case == "no_fmt":
        p.write_bytes(_riff([(b"data", data)]))
    elif case == "not_riff":
        p.write_bytes(b"RIFX" + bytes(40))
    elif case == "empty_data":
        p.write_bytes(_riff([(b"fmt ", _fmt()), (b"data", b"")]))
    elif case == "bits12":
        p.write_bytes(_riff([(b"fmt ", _fmt(bits=12)), (b"data", data)]))
    with pytest.raises(OSError):
        audio.read_wav(str(p))


def test_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    x = (rng.uniform(-0.99, 0.99, 5000)).astype(np.float32)
    p = tmp_path / "t.wav"
    audio.write_wav(str(p), x)
    y = audio.read_wav(str(p))
    q = np.trunc(x.astype(np.float64) * 32768.0)
    assert np.array_equal(y, (q / 32768.0).astype(np.float32))


# ---------------------------------------------------------------- oracle buffer rules
def test_fresh_buffer_lengths():
    I, S = 3 * FS, FS
    assert OA.fresh_buffer_length(0, 10) == I
    assert OA.fresh_buffer_length(0, I - 1) == I
    assert OA.fresh_buffer_length(0, I) == I + S
    assert OA.fresh_buffer_length(5, I + S) == I + 2 * S
    assert OA.fresh_buffer_length(I + 7, I + 9) == I + 7 + S
    assert OA.fresh_buffer_length(I + 7, I + 7 + S) == I + 7 + 2 * S


def test_track_lengths_of_a_processed_section():
    ir = np.zeros(100)
    ir[3] = 1.0
    sig = np.ones(50, np.float32)
    t = OA.process_track(ir, sig, 1000)
    # M + N + 1 outputs written from the offset; real_length is the last index
    assert t.real_length == 1000 + 100 + 50
    assert len(t.data) == 3 * FS and t.first_sample == 1000
    np.testing.assert_allclose(t.data[1003:1053], 1.0, atol=1e-12)
    assert np.abs(t.data[:1003]).max() < 1e-12 and np.abs(t.data[1053:]).max() < 1e-12
    total = OA.Track()
    total.add(t)
    assert total.real_length == 3 * FS - 1  # the whole buffer of the added track


def test_convolve_matches_direct_sum():
    rng = np.random.default_rng(4)
    a = rng.standard_normal(300)
    b = rng.standard_normal(77)
    c = OA.convolve(a, b)
    assert len(c) == 300 + 77 + 1
    np.testing.assert_allclose(c[:376], np.convolve(a, b), atol=1e-12)
    assert c[376] == 0 and c[377] == 0


def test_fades_are_running_float_sums():
    fo, fi = OA.fades(3)
    d = np.float32(1) / np.float32(3)
    assert fo.tolist() == [1.0, np.float32(1 - d), np.float32(np.float32(1 - d) - d)]
    assert fi.tolist() == [0.0, d, np.float32(d + d)]


def test_section_rules():
    sig = np.arange(44100 * 2, dtype=np.float32)
    s, off = OA.section(sig, 0.5, 0.25)
    assert off == 22050 and len(s) == 11025 and s[0] == 22050
    s, off = OA.section(sig, 1.5)
    assert off == 66150 and len(s) == 88200 - 66150
    s, off = OA.section(sig, 3.0)
    assert off == 0 and len(s) == 0


def test_host_library_exports_audio_abi():
    import ctypes as C
    import re

    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "ear_audio.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    declared = sorted(set(re.findall(r"\b(ear_[a-z_0-9]+)\s*\(", src)))
    assert declared == sorted(audio.EXPORTS)
    lib = C.CDLL(audio.HOST_LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
