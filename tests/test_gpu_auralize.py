"""`EAR render` end to end on the device: the listener's WAV file against the
oracle's restatement of the auralization stage (src/EAR.cpp:293-386): the
equalizer bands of the dry signal, each band's impulse response convolved
with its band (keyframes cross-faded), the reference's buffer lengths, the
sum over contexts, Normalize(0.8) and 16-bit samples.

The responses are the ones the CLI itself dumps under SET debugdir (the
FloatBuffer::Write format); how the tracer makes them is checked against the
oracle in the parity tests.  Tolerance: one 16-bit step (float FFTs on the
device against float64 ones here; the equalizer is bit-exact)."""
import os
import subprocess
import wave

import numpy as np
import pytest

from ear_amd import earfile, scenes
from oracle import audio as OA

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EAR = os.path.join(ROOT, "ear_amd", "bin", "EAR")
NAMES = ("low", "mid", "high")


def write_dry(path, n, seed=5):
    """A 16-bit mono WAV of noise bursts and tones; returns the samples as
    the reader scales them (v / 32768)."""
    rng = np.random.default_rng(seed)
    t = np.arange(n) / 44100.0
    x = 0.25 * np.sin(2 * np.pi * 220 * t) + 0.15 * np.sin(2 * np.pi * 3100 * t)
    x += 0.2 * rng.standard_normal(n) * (np.sin(2 * np.pi * 3 * t) > 0)
    v = np.clip(np.round(x * 32767), -32768, 32767).astype("<i2")
    with wave.open(str(path), "wb") as w:
        w.setnchannels(1)
        w.setsampwidth(2)
        w.setframerate(44100)
        w.writeframes(v.tobytes())
    return v.astype(np.float32) / np.float32(32768.0)


def read_wave(path):
    with wave.open(str(path), "rb") as w:
        ch = w.getnchannels()
        s = np.frombuffer(w.readframes(w.getnframes()), "<i2").reshape(-1, ch)
    return s.astype(np.int64)


def response(debug, rec, sound, band, frame=None):
    """The impulse response a context was convolved with: the first
    real_length samples of the dumped track (the file holds real_length + 1)."""
    name = f"response-{rec}.sound-{sound}" + (f".frame-{frame:02d}" if frame is not None else "")
    d = np.fromfile(os.path.join(debug, name + f".band-{band}{NAMES[band]}.bin"), "<f4")
    return d[:-1].astype(np.float64)


def run(tmp_path, ear_path):
    print(f"EAR render {ear_path}", flush=True)
    r = subprocess.run([EAR, "render", str(ear_path)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def check_close(got, want):
    assert got.shape == want.shape, (got.shape, want.shape)
    d = np.abs(got - want.astype(np.int64))
    assert d.max() <= 1, (d.max(), int(np.argmax(d)))


def test_render_static_mono(tmp_path):
    dry = write_dry(tmp_path / "dry.wav", 15000)
    debug = tmp_path / "debug"
    debug.mkdir()
    s = scenes.shoebox(spec=0.2)
    p = tmp_path / "box.EAR"
    earfile.scene_to_ear(s, p, samples=20_000, seed=77, maxbounces=120, wav=str(tmp_path / "dry.wav"),
                         debugdir=str(debug))
    out = run(tmp_path, p)
    assert "Processing data..." in out and "Merging result..." in out
    bands = OA.equalizer_split_f32(dry)
    for b in range(3):
        # equalizer dumps: bit-exact bands, written without normalisation
        got = read_wave(debug / f"sound-0.band-{b}{NAMES[b]}.wav")[:, 0]
        assert np.array_equal(got, OA.wav_quantize(bands[b]).astype(np.int64))
    tracks = [OA.process_track(response(debug, 0, 0, b), bands[b], 0) for b in range(3)]
    for b in range(3):
        # per-context results (Save with the base class's defaults: normalised)
        t = tracks[b]
        got = read_wave(debug / f"rec-0.sound-0.band-{b}.wav")[:, 0]
        check_close(got, OA.wav_quantize(t.data[: t.real_length], norm=True))
    got = read_wave(tmp_path / "out.wav")[:, 0]
    want = OA.listener_output(tracks)
    check_close(got, want)
    assert len(got) >= 3 * 44100 - 1  # at least the three-second buffer, less its last index
    assert np.abs(got).max() >= 26000  # normalised to 0.8 of full scale


def test_render_keyframes_cross_fade(tmp_path):
    dry = write_dry(tmp_path / "dry.wav", 22050, seed=9)
    debug = tmp_path / "debug"
    debug.mkdir()
    s = scenes.shoebox(spec=0.3)
    w = earfile.EarWriter()
    for i, (refl, refr, spec) in enumerate(s["materials"]):
        w.material(f"m{i}", refl, refr, spec)
    w.settings(samples=20_000, seed=3, maxbounces=120, debugdir=str(debug))
    keys = [0.0, 0.2]
    w.keys(keys)
    tris = np.asarray(s["tris"], np.float64)
    for i in range(len(s["materials"])):
        sel = tris[np.asarray(s["mat"]) == i]
        if len(sel):
            w.mesh(f"m{i}", sel.tolist())
    w.source(str(tmp_path / "dry.wav"), location=s["src"])
    p0 = np.asarray(s["recorders"][0]["pos"], np.float64)
    w.mono("out.wav", anim=[tuple(p0), tuple(p0 + np.array([1.5, 0.5, 0.0]))])
    p = tmp_path / "kf.EAR"
    w.save(p)
    run(tmp_path, p)
    bands = OA.equalizer_split_f32(dry)
    k0, k1 = np.float32(keys[0]), np.float32(keys[1])
    tracks = []
    for b in range(3):
        ir0 = response(debug, 0, 0, b, frame=0)
        ir1 = response(debug, 0, 0, b, frame=1)
        sig, off = OA.section(bands[b], k0, np.float32(k1 - k0))
        assert off == 0 and len(sig) == 8820
        tracks.append(OA.process_track(ir0, sig, off, ir2=ir1))
    for b in range(3):
        ir1 = response(debug, 0, 0, b, frame=1)
        sig, off = OA.section(bands[b], k1)
        assert off == 8820 and len(sig) == 22050 - 8820
        tracks.append(OA.process_track(ir1, sig, off))
    # context order: keyframe 0 bands 0..2, keyframe 1 bands 0..2
    order = [0, 1, 2, 3, 4, 5]
    got = read_wave(tmp_path / "out.wav")[:, 0]
    check_close(got, OA.listener_output([tracks[i] for i in order]))
    for i, (kf, b) in enumerate([(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2)]):
        t = tracks[i]
        g = read_wave(debug / f"rec-0.sound-0.frame-{kf:02d}.band-{b}.wav")[:, 0]
        check_close(g, OA.wav_quantize(t.data[: t.real_length], norm=True))


def test_render_stereo_writes_two_channels(tmp_path):
    write_dry(tmp_path / "dry.wav", 8000, seed=2)
    s = scenes.shoebox(stereo=True)
    p = tmp_path / "st.EAR"
    earfile.scene_to_ear(s, p, samples=20_000, seed=5, maxbounces=80, wav=str(tmp_path / "dry.wav"))
    run(tmp_path, p)
    got = read_wave(tmp_path / "out.wav")
    assert got.shape[1] == 2 and got.shape[0] >= 3 * 44100 - 1
    peak = np.abs(got).max()
    assert 25000 <= peak <= 26215  # Normalize(0.8) over both channels
    # the two ears hear different things (interaural delay and head shadow)
    assert np.abs(got[:, 0] - got[:, 1]).max() > 100



def test_render_two_sources(tmp_path):
    """Two sound sources: the CLI renders the (sound, band) contexts of both in
    one batch (sound 0's bands, then sound 1's) and the listener output is the
    sum of all six processed contexts."""
    dry0 = write_dry(tmp_path / "dry0.wav", 12000, seed=11)
    dry1 = write_dry(tmp_path / "dry1.wav", 9000, seed=12)
    debug = tmp_path / "debug"
    debug.mkdir()
    s = scenes.shoebox(spec=0.2)
    w = earfile.EarWriter()
    for i, (refl, refr, spec) in enumerate(s["materials"]):
        w.material(f"m{i}", refl, refr, spec)
    w.settings(samples=20_000, seed=21, maxbounces=120, debugdir=str(debug))
    tris = np.asarray(s["tris"], np.float64)
    for i in range(len(s["materials"])):
        sel = tris[np.asarray(s["mat"]) == i]
        if len(sel):
            w.mesh(f"m{i}", sel.tolist())
    w.source(str(tmp_path / "dry0.wav"), location=s["src"])
    w.source(str(tmp_path / "dry1.wav"), location=(8.0, 2.0, 2.5))
    w.mono("out.wav", s["recorders"][0]["pos"])
    p = tmp_path / "two.EAR"
    w.save(p)
    run(tmp_path, p)
    tracks = []
    for sound, dry in ((0, dry0), (1, dry1)):
        bands = OA.equalizer_split_f32(dry)
        for b in range(3):
            tracks.append(OA.process_track(response(debug, 0, sound, b), bands[b], 0))
    # the sources sit at different places: their responses differ
    assert not np.array_equal(response(debug, 0, 0, 1), response(debug, 0, 1, 1))
    got = read_wave(tmp_path / "out.wav")[:, 0]
    check_close(got, OA.listener_output(tracks))


def test_multiply_quirk_keeps_the_highest_written_bin_unscaled(tmp_path):
    """SET multiplyquirk 1: the reference's FloatBuffer::Multiply stops one
    bin short of real_length, the highest index written (src/Recorder.cpp:
    52-57, 85-89), so after Scene::Render (src/Scene.cpp:286-316) that bin of
    each track holds its paths' sum without 1/N and without gain^2, and Power
    (the dumps are raised to 0.335) skips it as well.  With one bounce and a
    quiet source (gain 0.001) that bin stays above the truncation threshold
    (max / 256), so the dumps show it: the quirk's response runs to it (the
    dump holds real_length + 1 values), its earlier bins equal the default
    render's, and it holds the paths' raw sum.  The bin is the last one whose
    |ramp| sum is nonzero in the same render through the C ABI (the CLI
    traces samples / 10 rays, src/EAR.cpp:78-82)."""
    from ear_amd import api

    s = scenes.shoebox(spec=0.2)
    samples, seed, gain = 2000, 77, 0.001
    n_rays = samples // 10
    dumps = {}
    for quirk in (0, 1):
        d = tmp_path / f"q{quirk}"
        d.mkdir()
        (d / "debug").mkdir()
        write_dry(d / "dry.wav", 4000)
        earfile.scene_to_ear(s, d / "box.EAR", samples=samples, seed=seed, maxbounces=1, wav=str(d / "dry.wav"),
                             src_gain=gain, debugdir=str(d / "debug"), noprocessing=1, multiplyquirk=quirk)
        run(d, d / "box.EAR")
        dumps[quirk] = [np.fromfile(d / "debug" / f"response-0.sound-0.band-{b}{NAMES[b]}.bin", "<f4")
                        for b in range(3)]
    air = tuple(1.0 - a for a in (0.01, 0.02, 0.05))
    with api.Context(0) as ctx, api.Scene(ctx, s["tris"], s["mat"], s["materials"]) as sc:
        r = api.render(ctx, sc, recorders=s["recorders"], src_pos=s["src"], air=air, n_rays=n_rays, seed=seed,
                       max_bounces=1, cutoff=1e-8, n_bins=3 * 44100, abs_hist=True, gain=gain)
    for b in range(3):
        written = np.flatnonzero(np.any(r.abs_raw[b, 0] != 0, axis=-1))
        L = int(written[-1])
        P = float(api.raw_to_float(r.raw[b, 0, L]))
        n, q = dumps[0][b], dumps[1][b]
        print("band", b, "last written bin", L, "dump lengths", len(n), len(q), "P", P, flush=True)
        assert len(q) == L + 2 and len(n) <= L + 2
        np.testing.assert_array_equal(q[:L], n[:L])
        # the paths' sum, neither divided by N nor multiplied by gain^2 nor powered
        np.testing.assert_allclose(q[L], np.float32(P), rtol=1e-6)
        if len(n) > L:
            # the default scales and powers it like every other bin
            want = np.sign(P) * (abs(P) / n_rays * gain * gain) ** 0.335
            np.testing.assert_allclose(n[L], want, rtol=1e-5)
