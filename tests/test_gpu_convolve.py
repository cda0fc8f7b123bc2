"""Auralization convolution (ear_convolve, RecorderTrack::Process) against an
FP64 direct convolution, plain and with the keyframe cross-fade."""
import numpy as np
import pytest

from ear_amd import api

pytestmark = pytest.mark.gpu


def _fades(n):
    """The reference's running float32 fade factors (src/Recorder.cpp:172-181)."""
    df = np.float32(1.0) / np.float32(n)
    f_out, f_in = np.float32(1.0), np.float32(0.0)
    out, inn = np.empty(n, np.float32), np.empty(n, np.float32)
    for i in range(n):
        out[i], inn[i] = f_out, f_in
        f_out = np.float32(f_out - df)
        f_in = np.float32(f_in + df)
    return out, inn


@pytest.mark.parametrize("m,n", [(1, 1), (7, 300), (5000, 20000), (44100, 3000)])
def test_convolve_matches_fp64(gpu_ctx, m, n):
    rng = np.random.default_rng(m + n)
    ir = rng.normal(size=m).astype(np.float32) * np.exp(-np.arange(m) / max(1, m / 5)).astype(np.float32)
    sig = rng.uniform(-1, 1, size=n).astype(np.float32)
    out = api.convolve(gpu_ctx, ir, sig)
    ref = np.convolve(ir.astype(np.float64), sig.astype(np.float64))
    assert len(out) == m + n + 1
    scale = np.abs(ref).max()
    np.testing.assert_allclose(out[: m + n - 1], ref, rtol=0, atol=2e-6 * scale * np.log2(m + n + 1))
    assert np.all(np.abs(out[m + n - 1:]) <= 2e-6 * scale * np.log2(m + n + 1))


def test_convolve_crossfade_matches_fp64(gpu_ctx):
    rng = np.random.default_rng(5)
    m1, m2, n = 3000, 4500, 12000
    ir1 = rng.normal(size=m1).astype(np.float32)
    ir2 = rng.normal(size=m2).astype(np.float32)
    sig = rng.uniform(-1, 1, size=n).astype(np.float32)
    out = api.convolve(gpu_ctx, ir1, sig, ir2=ir2)
    f_out, f_in = _fades(n)
    s1 = (sig * f_out).astype(np.float64)
    s2 = (sig * f_in).astype(np.float64)
    ref = np.zeros(max(m1, m2) + n + 1)
    c1 = np.convolve(ir1.astype(np.float64), s1)
    c2 = np.convolve(ir2.astype(np.float64), s2)
    ref[: len(c1)] += c1
    ref[: len(c2)] += c2
    assert len(out) == len(ref)
    np.testing.assert_allclose(out, ref, rtol=0, atol=2e-6 * np.abs(ref).max() * 15)
    # the fades are linear: the mid sample of a constant signal mixes both tracks equally
    assert abs(float(f_out[n // 2]) - 0.5) < 1e-3 and abs(float(f_in[n // 2]) - 0.5) < 1e-3


def test_convolve_argument_errors(gpu_ctx):
    from ear_amd import _ffi as F

    assert gpu_ctx.lib.ear_convolve(gpu_ctx.h, None, 0, None, 0, None, 0, None) == 1  # EAR_ERR_INVALID_ARGUMENT
    x = np.ones(4, np.float32)
    # a second track given with a null pointer
    assert gpu_ctx.lib.ear_convolve(gpu_ctx.h, F.C.c_void_p(x.ctypes.data), 4, None, 4, F.C.c_void_p(x.ctypes.data), 4,
                                    F.C.c_void_p(x.ctypes.data)) == 1
