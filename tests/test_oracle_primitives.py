"""Pins of the CPU oracle's building blocks against published vectors and
mathematics (not against the oracle's own formulas)."""
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def load_kat():
    rows = []
    with open(os.path.join(HERE, "golden", "philox4x32_10_kat.txt")) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            w = [int(x, 16) for x in line.split()]
            rows.append((w[0:4], w[4:6], w[6:10]))
    return rows


def test_philox_known_answers(oracle):
    kat = load_kat()
    assert len(kat) == 3
    for ctr, key, expect in kat:
        assert oracle.philox(ctr, key) == expect


def test_uniform_mapping_is_top24_bits(oracle):
    # u = (x >> 8) * 2^-24 of word (draw & 3) of block (draw >> 2)
    seed = 0x0123456789ABCDEF
    for ray, ctx, inst in [(0, 0, 0), (7, 2, 5), (123456, 0x10001, 3)]:
        for draw in range(12):
            x = oracle.philox([draw >> 2, ray, ctx, inst], [seed & 0xFFFFFFFF, seed >> 32])[draw & 3]
            assert oracle.uniform(seed, ray, ctx, inst, draw) == (x >> 8) / 16777216.0


def test_pow_against_libm(oracle):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(0, 1, 4000), 1 - 10.0 ** -rng.uniform(1, 7, 2000), rng.uniform(0.9, 1, 2000)])
    ys = np.concatenate([rng.uniform(0, 60, 4000), np.full(2000, 1000.0), rng.uniform(0, 5, 2000)])
    worst_d = 0.0
    ulps = []
    for x, y in zip(xs.astype(np.float32), ys.astype(np.float32)):
        ref = math.pow(float(x), float(y))
        got = oracle.pow_d(float(x), float(y))
        if ref > 1e-300:
            worst_d = max(worst_d, abs(got - ref) / ref)
        f = np.float32(oracle.powf(float(x), float(y)))
        r32 = np.float32(ref)
        if r32 != 0 and np.isfinite(r32):
            ulps.append(abs(int(f.view(np.int32)) - int(r32.view(np.int32))))
    assert worst_d < 1e-12
    ulps = np.array(ulps)
    assert ulps.max() <= 1
    assert (ulps == 0).mean() > 0.9999


def test_pow_special_cases(oracle):
    assert oracle.powf(1.0, 1e6) == 1.0
    assert oracle.powf(0.5, 0.0) == 1.0
    assert oracle.powf(0.0, 3.0) == 0.0
    assert oracle.powf(0.5, 2000.0) == 0.0  # underflows binary32
    assert math.isnan(oracle.powf(-0.5, 2.0))


def _fp64_triangle(v0, e1, e2, o, d):
    """o + t d = v0 + u e1 + v e2 solved in float64 by Cramer's rule."""
    A = np.stack([-d, e1, e2], axis=1).astype(np.float64)
    rhs = (o - v0).astype(np.float64)
    det = np.linalg.det(A)
    sol = np.linalg.solve(A, rhs)
    return det, sol  # t, u, v


def test_moeller_trumbore_against_fp64(oracle):
    rng = np.random.default_rng(2)
    agree = checked = 0
    for _ in range(4000):
        v0 = rng.uniform(-2, 2, 3).astype(np.float32)
        e1 = rng.uniform(-2, 2, 3).astype(np.float32)
        e2 = rng.uniform(-2, 2, 3).astype(np.float32)
        o = rng.uniform(-4, 4, 3).astype(np.float32)
        target = v0 + e1 * rng.uniform(-0.3, 1.0) + e2 * rng.uniform(-0.3, 1.0)
        d = (target - o).astype(np.float32)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        hit, t = oracle.mt_intersect(v0, e1, e2, o, d)
        mt_det = float(np.dot(e1.astype(np.float64), np.cross(d.astype(np.float64), e2.astype(np.float64))))
        det, (t64, u, v) = _fp64_triangle(v0, e1, e2, o, d)
        margin = min(abs(u), abs(v), abs(1 - u - v), abs(u - 1), abs(v - 1))
        if margin < 1e-4 or abs(abs(mt_det) - 1e-5) < 1e-7:
            continue
        checked += 1
        inside = u >= 0 and v >= 0 and u + v <= 1 and abs(mt_det) >= 1e-5
        assert hit == inside
        if hit:
            # binary32 error grows with the conditioning ||e1|| ||e2|| / |det|
            kappa = float(np.linalg.norm(e1) * np.linalg.norm(e2)) / abs(mt_det)
            assert abs(t - t64) <= 2e-6 * (1.0 + abs(t64)) * kappa
            agree += 1
    assert checked > 3000 and agree > 500


def test_closest_hit_against_fp64_bruteforce(oracle):
    rng = np.random.default_rng(3)
    tris = rng.uniform(-5, 5, size=(300, 3, 3)).astype(np.float32)
    sc = oracle.OracleScene(tris, np.zeros(300, np.uint16), [([0.5] * 3, None, None)])
    e1 = (tris[:, 1] - tris[:, 0]).astype(np.float64)
    e2 = (tris[:, 2] - tris[:, 0]).astype(np.float64)
    v0 = tris[:, 0].astype(np.float64)
    n_checked = 0
    for _ in range(300):
        o = rng.uniform(-6, 6, 3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        od, dd = o.astype(np.float64), d.astype(np.float64)
        p = np.cross(dd, e2)
        det = np.einsum("ij,ij->i", e1, p)
        tv = od - v0
        u = np.einsum("ij,ij->i", tv, p) / det
        q = np.cross(tv, e1)
        v = (q @ dd) / det
        t = np.einsum("ij,ij->i", e2, q) / det
        ok = (np.abs(det) >= 1e-5) & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 0.001) & (t < 1e6)
        near = (np.minimum.reduce([np.abs(u), np.abs(v), np.abs(1 - u - v)]) < 1e-4) & (t > 0)
        if near.any():
            continue
        cand = np.nonzero(ok)[0]
        ot, oi = sc.closest_hit(o, d)
        if len(cand) == 0:
            assert oi == -1
            continue
        best = cand[np.argmin(t[cand])]
        srt = np.sort(t[cand])
        if len(srt) > 1 and srt[1] - srt[0] < 1e-4:
            continue
        n_checked += 1
        assert oi == best
        assert abs(ot - t[best]) < 1e-5 * max(1, t[best])
    assert n_checked > 100


def test_triangle_normal_is_unit_cross(oracle):
    v = np.array([[0, 0, 0], [2, 0, 0], [0, 3, 0]], np.float32)
    assert np.array_equal(oracle.tri_normal(v), np.array([0, 0, 1], np.float32))
    rng = np.random.default_rng(4)
    for _ in range(100):
        v = rng.normal(size=(3, 3)).astype(np.float32)
        n = oracle.tri_normal(v)
        c = np.cross((v[1] - v[0]).astype(np.float64), (v[2] - v[0]).astype(np.float64))
        np.testing.assert_allclose(n, c / np.linalg.norm(c), atol=1e-6)


def test_material_retention_and_invalid(oracle):
    # retention = refl + refr as the Material ctor computes it (src/Material.cpp:32-59)
    r = oracle.material_retention([0.7, 0.5, 0.3])
    np.testing.assert_allclose(r, [0.7, 0.5, 0.3], atol=1e-7)
    r = oracle.material_retention([0.5, 0.5, 0.5], [0.25, 0.1, 0.0])
    np.testing.assert_allclose(r, [0.75, 0.6, 0.5], atol=1e-7)
    assert oracle.material_retention([0.8, 0.5, 0.5], [0.3, 0.1, 0.0]) is None  # refl + refr > 1 throws


def test_sphere_sampling_is_uniform_and_draw_discipline(oracle):
    n = 200_000
    v, nd = oracle.sample_dirs(11, n, mode=0)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=2e-7)
    # Archimedes: z is uniform on [-1, 1] for a uniform direction
    counts, _ = np.histogram(v[:, 2], bins=20, range=(-1, 1))
    chi2 = ((counts - n / 20) ** 2 / (n / 20)).sum()
    assert chi2 < 60  # 19 dof, p ~ 1e-5
    oct_ = (v[:, 0] > 0) * 4 + (v[:, 1] > 0) * 2 + (v[:, 2] > 0)
    c8 = np.bincount(oct_, minlength=8)
    assert ((c8 - n / 8) ** 2 / (n / 8)).sum() < 40
    # 3 uniforms per try; acceptance pi/6 (1 - 0.001^1.5) of the cube
    assert np.all(nd % 3 == 0)
    p = math.pi / 6 * (1 - 0.001 ** 1.5)
    assert abs(nd.mean() - 3 / p) < 0.02 * 3 / p


def test_hemisphere_sampling(oracle):
    n = 200_000
    nrm = np.array([0.3, -0.5, 0.8], np.float32)
    nrm /= np.linalg.norm(nrm)
    v, nd = oracle.sample_dirs(12, n, mode=1, normal=nrm)
    c = v @ nrm
    assert c.min() >= 0
    # uniform (not cosine) over the hemisphere: cos is uniform on [0,1], mean 1/2
    assert abs(c.mean() - 0.5) < 4 * (1 / math.sqrt(12 * n))
    counts, _ = np.histogram(c, bins=10, range=(0, 1))
    assert ((counts - n / 10) ** 2 / (n / 10)).sum() < 40
    p = math.pi / 6 * (1 - 0.001 ** 1.5)
    assert abs(nd.mean() - 2 * 3 / p) < 0.03 * 6 / p


def test_reference_fp32_ramp_against_exact(oracle):
    rng = np.random.default_rng(5)
    for _ in range(300):
        a = float(rng.uniform(-1, 1) * 10 ** rng.uniform(-8, 1))
        dist = float(rng.uniform(0.5, 1500))
        t = dist / 343.0
        buf = oracle.record_f32(200_000, a, t, dist)
        width = np.float32(math.sqrt(np.float32(dist)))
        w = int(math.ceil(width))
        s = int(float(np.float32(t)) * 44100.0)  # (int)(t * 44100.0), double product
        A = np.float32(2.0) * np.float32(a) / np.float32(np.sqrt(np.float32(dist)))
        exact = (w - np.arange(w)) * (float(A) / w)
        got = buf[s:s + w].astype(np.float64)
        assert np.count_nonzero(buf) <= w
        np.testing.assert_allclose(got, exact, rtol=w * 2 ** -22, atol=abs(float(A)) * w * 2 ** -23)
        # total of the ramp: A (w + 1) / 2
        assert abs(got.sum() - float(A) * (w + 1) / 2) <= abs(float(A)) * w * w * 2 ** -22
