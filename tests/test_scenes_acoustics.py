"""Scene generators and acoustic metrics (host logic, CPU only)."""
import math

import numpy as np

from ear_amd import acoustics, scenes


def test_shoebox_is_closed_and_outward():
    s = scenes.shoebox()
    assert s["tris"].shape == (12, 3, 3)
    assert abs(scenes.signed_volume(s["tris"]) - 320.0) < 1e-3
    assert abs(scenes.surface_area(s["tris"]) - 304.0) < 1e-3
    assert len(s["materials"]) == 6
    for refl, refr, spec in s["materials"]:
        assert all(0.6 - 1e-6 <= x <= 0.95 + 1e-6 for x in refl)
        assert refr == [0.0, 0.0, 0.0] and spec == [0.5, 0.5, 0.5]


def test_hall_size_and_materials_are_seeded():
    a = scenes.hall(target_tris=20000)
    b = scenes.hall(target_tris=20000)
    assert np.array_equal(a["tris"], b["tris"])
    assert abs(len(a["tris"]) - 20000) < 300
    assert a["mat"].max() == 7
    alpha = a["alpha"]
    assert alpha.min() >= 0.02 and alpha.max() <= 0.6
    # recorder above every seat block, source inside the shell
    z = a["tris"][:, :, 2]
    assert a["recorders"][0]["pos"][2] > 1.2 and a["recorders"][0]["kind"] == 1
    assert 0 < a["src"][0] < 40 and 0 < a["src"][1] < 28 and z.min() == 0.0


def test_hall_full_size_triangle_count():
    h = scenes.hall()
    assert 195_000 <= len(h["tris"]) <= 205_000


def test_rooms_span_sizes():
    for seed in range(6):
        r = scenes.room(seed=seed, target_tris=3000)
        assert 100 < len(r["tris"]) < 6000
        assert scenes.signed_volume(r["tris"]) > 0


def test_edc_and_t60_of_exponential_decay():
    fs = acoustics.FS
    T = 1.3
    t = np.arange(int(2.5 * fs)) / fs
    e = np.exp(-t * math.log(1e6) / T)
    assert abs(acoustics.t60_iso(e) - T) < 1e-3 * T
    L = acoustics.edc_db(e)
    assert L[0] == 0.0 and L[int(T * fs / 2)] < -29.0
    assert acoustics.edc_error(e, e) == 0.0
    assert abs(acoustics.edc_error(1.01 * e, e) - 0.01) < 1e-12


def test_reference_t60_rule():
    # direct lobe rises then falls; tail above direct/1000 ends at bin 900
    x = np.zeros(2000, np.float32)
    x[100:104] = [0.2, 0.6, 1.0, 0.8]
    x[104:900] = 0.01
    x[900:1500] = 0.0005
    t = acoustics.t60_reference(x)
    assert abs(t - (899 - 103) / 44100.0) < 1e-12


def test_sabine_eyring():
    V, S, a, m = 1000.0, 600.0, 0.2, 0.001
    assert abs(acoustics.sabine(V, S * a, m) - 0.1611 * V / (S * a + 4 * m * V)) < 1e-12
    assert acoustics.eyring(V, S, a, m) < acoustics.sabine(V, S * a, m)


def test_power_is_sign_preserving():
    x = np.array([-8.0, 0.0, 8.0], np.float32)
    p = acoustics.power(x, 1.0 / 3.0)
    np.testing.assert_allclose(p, [-2.0, 0.0, 2.0], rtol=1e-6)
