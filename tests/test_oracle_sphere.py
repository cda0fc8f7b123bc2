"""The receiver-sphere estimator of the oracle (EAR_EST_SPHERE) against
closed forms.

In free field every ray is one straight segment.  A ray crossing a sphere of
radius r deposits chord / (4/3 pi r^3), so the mean deposit per emitted ray
is (1 / V) * (1 / 4 pi) * integral over the ball of dV / s^2 (s = distance
from the source), whose closed form is

    (3 / (4 pi r^3)) * (1 / 2) * (r - (d^2 - r^2) / (2 d) * ln((d + r) / (d - r)))

(-> 1 / (4 pi d^2) for r << d).  A mirror plane (spec = 1) adds the same
form at the image source's distance, times the reflection coefficient and
with the sign of one reflection."""
import numpy as np
import pytest

from oracle import oracle as O


def ball_average(d, r):
    return (3.0 / (4.0 * np.pi * r ** 3)) * 0.5 * (r - (d * d - r * r) / (2.0 * d) * np.log((d + r) / (d - r)))


def free_field():
    return O.OracleScene(np.zeros((0, 3, 3), np.float32), np.zeros(0, np.uint16), [((0.5, 0.5, 0.5), None, None)])


def test_closed_form_limit():
    # small spheres: the inverse-square law of the direct sound, times
    # 1 + (r/d)^2 / 5 + O((r/d)^4)
    d, r = 2.0, 0.1
    assert abs(ball_average(d, r) * 4 * np.pi * d * d - (1.0 + (r / d) ** 2 / 5.0)) < 1e-5


@pytest.mark.parametrize("d,r,n,tol", [(3.0, 1.0, 8_000_000, 0.015), (5.0, 0.5, 16_000_000, 0.03)])
@pytest.mark.parametrize("axis", [0, 2])
def test_free_field_sphere(d, r, n, tol, axis):
    pos = [0.0, 0.0, 0.0]
    pos[axis] = d
    o = O.render(free_field(), band=1, air=1.0, recorders=[dict(pos=pos, radius=r)], n_rays=n, src_pos=(0, 0, 0),
                 n_threads=8, n_bins=4096, estimator=1)
    assert o["segments"] == n  # one escaping segment per ray
    assert o["shadow_rays"] == 0
    got = o["tally"][0] / n
    # ~4 standard deviations of the Monte-Carlo mean at these ray counts
    assert abs(got / ball_average(d, r) - 1.0) < tol, (got, ball_average(d, r))
    # hit fraction: the sphere's solid angle
    frac = o["contributions"] / n
    p = (1.0 - np.sqrt(1.0 - r * r / (d * d))) / 2.0
    assert abs(frac / p - 1.0) < tol
    # the histogram holds the same deposits, spread by the a9 ramp (total A (w + 1) / 2 each)
    assert o["hist_signed"].sum() > 0 and np.all(o["hist_signed"] >= 0)


def test_dry_scales_first_segment_and_air_attenuates():
    sc = free_field()
    rec = [dict(pos=(3.0, 0, 0), radius=1.0)]
    a = O.render(sc, band=1, air=1.0, recorders=rec, n_rays=200_000, n_threads=8, n_bins=4096, estimator=1)
    b = O.render(sc, band=1, air=1.0, dry=0.25, recorders=rec, n_rays=200_000, n_threads=8, n_bins=4096,
                 estimator=1)
    assert b["contributions"] == a["contributions"]
    np.testing.assert_allclose(b["tally"], 0.25 * a["tally"], rtol=1e-6)
    c = O.render(sc, band=1, air=0.9, recorders=rec, n_rays=200_000, n_threads=8, n_bins=4096, estimator=1)
    # entry distances lie in [d - r, d]: the attenuation is between 0.9^3 and 0.9^2
    ratio = c["tally"][0] / a["tally"][0]
    assert 0.9 ** 3 - 1e-6 <= ratio <= 0.9 ** 2 + 1e-6


def test_mirror_plane_image_source():
    big = 1000.0
    tris = np.array([[[-big, -big, 0], [big, -big, 0], [big, big, 0]],
                     [[-big, -big, 0], [big, big, 0], [-big, big, 0]]], np.float32)
    refl = 0.7
    sc = O.OracleScene(tris, np.zeros(2, np.uint16), [((refl,) * 3, (0.0,) * 3, (1.0,) * 3)])
    src = (0.0, 0.0, 2.0)
    cen = (3.0, 0.0, 1.5)
    r = 0.5
    n = 4_000_000
    o = O.render(sc, band=1, air=1.0, recorders=[dict(pos=cen, radius=r)], n_rays=n, src_pos=src, n_threads=8,
                 n_bins=8192, estimator=1, cutoff=0.0)
    d = np.linalg.norm(np.subtract(cen, src))
    d_img = np.linalg.norm(np.subtract(cen, (0.0, 0.0, -2.0)))
    direct = (o["tally_abs"][0] + o["tally"][0]) / 2 / n   # positive deposits (no reflection)
    mirror = (o["tally_abs"][0] - o["tally"][0]) / 2 / n   # negative deposits (one reflection)
    assert abs(direct / ball_average(d, r) - 1.0) < 0.05
    assert abs(mirror / (refl * ball_average(d_img, r)) - 1.0) < 0.06


def test_nee_histograms_unchanged_by_radius():
    # the default estimator ignores the radius
    sc = free_field()
    a = O.render(sc, band=1, air=1.0, recorders=[dict(pos=(3.0, 0, 0))], n_rays=1000, n_bins=4096)
    b = O.render(sc, band=1, air=1.0, recorders=[dict(pos=(3.0, 0, 0), radius=0.7)], n_rays=1000, n_bins=4096)
    assert np.array_equal(a["hist_signed"], b["hist_signed"]) and a["direct"]["visible"][0] == 1


def test_invalid_sphere_radius():
    with pytest.raises(RuntimeError):
        O.render(free_field(), band=1, air=1.0, recorders=[dict(pos=(3.0, 0, 0))], n_rays=10, n_bins=64,
                 estimator=1)
