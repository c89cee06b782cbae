"""Geometric arrival times and time residuals (src/calc.jl). The three
closest_approach_distance cases are the reference's own test
(test/test_calc.jl:4-24); the rest are closed forms."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def C(P):
    from photonprop import calc
    return calc


@pytest.mark.parametrize("p0,d,pos,expect", [
    ([0, 0, 0], [1, 0, 0], [0, 1, 0], 1.0),
    ([0, 0, 0], [0, 1, 0], [1, 1, 0], 1.0),
    ([0, 0, 0], [1, 1, 1], [2, 2, 2], 0.0)])
def test_closest_approach_distance_reference_cases(C, p0, d, pos, expect):
    assert C.closest_approach_distance(p0, d, pos) == pytest.approx(expect, abs=1e-12)


def test_tgeo_cascade(C, P):
    m = P.CascadiaMediumProperties()
    cg = float(m.group_velocity(800.0))
    assert C.calc_tgeo(30.0, 0.2) == pytest.approx(150.0)
    assert C.calc_tgeo(30.0, m) == pytest.approx(30.0 / cg)
    t = P.HomogeneousDetector(P.Spherical(np.float32([0, 0, 30]), 0.3), 1.0, 7)
    assert C.calc_tgeo(30.0, m, t) == pytest.approx(29.7 / cg)


def test_tgeo_track_is_the_cherenkov_cone_minimum(C):
    """For a track the Cherenkov-cone arrival time equals the minimum over
    emission points s of s/c + |pos - p(s)| n_grp / c when the emission
    angle satisfies cos = 1/n_ph (n_grp = n_ph: the classical cone)."""
    n = 1.34
    p0, d, pos = np.zeros(3), np.array([0.0, 0.0, 1.0]), np.array([12.0, 0.0, 40.0])
    s = np.linspace(-200, 200, 2_000_001)
    t = (s + np.linalg.norm(pos[None, :] - s[:, None] * d[None, :], axis=1) * n) / 0.299792458
    assert C.calc_tgeo_tracks(p0, d, pos, n, n) == pytest.approx(t.min(), rel=1e-9)
    assert C.closest_approach_param(p0, d, pos) == pytest.approx(40.0)


def test_shift_particle(C, P):
    from photonprop import lightyield as L
    p = L.Particle(np.array([1.0, 2.0, 3.0]), np.array([0.0, 0.6, 0.8]), 5.0, 1e3, 10.0,
                   L.PMuMinus)
    q = C.shift_particle(p, 10.0)
    np.testing.assert_allclose(q.position, [1.0, 8.0, 11.0])
    assert q.time == pytest.approx(5.0 + 10.0 / 0.299792458)
    r = C.shift_to_closest_approach(p, np.array([1.0, 2.0 + 6.0, 3.0 + 8.0]))
    np.testing.assert_allclose(r.position, [1.0, 8.0, 11.0])


def test_time_residual_columns(C, P):
    import pandas as pd
    m = P.CascadiaMediumProperties()
    t1 = P.HomogeneousDetector(P.Spherical(np.float32([0, 0, 20]), 0.3), 1.0, 4)
    t2 = P.HomogeneousDetector(P.Spherical(np.float32([0, 30, 0]), 0.3), 1.0, 9)
    src = P.PointlikeIsotropicEmitter(np.float32([0, 0, 0]), 2.0, 100)
    setup = P.PhotonPropSetup([src], [t1, t2], m, P.Monochromatic(450.0), 1)
    df = pd.DataFrame({"module_id": [4, 9, 4], "time": [100.0, 150.0, 120.0]})
    C.calc_time_residual(df, setup)
    cg = float(m.group_velocity(800.0))
    np.testing.assert_allclose(df["tres"], [100 - 19.7 / cg - 2, 150 - 29.7 / cg - 2,
                                            120 - 19.7 / cg - 2])
