"""Host light yield and emitter construction (src/lightyield.jl:140-488,
:849-873), pinned by the reference's constants and closed forms."""
import math

import numpy as np
import pytest
from scipy import integrate


def test_longitudinal_parameters_of_a_10_tev_electron(P):
    from photonprop import lightyield as L
    a, b = L.get_longitudinal_params(L.PEMinus, 1e4)
    assert a == pytest.approx(2.01849 + 1.45469 * 4.0, abs=1e-12)  # 7.83725
    assert b == 0.63207
    m = P.CascadiaMediumProperties()
    lp = L.LongitudinalParameterisation.from_energy(1e4, m, L.PEMinus)
    assert lp.a == pytest.approx(7.83725, rel=1e-7)
    # lrad = X0 / rho * 10 (g/cm^2 over kg/m^3 -> m), src/lightyield.jl:191-192
    assert lp.lrad == pytest.approx(m.radiation_length() / m.material_density() * 10, rel=1e-7)
    assert 0.33 < lp.lrad < 0.37
    # gamma profile: mean depth (a-1)/b radiation lengths
    assert (lp.a - 1) / lp.b * lp.lrad == pytest.approx(10.82 * lp.lrad, rel=1e-3)


def test_cascade_track_length_and_total_yield(P):
    from photonprop import lightyield as L
    from photonprop.spectrum import frank_tamm
    assert L.cascade_cherenkov_track_length(1.0, L.PEMinus) == pytest.approx(5.3207078881)
    r = L.cascade_cherenkov_track_length(2e5, L.PEMinus) / L.cascade_cherenkov_track_length(
        1e5, L.PEMinus)
    assert r == pytest.approx(2 ** 1.00000211, rel=1e-12)
    m = P.CascadiaMediumProperties()
    spec = P.make_cherenkov_spectrum((300.0, 800.0), m)
    # independent quadrature of the Frank-Tamm spectrum: 2 pi alpha (1 - 1/n^2) / wl^2
    # with wl in nm, x 1e9 = photons per m of track per nm
    norm, _ = integrate.quad(lambda wl: frank_tamm(wl, m.phase_refractive_index(wl)) * 1e9,
                             300, 800, limit=200)
    assert spec.spectral_dist.normalization == pytest.approx(norm, rel=1e-6)
    part = L.Particle(np.zeros(3), np.array([0, 0, 1.0]), 0.0, 1.0, 0.0, L.PEMinus)
    y = L.total_lightyield(part, m, spec)
    assert y == pytest.approx(norm * 5.3207078881, rel=1e-6)
    # ~2.3e5 photons per GeV in 300-800 nm (the scale the reference's
    # benchmark sizing implies)
    assert 2.2e5 < y < 2.4e5


def test_rel_additional_track_length(P):
    from photonprop import lightyield as L
    assert L.rel_additional_track_length(1.30, math.e) == pytest.approx(0.1842 + 0.0204)
    assert L.rel_additional_track_length(1.36, 1.0) == pytest.approx(0.1916)


def test_lightsabre_muon_yield(P):
    from photonprop import lightyield as L
    m = P.CascadiaMediumProperties()
    spec = P.make_cherenkov_spectrum((300.0, 800.0), m)
    part = L.Particle(np.zeros(3), np.array([0, 0, 1.0]), 0.0, 1e4, 100.0, L.PMuMinus)
    le = 4.0
    coeffs = L._LIGHTSABRE_COEFFS
    stoch = 10 ** sum(c * le ** i for i, c in enumerate(coeffs)) * 100.0
    bare = L.total_lightyield(part, m, spec)
    rng = np.random.default_rng(1)
    n = [L.FastLightsabreMuonEmitter(part, m, spec, rng=rng).photons for _ in range(200)]
    mean = bare + stoch
    assert np.mean(n) == pytest.approx(mean, rel=5 * math.sqrt(mean / 200) / mean + 1e-9)
    em = L.FastLightsabreMuonEmitter(part, m, spec, rng=rng)
    assert em.length == 100.0 and isinstance(em, L.CherenkovTrackEmitter)
    with pytest.raises(ValueError):
        L.FastLightsabreMuonEmitter(
            L.Particle(np.zeros(3), np.array([0, 0, 1.0]), 0.0, 5.0, 100.0, L.PMuMinus), m, spec)


def test_split_and_rescale_source(P):
    from photonprop import lightyield as L
    s = L.PointlikeIsotropicEmitter(np.float32([0, 0, 0]), 0.0, 103)
    parts = L.split_source(s, 4)
    assert [p.photons for p in parts] == [25, 25, 25, 28]
    with pytest.raises(ValueError):
        L.split_source(L.PointlikeIsotropicEmitter(np.float32([0, 0, 0]), 0.0, 3), 4)
    assert L.rescale_source(s, 0.5).photons == 52  # ceil(51.5)
