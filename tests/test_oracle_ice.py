"""Ice medium and the simplified-Liu phase function (src/ice_properties.jl)
in the oracle, pinned by closed forms: the SL inverse CDF against its CDF
((1+c)/2)^((1+g)/(1-g)) (KS), E[cos] = g for SL and for the HG+SL mixture,
and the medium values against the host formulas of the same reference lines."""
import numpy as np
import pytest
from scipy import stats


@pytest.mark.parametrize("g", [0.9, 0.6])
def test_simplified_liu_cdf_and_mean(O, g):
    rng = np.random.default_rng(int(g * 10))
    u = rng.random(400_000).astype(np.float32)
    c = O.eval_fn("SL_COS", np.stack([np.full_like(u, g), u], 1))[:, 0].astype(np.float64)
    a = (1 + g) / (1 - g)
    assert stats.kstest(c, lambda x: ((1 + np.clip(x, -1, 1)) / 2) ** a).pvalue > 1e-3
    assert c.mean() == pytest.approx(g, abs=4 * c.std() / np.sqrt(len(c)))
    assert c.min() >= -1 and c.max() <= 1


def test_mixture_mean_is_g(O, P):
    m = P.make_homogenous_clearice_properties()
    rng = np.random.default_rng(3)
    n = 400_000
    u = rng.random((n, 2)).astype(np.float32)
    hg = O.eval_fn("HG_COS", np.stack([np.full(n, 0.9, np.float32), u[:, 1]], 1))[:, 0]
    sl = O.eval_fn("SL_COS", np.stack([np.full(n, 0.9, np.float32), u[:, 1]], 1))[:, 0]
    c = np.where(u[:, 0] < m.hg_fraction, hg, sl).astype(np.float64)
    assert c.mean() == pytest.approx(0.9, abs=4 * c.std() / np.sqrt(n))


def test_ice_medium_values_match_host(O, P):
    m = P.make_homogenous_clearice_properties(sca_scale=1.3)
    wl = np.linspace(300.0, 800.0, 501, dtype=np.float32)
    r = O.eval_fn("CASCADIA", wl[:, None], medium=m.to_pod())
    # the medium POD is float32: 1 - float32(0.9) alone differs by 2e-7
    np.testing.assert_allclose(r[:, 0], m.scattering_length(wl), rtol=1e-6)
    np.testing.assert_allclose(r[:, 1], m.group_velocity(wl), rtol=2e-7)
    np.testing.assert_allclose(r[:, 2], m.phase_refractive_index(wl), rtol=2e-7)
    # sanity of the reference's clear-ice model: effective scattering length
    # sca_len / (1 - g) ~ 43 m at 400 nm, n ~ 1.32, absorption length ~150 m
    assert 30 < m.scattering_length(400.0) / (1 - 0.9) / 1.3 < 60
    assert 1.30 < m.phase_refractive_index(400.0) < 1.34
    assert 50 < m.absorption_length(400.0) < 300
