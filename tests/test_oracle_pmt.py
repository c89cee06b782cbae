"""PMT-level detection of the POM: the oracle (oracle/pp_pmt_oracle.c) and the
host model, pinned by the reference's PMT layout (src/pom.jl:108-141) and the
uniform-flux identity its own comment states (src/pom.jl:285-287): for photons
spread uniformly over the sphere the summed per-PMT probability averages to
acc_1(wl) + acc_2(wl)."""
import numpy as np
import pytest
from scipy import stats


@pytest.fixture(scope="module")
def pom(P):
    from photonprop import pom
    return pom


def uniform_on_sphere(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_pmt_layout(O):
    d = O.pom_pmt_dirs()
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    # calc_rot_matrix(e_z, e_x) turns the module axis to +x: polar angles
    # 65, 32.5, 115, 147.5 deg about +x, four PMTs each, 90 deg apart
    ax = d[:, 0]
    for k, th in enumerate([65.0, 32.5, 115.0, 147.5]):
        np.testing.assert_allclose(ax[4 * k:4 * k + 4], np.cos(np.deg2rad(th)), atol=1e-12)
        az = np.arctan2(d[4 * k:4 * k + 4, 2], d[4 * k:4 * k + 4, 1])
        gaps = np.sort(np.mod(np.diff(np.sort(az)), 2 * np.pi))
        np.testing.assert_allclose(gaps, np.pi / 2, atol=1e-12)
    # an orientation rotates every PMT direction
    rz = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    np.testing.assert_allclose(O.pom_pmt_dirs(rz), d @ rz.T, atol=1e-7)


def test_acceptance_table(pom):
    acc = pom.default_acceptance()
    assert len(acc.wavelengths) == 26 and acc.wavelengths[0] == 300 and acc.wavelengths[-1] == 800
    assert 0.1 < acc.sigma_1 < acc.sigma_2 < 0.5
    assert np.all(acc.acc_1 >= 0) and np.all(acc.acc_2 >= 0)
    assert acc.pos_wl_acc(1, 250.0) == 0.0 and acc.pos_wl_acc(2, 900.0) == 0.0


@pytest.mark.parametrize("wl", [350.0, 420.0, 555.0])
def test_uniform_flux_identity(O, pom, wl):
    acc = pom.default_acceptance()
    rng = np.random.default_rng(int(wl))
    n = 400_000
    rel = uniform_on_sphere(rng, n)
    ids, psum = O.pom_pmt_hits(rel, np.full(n, wl), np.arange(n), np.zeros(n), 5,
                               acc.wavelengths, acc.acc_1, acc.acc_2, acc.sigma_1, acc.sigma_2)
    expect = acc.pos_wl_acc(1, wl) + acc.pos_wl_acc(2, wl)
    # the Rayleigh mass beyond pi is negligible for these widths
    assert psum.mean() == pytest.approx(expect, rel=5e-3)
    # detected fraction is the mean probability (binomial)
    k = int((ids > 0).sum())
    assert stats.binomtest(k, n, psum.mean()).pvalue > 1e-3


def test_pmt_choice_follows_the_weights(O, pom):
    acc = pom.default_acceptance()
    d = O.pom_pmt_dirs()
    # a point between PMTs 0 and 4, many photons (distinct ids) at one wavelength
    p = d[0] + d[4]
    p /= np.linalg.norm(p)
    n = 300_000
    ids, psum = O.pom_pmt_hits(np.tile(p, (n, 1)), np.full(n, 400.0), np.arange(n), np.zeros(n),
                               9, acc.wavelengths, acc.acc_1, acc.acc_2, acc.sigma_1,
                               acc.sigma_2)
    sig = np.array([acc.sigma_1, acc.sigma_2])[(np.arange(16) >> 2) & 1]
    a = np.array([acc.pos_wl_acc(1, 400.0), acc.pos_wl_acc(2, 400.0)])[(np.arange(16) >> 2) & 1]
    pt = np.arccos(np.clip(d @ p, -1, 1))
    w = pt / sig**2 * np.exp(-pt**2 / (2 * sig**2)) / (0.5 * np.sin(pt)) * a / 8
    assert psum[0] == pytest.approx(w.sum(), rel=1e-6)  # float32 inputs
    counts = np.bincount(ids[ids > 0], minlength=17)[1:]
    big = w / w.sum() * counts.sum() > 20
    chi = stats.chisquare(counts[big], w[big] / w[big].sum() * counts[big].sum())
    assert chi.pvalue > 1e-3
    assert set(np.nonzero(counts)[0]) <= set(np.nonzero(w)[0])


def test_hit_at_a_pmt_centre_and_outside_the_table(O, pom):
    acc = pom.default_acceptance()
    d = O.pom_pmt_dirs()
    ids, psum = O.pom_pmt_hits(d[:1], [450.0], [1], [0], 3, acc.wavelengths, acc.acc_1,
                               acc.acc_2, acc.sigma_1, acc.sigma_2)
    # sin(pt) = 0 gives that PMT weight 0 (src/pom.jl:279), the others still count
    assert psum[0] > 0
    ids, psum = O.pom_pmt_hits(d[:1] * 0.5 + d[1:2] * 0.5, [900.0], [1], [0], 3, acc.wavelengths,
                               acc.acc_1, acc.acc_2, acc.sigma_1, acc.sigma_2)
    assert psum[0] == 0.0 and ids[0] == 0
