"""Whole-photon oracle checks against closed forms and the reference's
first-hit and photon-count semantics (src/photon_prop_cuda.jl:456-537)."""
import numpy as np
from scipy import stats


def pod_sources(P, srcs):
    return P.processing.sources_to_pod(srcs)


def run(O, P, srcs, medium, targets, seed=1, n_steps=15, **kw):
    return O.propagate(pod_sources(P, srcs), medium.to_pod(), P.targets_to_pod(targets), seed,
                       n_steps=n_steps, **kw)


def test_no_scatter_direct_hits(O, P):
    """sca_len -> inf: only direct hits, fraction (1 - sqrt(1 - r^2/d^2))/2,
    t = |x| / c_grp, dist = |x|, one step, direction unchanged."""
    d, r, cg = 10.0, 0.3, 0.2
    medium = P.SimpleMediumProperties(30.0, 1e9, 1.35, 0.9, cg)
    targets = [P.spherical_target((0, 0, d), r, 1)]
    n = 2_000_000
    hits, nsim = run(O, P, [P.PointlikeIsotropicEmitter(np.zeros(3), 0.0, n)], medium, targets,
                     mono_wl=450.0)
    assert nsim == n
    p = (1 - np.sqrt(1 - r * r / (d * d))) / 2
    assert stats.binomtest(len(hits), n, p).pvalue > 1e-3
    assert np.all(hits["n_steps"] == 1)
    dist = np.linalg.norm(hits["pos"].astype(np.float64), axis=1)
    np.testing.assert_allclose(hits["dist"], dist, rtol=1e-5)
    np.testing.assert_allclose(hits["time"], dist / cg, rtol=1e-5)
    np.testing.assert_allclose(np.linalg.norm(hits["pos"] - [0, 0, d], axis=1), r, rtol=1e-4)
    np.testing.assert_array_equal(hits["dir"], hits["init_dir"])


def test_first_target_in_index_order_wins(O, P):
    """Two spheres on the beam: the first in index order whose near root lies
    within the step is reported, even if another sphere is nearer."""
    medium = P.SimpleMediumProperties(30.0, 1e9, 1.35, 0.9, 0.2)
    beam = P.PencilEmitter(np.zeros(3), np.float32([0, 0, 1]), 0.0, 0.0, 100)
    far_first = [P.spherical_target((0, 0, 20), 0.3, 7), P.spherical_target((0, 0, 10), 0.3, 8)]
    hits, _ = run(O, P, [beam], medium, far_first, mono_wl=400.0)
    assert len(hits) == 100
    assert np.all(hits["module_ix"] == 1)
    np.testing.assert_allclose(hits["dist"], 19.7, rtol=1e-6)
    hits, _ = run(O, P, [beam], medium, far_first[::-1], mono_wl=400.0)
    assert np.all(hits["module_ix"] == 1)
    np.testing.assert_allclose(hits["dist"], 9.7, rtol=1e-6)


def test_short_steps_reach_near_sphere_first(O, P):
    """With sca_len ~ distance both spheres are reachable: every hit on the far
    sphere has crossed the near sphere's plane in an earlier step."""
    medium = P.SimpleMediumProperties(30.0, 5.0, 1.35, 0.99, 0.2)
    beam = P.PencilEmitter(np.zeros(3), np.float32([0, 0, 1]), 0.0, 0.0, 20_000)
    t = [P.spherical_target((0, 0, 20), 0.3, 1), P.spherical_target((0, 0, 10), 0.3, 2)]
    hits, _ = run(O, P, [beam], medium, t, mono_wl=400.0)
    assert np.any(hits["module_ix"] == 2) and np.any(hits["module_ix"] == 1)


def test_photon_count_float32_product(O):
    """photons * photon_scaling is a Float32 product, floored
    (src/photon_prop_cuda.jl:456-460)."""
    assert O.photon_count(123456789, 1.0) == 123456792
    assert O.photon_count(1001, 0.5) == 500
    assert O.photon_count(7, 0.3) == 2
    assert O.photon_count(0, 3.0) == 0


def test_scaling_changes_photon_count(O, P):
    medium = P.CascadiaMediumProperties()
    t = [P.spherical_target((0, 0, 10), 0.3, 1)]
    _, nsim = run(O, P, [P.PointlikeIsotropicEmitter(np.zeros(3), 0.0, 1001)], medium, t,
                  photon_scaling=0.5, mono_wl=450.0)
    assert nsim == 500


def test_oracle_is_deterministic_and_thread_independent(O, P):
    medium = P.CascadiaMediumProperties()
    t = P.make_detector_grid(2, 2, 20.0, 5, 10.0)
    src = P.ExtendedCherenkovEmitter(np.float32([3, 4, 5]), np.float32([0, 0.6, 0.8]), 0.0,
                                     200_000, P.LongitudinalParameterisation(7.837, 0.632, 0.356))
    h1, n1 = run(O, P, [src], medium, t, mono_wl=420.0, n_threads=1)
    h2, n2 = run(O, P, [src], medium, t, mono_wl=420.0, n_threads=8)
    assert n1 == n2 == 200_000
    assert len(h1) > 50
    assert h1.tobytes() == h2.tobytes()


def test_zero_steps_and_zero_targets(O, P):
    medium = P.CascadiaMediumProperties()
    src = P.PointlikeIsotropicEmitter(np.zeros(3), 0.0, 1000)
    hits, n = run(O, P, [src], medium, [P.spherical_target((0, 0, 1), 0.5, 1)], n_steps=0,
                  mono_wl=450.0)
    assert len(hits) == 0 and n == 1000
