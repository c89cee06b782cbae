"""Planar targets in the oracle: check_intersection(::Rectangular /
::Circular) (src/photon_prop_cuda.jl:353-404), pinned by construction and by
closed-form solid angles.

Rays are built backwards from a chosen crossing point, so the expected
outcome (inside or outside the shape, crossing distance) is known without
evaluating the formula under test; the no-scatter hit fractions are the
solid angles of a disc and a rectangle seen from a point on their axis."""
import numpy as np
from scipy import stats

RNG = np.random.default_rng(77)
RECT, DISC = 1, 2


def rows(kind, c, a, b, pos, d, step):
    n = len(pos)
    x = np.zeros((n, 13), np.float32)
    x[:, 0] = np.array([kind], np.uint32).view(np.float32)[0]
    x[:, 1:4] = c
    x[:, 4] = a
    x[:, 5] = b
    x[:, 6:9] = pos
    x[:, 9:12] = d
    x[:, 12] = step
    return x


def unit_rows(n, up=True):
    v = RNG.normal(size=(n, 3))
    v[:, 2] = np.abs(v[:, 2]) + 0.05
    if not up:
        v[:, 2] *= -1
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def backwards(cross, d, dist):
    return cross - d * dist[:, None]


def test_rectangle_crossings_by_construction(O):
    n = 50_000
    c = np.array([3.0, -2.0, 7.0])
    lx, ly = 4.0, 1.5
    # crossing points well inside (|u| < 0.95 half side) and well outside
    # (|u| > 1.05 half side on one axis)
    u = RNG.uniform(-0.95, 0.95, (n, 2)) * [lx / 2, ly / 2]
    out = RNG.uniform(1.05, 3.0, (n, 2)) * [lx / 2, ly / 2] * RNG.choice([-1, 1], (n, 2))
    pick = RNG.random(n) < 0.5
    off = np.where(pick[:, None], u, np.stack([out[:, 0], u[:, 1]], 1))
    cross = np.column_stack([c[0] + off[:, 0], c[1] + off[:, 1], np.full(n, c[2])])
    for up in (True, False):
        d = unit_rows(n, up)
        dist = RNG.uniform(0.5, 30.0, n)
        pos = backwards(cross, d, dist)
        y = O.eval_fn("PLANE_ISEC", rows(RECT, c, lx, ly, pos, d, 40.0))
        np.testing.assert_array_equal(y[:, 0] == 1, pick)
        np.testing.assert_allclose(y[pick, 1], dist[pick], rtol=2e-5, atol=2e-5)
        # step shorter than the crossing: no hit
        y = O.eval_fn("PLANE_ISEC", rows(RECT, c, lx, ly, pos, d, 0.25))
        assert not np.any(y[:, 0])


def test_disc_crossings_by_construction(O):
    n = 50_000
    c = np.array([-1.0, 4.0, -3.0])
    r = 2.5
    rad = np.where(RNG.random(n) < 0.5, RNG.uniform(0, 0.97, n), RNG.uniform(1.03, 3.0, n)) * r
    phi = RNG.uniform(0, 2 * np.pi, n)
    cross = np.column_stack([c[0] + rad * np.cos(phi), c[1] + rad * np.sin(phi), np.full(n, c[2])])
    d = unit_rows(n)
    dist = RNG.uniform(0.5, 30.0, n)
    y = O.eval_fn("PLANE_ISEC", rows(DISC, c, r, 0.0, backwards(cross, d, dist), d, 40.0))
    inside = rad < r
    np.testing.assert_array_equal(y[:, 0] == 1, inside)
    np.testing.assert_allclose(y[inside, 1], dist[inside], rtol=2e-5, atol=2e-5)


def test_plane_window_and_edges(O):
    """0 <= d <= step inclusive, strict inequality on the shape's edge, no hit
    parallel to the plane, and the plane is crossed from either side."""
    c = np.array([0.0, 0.0, 10.0])
    up = np.array([[0.0, 0.0, 1.0]])
    cases = [
        (RECT, (0, 0, 10), up, 5.0, 1, 0.0),    # starts on the plane: d = 0 hits
        (RECT, (0, 0, 6), up, 4.0, 1, 4.0),     # d == step hits
        (RECT, (0, 0, 6), up, 3.999, 0, 0.0),   # d > step misses
        (RECT, (1, 0, 6), up, 9.0, 0, 0.0),     # |x| == length_x/2 is outside
        (RECT, (0.999, 0, 6), up, 9.0, 1, 4.0),
        (DISC, (2, 0, 6), up, 9.0, 0, 0.0),     # on the rim (radius 2) is outside
        (DISC, (1.999, 0, 6), up, 9.0, 1, 4.0),
        (DISC, (0, 0, 12), -up, 9.0, 1, 2.0),   # crossed from above
        (DISC, (0, 0, 12), up, 9.0, 0, 0.0),    # moving away: d < 0
    ]
    for kind, p, d, step, hit, dd in cases:
        y = O.eval_fn("PLANE_ISEC", rows(kind, c, 2.0, 2.0, np.array([p], float), d, step))
        assert y[0, 0] == hit, (kind, p, d, step)
        assert y[0, 1] == np.float32(dd)
    flat = np.array([[1.0, 0.0, 0.0]])
    y = O.eval_fn("PLANE_ISEC", rows(DISC, c, 5.0, 0.0, np.array([[0, 0, 10.0]]), flat, 9.0))
    assert y[0, 0] == 0


def _run(O, P, targets, n, pos=(0, 0, 0), seed=3):
    medium = P.SimpleMediumProperties(30.0, 1e9, 1.35, 0.9, 0.2)
    src = P.PointlikeIsotropicEmitter(np.float32(pos), 0.0, n)
    return O.propagate(P.processing.sources_to_pod([src]), medium.to_pod(),
                       P.targets_to_pod(targets), seed, n_steps=15, mono_wl=450.0)


def test_disc_solid_angle_no_scatter(O, P):
    """Isotropic point at distance h below a disc of radius R: hit fraction
    (1 - h / sqrt(h^2 + R^2)) / 2; every hit lies on the disc at t = dist / c."""
    h, R, n = 5.0, 3.0, 1_000_000
    hits, nsim = _run(O, P, [P.circular_target((0, 0, h), R, 4)], n)
    p = (1 - h / np.hypot(h, R)) / 2
    assert nsim == n
    assert stats.binomtest(len(hits), n, p).pvalue > 1e-3
    assert np.all(hits["n_steps"] == 1) and np.all(hits["module_ix"] == 1)
    np.testing.assert_allclose(hits["pos"][:, 2], h, rtol=1e-6)
    assert np.all(np.hypot(hits["pos"][:, 0], hits["pos"][:, 1]) < R * (1 + 1e-6))
    dist = np.linalg.norm(hits["pos"].astype(np.float64), axis=1)
    np.testing.assert_allclose(hits["dist"], dist, rtol=1e-5)
    np.testing.assert_allclose(hits["time"], dist / 0.2, rtol=1e-5)


def test_rectangle_solid_angle_no_scatter(O, P):
    """On-axis solid angle of a 2a x 2b rectangle at distance h:
    4 arcsin(a b / sqrt((a^2 + h^2)(b^2 + h^2)))."""
    h, lx, ly, n = 4.0, 6.0, 2.0, 1_000_000
    hits, _ = _run(O, P, [P.rectangular_target((0, 0, -h), lx, ly, 9)], n)
    a, b = lx / 2, ly / 2
    omega = 4 * np.arcsin(a * b / np.sqrt((a * a + h * h) * (b * b + h * h)))
    assert stats.binomtest(len(hits), n, omega / (4 * np.pi)).pvalue > 1e-3
    np.testing.assert_allclose(hits["pos"][:, 2], -h, rtol=1e-6)
    assert np.all(np.abs(hits["pos"][:, 0]) < a * (1 + 1e-6))
    assert np.all(np.abs(hits["pos"][:, 1]) < b * (1 + 1e-6))


def test_mixed_list_first_in_index_order(O, P):
    """A beam through a disc (far, index 1) and a sphere (near, index 2)
    reports the disc; reversed order reports the sphere."""
    medium = P.SimpleMediumProperties(30.0, 1e9, 1.35, 0.9, 0.2)
    beam = P.PencilEmitter(np.zeros(3), np.float32([0, 0, 1]), 0.0, 0.0, 100)
    t = [P.circular_target((0, 0, 20), 1.0, 5), P.spherical_target((0, 0, 10), 0.3, 6)]
    for order, ix, dist in ((t, 1, 20.0), (t[::-1], 1, 9.7)):
        hits, _ = O.propagate(P.processing.sources_to_pod([beam]), medium.to_pod(),
                              P.targets_to_pod(order), 1, n_steps=15, mono_wl=400.0)
        assert len(hits) == 100
        assert np.all(hits["module_ix"] == ix)
        np.testing.assert_allclose(hits["dist"], dist, rtol=1e-6)
