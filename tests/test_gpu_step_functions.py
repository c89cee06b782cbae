"""GPU step functions (pp_eval_fn, the kernel's own inlined code) against the
CPU oracle on identical inputs.

The oracle evaluates each function exactly (double precision, one rounding
of the Float32 output). Tolerance: 4 FP32 ulp of the output's scale plus the
function's own conditioning, i.e. the largest change of the exact output
when any one Float32 input moves by 1 ulp (kappa). Unit-vector components
are measured at scale 1 (ulp(1) = 2^-23); scalars at their magnitude.
Philox and the spectrum lookup are bit-exact. Rejection samplers (gamma)
are compared draw by draw; decisions flipped by a near-tie are counted
apart from the 4-ulp bound."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(2024)
ULP1 = 2.0**-23


def ulp(x):
    x = np.maximum(np.abs(np.asarray(x, dtype=np.float64)), 2.0**-126)
    return 2.0 ** (np.floor(np.log2(x)) - 23)


def report(name, err):
    print(f"{name}: max {np.max(err):.2f} ulp, p99.9 {np.quantile(err, 0.999):.2f} ulp")


def unit(n):
    v = RNG.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def both(P, O, dev, fn, x, **kw):
    g = P._lib.eval_fn(dev, fn, x).astype(np.float64)
    r = O.eval_fn(fn, x, **kw).astype(np.float64)
    return g, r


def kappa(O, fn, x, **kw):
    """Largest change of the oracle output when one input moves by 1 ulp."""
    x = np.asarray(x, dtype=np.float32)
    base = O.eval_fn(fn, x, **kw).astype(np.float64)
    k = np.zeros_like(base)
    for j in range(x.shape[1]):
        for to in (np.float32(np.inf), np.float32(-np.inf)):
            xp = x.copy()
            xp[:, j] = np.nextafter(x[:, j], to)
            k = np.maximum(k, np.abs(O.eval_fn(fn, xp, **kw).astype(np.float64) - base))
    return k


def check_vec(P, O, dev, fn, x, name, scale=None, **kw):
    """Assert |gpu - oracle| <= 4 ulp(scale) + kappa; print the error in ulp."""
    g, r = both(P, O, dev, fn, x, **kw)
    k = kappa(O, fn, x, **kw)
    u = ULP1 if scale is None else ulp(scale)
    err = np.abs(g - r) / u
    excess = (np.abs(g - r) - k) / u
    report(name, err)
    print(f"{name}: max over kappa {np.max(excess):.2f} ulp, "
          f"within 4 ulp without kappa {np.mean(err <= 4):.6f}")
    assert np.max(excess) <= 4
    return g, r


def test_philox_bit_exact(P, O, dev):
    w = RNG.integers(0, 2**32, (100_000, 6), dtype=np.uint64).astype(np.uint32)
    g = P._lib.eval_fn(dev, "PHILOX", w.view(np.float32)).view(np.uint32)
    r = O.eval_fn("PHILOX", w.view(np.float32)).view(np.uint32)
    np.testing.assert_array_equal(g, r)


def test_isotropic_dir(P, O, dev):
    x = RNG.random((200_000, 2), dtype=np.float32)
    check_vec(P, O, dev, "ISOTROPIC_DIR", x, "isotropic_dir")


def test_rot_from_ez(P, O, dev):
    n = 200_000
    x = np.concatenate([unit(n), unit(n)], axis=1)
    x[:10, :3] = [0, 0, 1]
    x[10:20, :3] = [0, 0, -1]
    check_vec(P, O, dev, "ROT_FROM_EZ", x, "rot_from_ez")


@pytest.mark.parametrize("g_", [0.5, 0.95, 0.99])
def test_hg_cos(P, O, dev, g_):
    u = RNG.random(200_000, dtype=np.float32)
    x = np.stack([np.full_like(u, g_), u], 1)
    check_vec(P, O, dev, "HG_COS", x, f"hg_cos g={g_}")


@pytest.mark.parametrize("b", [0.835, 0.3])
def test_es_cos(P, O, dev, b):
    u = RNG.random(200_000, dtype=np.float32)
    x = np.stack([np.full_like(u, b), u], 1)
    check_vec(P, O, dev, "ES_COS", x, f"es_cos b={b}")


@pytest.mark.parametrize("g_", [0.9, 0.5])
def test_sl_cos(P, O, dev, g_):
    u = RNG.random(200_000, dtype=np.float32)
    x = np.stack([np.full_like(u, g_), u], 1)
    check_vec(P, O, dev, "SL_COS", x, f"sl_cos g={g_}")


def test_consecutive_evaluations_see_their_own_inputs(P, O, dev):
    """Regression: inputs are uploaded on the context's stream, so a call
    never reads the previous call's inputs left in a reused allocation."""
    u = RNG.random(500_000, dtype=np.float32)
    for g_ in (0.9, 0.2, 0.7, 0.45):
        x = np.stack([np.full_like(u, g_), u], 1)
        g, r = both(P, O, dev, "SL_COS", x)
        assert np.max(np.abs(g - r)) < 1e-5, g_


def test_update_direction(P, O, dev):
    n = 200_000
    x = np.concatenate([unit(n), RNG.uniform(-1, 1, (n, 1)).astype(np.float32),
                        RNG.random((n, 1), dtype=np.float32)], 1)
    check_vec(P, O, dev, "UPDATE_DIR", x, "update_direction")


def test_sphere_intersection(P, O, dev):
    n = 200_000
    c = RNG.uniform(-20, 20, (n, 3)).astype(np.float32)
    r = RNG.uniform(0.1, 1.0, n).astype(np.float32)
    p = RNG.uniform(-30, 30, (n, 3)).astype(np.float32)
    aim = c + RNG.normal(0, 0.5, (n, 3)).astype(np.float32) - p
    d = (aim / np.linalg.norm(aim, axis=1, keepdims=True)).astype(np.float32)
    step = RNG.exponential(40.0, n).astype(np.float32)
    x = np.concatenate([c, r[:, None], p, d, step[:, None]], 1)
    g, ref = both(P, O, dev, "SPHERE_ISEC", x)
    k = kappa(O, "SPHERE_ISEC", x)
    # exact geometry: discriminant b = r^2 - |dp - a dir|^2, near root
    pd, cd, dd = p.astype(np.float64), c.astype(np.float64), d.astype(np.float64)
    dp = pd - cd
    a = np.sum(dd * dp, 1)
    q = dp - a[:, None] * dd
    b = r.astype(np.float64) ** 2 - np.sum(q * q, 1)
    root = -a - np.sqrt(np.abs(b))
    ndp = np.linalg.norm(dp, axis=1)
    eps = 16 * 2.0**-24
    # decisions may differ only for rays within 16 float ulps of a boundary:
    # grazing (b ~ 0), root at the step end or at the start point
    edge = (k[:, 0] > 0) | (np.abs(b) < eps * (2 * ndp * np.linalg.norm(q, axis=1) + r**2)) | (
        np.abs(root - step) < eps * (np.abs(a) + step + ndp)) | (np.abs(root) < eps * ndp)
    print(f"sphere: {edge.mean():.2e} of rows within the decision band")
    assert np.all((g[:, 0] == ref[:, 0]) | edge)
    h = (g[:, 0] == 1) & (ref[:, 0] == 1) & ~edge
    assert h.sum() > 10_000
    err = np.abs(g[h, 1] - ref[h, 1]) / ulp(np.abs(a[h]))
    excess = (np.abs(g[h, 1] - ref[h, 1]) - k[h, 1]) / ulp(np.abs(a[h]))
    report("sphere d (ulp of |a|)", err)
    print(f"sphere d: max over kappa {np.max(excess):.2f} ulp")
    assert np.max(excess) <= 4


@pytest.mark.parametrize("kind", [1, 2], ids=["rectangle", "disc"])
def test_plane_intersection(P, O, dev, kind):
    """check_intersection(::Rectangular / ::Circular): decisions agree except
    for rays within 16 float ulps of a decision boundary (the shape's edge,
    d = 0, d = step); d within 4 ulp + kappa."""
    n = 200_000
    c = RNG.uniform(-20, 20, (n, 3)).astype(np.float32)
    a = RNG.uniform(0.5, 3.0, n).astype(np.float32)
    b = RNG.uniform(0.5, 3.0, n).astype(np.float32)
    cross = c + np.column_stack([RNG.uniform(-1.3, 1.3, (n, 2)) * np.column_stack([a, b]) *
                                 (0.5 if kind == 1 else 1.0), np.zeros(n)]).astype(np.float32)
    d = unit(n)
    dist = RNG.uniform(-5, 40, n).astype(np.float32)
    p = (cross - d * dist[:, None]).astype(np.float32)
    step = RNG.exponential(30.0, n).astype(np.float32)
    x = np.zeros((n, 13), np.float32)
    x[:, 0] = np.array([kind], np.uint32).view(np.float32)[0]
    x[:, 1:4], x[:, 4], x[:, 5], x[:, 6:9], x[:, 9:12], x[:, 12] = c, a, b, p, d, step
    g, ref = both(P, O, dev, "PLANE_ISEC", x)
    # exact crossing of the float inputs
    cd, pd, dd = c.astype(np.float64), p.astype(np.float64), d.astype(np.float64)
    t = (cd[:, 2] - pd[:, 2]) / dd[:, 2]
    ex = pd[:, 0] + dd[:, 0] * t - cd[:, 0]
    ey = pd[:, 1] + dd[:, 1] * t - cd[:, 1]
    scale = np.abs(pd[:, :2]).max(1) + np.abs(cd[:, :2]).max(1) + np.abs(t) + a
    eps = 16 * 2.0**-24
    if kind == 1:
        margin = np.minimum(np.abs(np.abs(ex) - a / 2.0), np.abs(np.abs(ey) - b / 2.0))
    else:
        rr = np.float32(a) * np.float32(a)
        margin = np.abs(ex * ex + ey * ey - rr) / (2 * np.sqrt(rr) + 1e-30)
    tz = np.abs(pd[:, 2]) + np.abs(cd[:, 2])
    edge = (margin < eps * scale) | (np.abs(t) < eps * tz / np.abs(dd[:, 2])) | (
        np.abs(t - step) < eps * (step + tz / np.abs(dd[:, 2])))
    print(f"plane kind {kind}: {edge.mean():.2e} of rows within the decision band, "
          f"hit rate {ref[:, 0].mean():.3f}")
    assert np.all((g[:, 0] == ref[:, 0]) | edge)
    h = (g[:, 0] == 1) & (ref[:, 0] == 1)
    assert h.sum() > 20_000
    err = np.abs(g[h, 1] - ref[h, 1]) / ulp(ref[h, 1])
    report(f"plane d kind {kind}", err)
    assert err.max() <= 4


def test_step_length(P, O, dev):
    u = ((RNG.integers(0, 2**24, 200_000) + 1) * 2.0**-24).astype(np.float32)
    x = np.stack([u, RNG.uniform(5, 100, len(u)).astype(np.float32)], 1)
    g, r = both(P, O, dev, "STEP_LEN", x)
    nz = r[:, 0] > 0
    err = np.abs(g[nz] - r[nz]) / ulp(r[nz])
    report("step_length", err)
    assert err.max() <= 4
    assert np.all(g[~nz] == 0)


def test_cascadia_medium_values(P, O, dev):
    m = P.CascadiaMediumProperties()
    P.processing.stage(dev, [P.spherical_target((0, 0, 0), 1.0, 1)], m, P.Monochromatic(450.0))
    wl = np.linspace(290.0, 810.0, 20_001, dtype=np.float32)
    g, r = both(P, O, dev, "CASCADIA", wl[:, None], medium=m.to_pod())
    for k, name in enumerate(("sca_len", "c_grp", "n_phase")):
        err = np.abs(g[:, k] - r[:, k]) / ulp(r[:, k])
        report(name, err)
        assert err.max() <= 4, name


def test_ice_medium_values(P, O, dev):
    m = P.make_homogenous_clearice_properties(sca_scale=1.3)
    P.processing.stage(dev, [P.spherical_target((0, 0, 0), 1.0, 1)], m, P.Monochromatic(450.0))
    wl = np.linspace(290.0, 810.0, 20_001, dtype=np.float32)
    g, r = both(P, O, dev, "CASCADIA", wl[:, None], medium=m.to_pod())
    for k, name in enumerate(("sca_len", "c_grp", "n_phase")):
        err = np.abs(g[:, k] - r[:, k]) / ulp(r[:, k])
        report(name, err)
        assert err.max() <= 4, name


def test_cherenkov_track_direction(P, O, dev):
    """cos(theta) is a Float32 variable before sin = sqrt(1 - cos^2) (as in
    the reference): one ulp of that stored cosine is added to kappa, since
    next to cos = 1 it moves sin(theta) by far more than 4 ulp."""
    x = RNG.random((200_000, 2), dtype=np.float32)
    g, r = both(P, O, dev, "TRACK_DIR", x)
    k = kappa(O, "TRACK_DIR", x)
    ct = r[:, 2]
    st = np.sqrt(np.maximum(1 - ct * ct, 0))
    for to in (np.inf, -np.inf):
        c2 = np.nextafter(ct.astype(np.float32), np.float32(to)).astype(np.float64)
        ds = np.abs(np.sqrt(np.maximum(1 - c2 * c2, 0)) - st)
        k[:, 0] = np.maximum(k[:, 0], ds + np.abs(c2 - ct))
        k[:, 1] = np.maximum(k[:, 1], ds + np.abs(c2 - ct))
        k[:, 2] = np.maximum(k[:, 2], np.abs(c2 - ct))
    err = np.abs(g - r) / ULP1
    excess = (np.abs(g - r) - k) / ULP1
    report("track_dir", err)
    print(f"track_dir: max over kappa {np.max(excess):.2f} ulp")
    assert np.max(excess) <= 4


def test_gamma_sampler_same_draws(P, O, dev):
    """Marsaglia-Tsang draws from the same Philox blocks. Both evaluators also
    return the accepted normal variate x = sqrt(-2 log u1) cos(2 pi u2): the
    kernel's, from Float32 log, sqrt, sincospi and product, agrees with the
    oracle's (double, rounded once) within 4 ulp. Given the variates, the
    draws agree within 4 ulp + kappa, kappa the change of d (1 + c x)^3 scale
    (boost) that the variates' difference plus 1 ulp of x causes, and that of
    c = 1/sqrt(9 d) rounded to Float32 (the reference's arithmetic is Float32
    throughout): near x = -1/c the cube amplifies either without bound. A
    decision flipped by the last bits of a near-tie accepts another variate;
    flips are counted separately and must stay rare."""
    for shape, scale in ((7.837, 1 / 0.632), (1.776, 1.0), (0.164, 1.0)):
        n = 50_000
        rows = np.zeros((n, 7), dtype=np.float32)
        rows[:, 0] = shape
        rows[:, 1] = scale
        w = rows.view(np.uint32)
        w[:, 2] = 99
        w[:, 4] = np.arange(n)
        g, r = both(P, O, dev, "GAMMA", rows)
        xg, xo = g[:, 1].astype(np.float64), r[:, 1].astype(np.float64)
        flip = np.abs(xg - xo) > 1e-4 * np.maximum(np.abs(xo), 1.0)
        same = ~flip
        dx = np.abs(xg - xo)
        x_err = dx[same] / ulp(r[same, 1])
        report(f"gamma k={shape} normal variate", x_err)
        assert np.max(x_err) <= 4
        k = shape + 1.0 if shape < 1.0 else shape
        d = k - 1.0 / 3.0
        c = 1.0 / np.sqrt(9.0 * d)
        v1 = np.maximum(1.0 + c * xo, 1e-30)
        eps32 = 2.0 ** -23
        kappa = np.abs(r[:, 0]) * 3.0 * c / v1 * (dx + ulp(r[:, 1]) + np.abs(xo) * eps32)
        if shape < 1.0:
            # the boost u^(1/shape) = exp(L), L = ln(u) / shape: the Float32
            # exponent 1/shape carries 0.5 ulp, which moves the boost by
            # |L| 2^-24 relative (counted twice); L from the draw itself
            boost = r[:, 0] / (d * v1 ** 3 * scale)
            kappa += np.abs(r[:, 0]) * np.abs(np.log(np.maximum(boost, 1e-38))) * eps32
        err = np.abs(g[same, 0] - r[same, 0]) / ulp(r[same, 0])
        excess = (np.abs(g[same, 0] - r[same, 0]) - kappa[same]) / ulp(r[same, 0])
        report(f"gamma k={shape}", err)
        print(f"gamma k={shape}: max over kappa {np.max(excess):.2f} ulp; "
              f"{flip.sum()} of {n} draws with a flipped decision")
        assert np.max(excess) <= 4
        assert flip.sum() <= 5


@pytest.mark.parametrize("g_", [0.95, 0.5])
def test_mixed_hges_kernel_sampler(P, O, dev, g_):
    """The branch-free HG/ES mixture the propagation kernel calls (both
    samplers share one Newton reciprocal; it takes the random words, u =
    k 2^-24) against the oracle's MixedHGES (branch choice, then the cosine),
    with the ends and the choice threshold among the draws."""
    n = 200_000
    x = np.zeros((n, 5), np.float32)
    x[:, 0] = g_
    x[:, 1] = 0.835
    x[:, 2] = 0.8
    x[:, 3] = RNG.integers(0, 1 << 24, n).astype(np.float32) * np.float32(2.0**-24)
    x[:, 4] = RNG.integers(0, 1 << 24, n).astype(np.float32) * np.float32(2.0**-24)
    t = np.float32(np.ceil(np.float64(np.float32(0.8)) * 2.0**24)) * np.float32(2.0**-24)
    x[:6, 3] = [0.0, 1.0 - 2.0**-24, t - np.float32(2.0**-24), t, t + np.float32(2.0**-24), 0.5]
    x[:6, 4] = [0.0, 1.0 - 2.0**-24, 0.5, 2.0**-24, 0.25, 0.75]
    g, r = check_vec(P, O, dev, "MIXED_HGES", x, f"mixed_hges g={g_}")
    hg = x[:, 3] < x[:, 2]
    assert 0.79 < hg.mean() < 0.81
    # the ES branch has E[cos] = 0, the HG branch E[cos] = g
    assert abs(g[~hg, 0].mean()) < 0.01
    assert abs(g[hg, 0].mean() - g_) < 0.01


def test_spectrum_lookup_bit_exact(P, O, dev):
    m = P.CascadiaMediumProperties()
    tab = P.workloads.acceptance_weighted_spectrum(m)
    for mapping in (0, 1):
        spec = P.DeviceSpectralDist(tab.table, tab.normalization, mapping)
        P.processing.stage(dev, [P.spherical_target((0, 0, 0), 1.0, 1)], m, spec)
        u = RNG.random((100_000, 1), dtype=np.float32)
        u[:4, 0] = [0.0, 1.0 - 2.0**-24, 0.5, 2.0**-24]
        g = P._lib.eval_fn(dev, "SPECTRUM", u)
        r = O.eval_fn("SPECTRUM", u, spectrum=(0.0, tab.table, mapping))
        np.testing.assert_array_equal(g.view(np.uint32), r.view(np.uint32))
