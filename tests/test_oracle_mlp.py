"""The CPU oracle of the surrogate MLP, pinned by mathematics: bf16 rounding
against torch's conversion, digamma against scipy, every analytic gradient
(loss head, network) against central finite differences, Adam's first step in
closed form, and the pair features in closed form."""
import numpy as np
import pytest
from scipy import special

DIMS = [8, 16, 16, 10]


def labels_for(rng, n, n_mc=1e5):
    lab = np.zeros((n, 66), np.float32)
    lab[:, 0] = n_mc
    t = rng.gamma(3.0, 40.0, size=(n, 200))
    for b in range(n):
        h, _ = np.histogram(t[b][: rng.integers(0, 200)], bins=64, range=(0, 2000))
        lab[b, 2:] = h * rng.uniform(0.5, 1.0)
        lab[b, 1] = lab[b, 2:].sum()
    return lab


def rand_params(rng, dims, scale=0.3):
    n = sum(dims[i + 1] * dims[i] + dims[i + 1] for i in range(len(dims) - 1))
    p = rng.normal(0, scale, n).astype(np.float32)
    p[-10:] = [-9.0, 0.1, -0.2, 0.0, 0.7, 1.2, 0.3, -0.5, 0.4, 1.5]  # output biases
    return p


def test_bf16_rounding_matches_torch(O):
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-30, 100), [0.0, -0.0, 1e38],
                        np.float32(1.0) + np.arange(1, 9) * np.float32(2.0 ** -9)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(O.bf16(x).view(np.uint32), ref.view(np.uint32))
    for v in x[:50]:
        assert O.lib().ppo_bf16(float(v)) == np.float32(ref[list(x).index(v)])


def test_digamma(O):
    for v in [0.02, 0.1, 0.5, 1.0, 2.5, 5.9, 6.0, 17.0, 300.0]:
        assert O.lib().ppo_digamma(v) == pytest.approx(special.digamma(v), rel=1e-10, abs=1e-12)


def test_loss_head_gradient_finite_differences(O):
    rng = np.random.default_rng(1)
    n = 6
    y = rng.normal(0, 1, (n, 10))
    y[:, 0] = rng.uniform(-9, -5, n)
    y[:, 4:7] = rng.uniform(-1, 2, (n, 3))
    y[:, 7:10] = rng.uniform(-2, 2, (n, 3))
    lab = labels_for(rng, n)
    ls, dy = O.loss_head(y, lab)
    h = 1e-6
    for b in range(n):
        for o in range(10):
            yp, ym = y.copy(), y.copy()
            yp[b, o] += h
            ym[b, o] -= h
            fd = (O.loss_head(yp, lab)[0][b] - O.loss_head(ym, lab)[0][b]) / (2 * h)
            assert dy[b, o] == pytest.approx(fd, rel=1e-5, abs=1e-5 * max(1.0, abs(ls[b])))


def test_poisson_term_minimum_and_mixture_normalisation(O):
    lab = np.zeros((1, 66), np.float32)
    lab[0, 0], lab[0, 1] = 1e6, 250.0
    y = np.zeros((1, 10))
    y[0, 0] = np.log(250.0 / 1e6)  # lambda = w: zero gradient of the Poisson term
    _, dy = O.loss_head(y, lab)
    assert abs(dy[0, 0]) < 1e-9
    # one component (logit +30): the binned density integrates to ~1 and the
    # NLL of a histogram equals -sum h log pdf, the gamma log density
    k, lth = 3.0, np.log(100.0)
    y = np.array([[-5.0, 30.0, 0.0, 0.0, np.log(k), 0.0, 0.0, lth - 4.0, 0.0, 0.0]])
    x = (np.arange(64) + 0.5) * 2000 / 64
    lab = np.zeros((1, 66), np.float32)
    lab[0, 0] = 1.0
    lab[0, 2 + 10] = 2.0
    ls, _ = O.loss_head(y, lab)
    pois = 1.0 * np.exp(-5.0)  # lambda - w log lambda with n_mc = 1, w = 0
    logpdf = (k - 1) * np.log(x[10]) - x[10] / 100.0 - k * np.log(100.0) - special.gammaln(k)
    assert ls[0] == pytest.approx(pois - 2.0 * logpdf, rel=1e-9)
    pdf = np.exp((k - 1) * np.log(x) - x / 100.0 - k * np.log(100.0) - special.gammaln(k))
    assert pdf.sum() * 2000 / 64 == pytest.approx(1.0, abs=2e-3)


@pytest.mark.parametrize("quant", [False])
def test_network_gradient_finite_differences(O, quant):
    rng = np.random.default_rng(2)
    p = rand_params(rng, DIMS).astype(np.float64)
    x = rng.normal(0, 1, (5, 8)).astype(np.float32)
    lab = labels_for(rng, 5)
    loss, g = O.mlp_loss_grad(DIMS, p.astype(np.float32), x, lab, quant=quant)
    p32 = p.astype(np.float32)
    idx = rng.choice(p.size, 60, replace=False)
    for i in idx:
        h = max(1e-3 * abs(float(p32[i])), 1e-3)
        pp, pm = p32.copy(), p32.copy()
        pp[i] += np.float32(h)
        pm[i] -= np.float32(h)
        step = float(pp[i]) - float(pm[i])
        fd = (O.mlp_loss_grad(DIMS, pp, x, lab, quant=False)[0] -
              O.mlp_loss_grad(DIMS, pm, x, lab, quant=False)[0]) / step
        assert g[i] == pytest.approx(fd, rel=2e-3, abs=2e-4 * max(1.0, abs(loss))), i


def test_quantised_gradient_close_to_exact(O):
    rng = np.random.default_rng(3)
    dims = [8, 64, 64, 10]
    p = rand_params(rng, dims, 0.15)
    x = rng.normal(0, 1, (64, 8)).astype(np.float32)
    lab = labels_for(rng, 64)
    l0, g0 = O.mlp_loss_grad(dims, p, x, lab, quant=False)
    l1, g1 = O.mlp_loss_grad(dims, p, x, lab, quant=True)
    assert l1 == pytest.approx(l0, rel=2e-2)
    cos = g0 @ g1 / np.linalg.norm(g0) / np.linalg.norm(g1)
    assert cos > 0.99


def test_adam_first_step_closed_form(O):
    rng = np.random.default_rng(4)
    p = rng.normal(0, 1, 100).astype(np.float32)
    g = rng.normal(0, 1e-2, 100).astype(np.float32)
    z = np.zeros(100, np.float32)
    p1, m1, v1 = O.adam(p, g, z, z, 0, 1e-3, eps=1e-8)
    # first step: m_hat = g, v_hat = g^2, update = lr * g / (|g| + eps)
    np.testing.assert_allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(m1, 0.1 * g, rtol=1e-6)
    np.testing.assert_allclose(v1, 0.001 * g * g, rtol=1e-5)


def test_pair_features_closed_form(O):
    c = np.array([[0, 0, 0, 0, 0, 1, 1e4, 0], [1, 2, 3, 1, 0, 0, 100.0, 0]], np.float32)
    m = np.array([[0, 0, 10], [1, 2 + 30, 3]], np.float32)
    f = O.sur_features(c, m)
    np.testing.assert_allclose(f[0], [(1 - 1.2) / 0.7, 1, 0, 0, 0.1, 0, 0, (1 - 1.2) / 0.7],
                               atol=1e-6)
    lr = (np.log10(30) - 1.2) / 0.7
    np.testing.assert_allclose(f[1], [lr, 0, 1, (2 - 4) / 1.2, 0, 0.3, 0.5 * np.log10(31), 0],
                               atol=1e-6)
