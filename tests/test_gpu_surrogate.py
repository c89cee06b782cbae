"""Surrogate MLP kernels on the GPU against the CPU oracle (oracle/pp_mlp_oracle.c).

The oracle rounds to bf16 at the same points as the kernels and computes in
double, so the remaining differences are the kernels' fp32 accumulation and
the occasional one-ulp bf16 flip it causes in a hidden activation (bf16 has
an 8-bit significand: one ulp is 0.4%). The tolerances allow for such flips:
outputs within 2e-3, gradients within 3% per layer in L2 and cosine
similarity > 0.999; the measured agreement is far closer (~5e-6, ~1e-4)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S(P):
    from photonprop import surrogate
    return surrogate


def labels_for(rng, n, n_mc=1e5):
    lab = np.zeros((n, 66), np.float32)
    lab[:, 0] = n_mc
    for b in range(n):
        t = rng.gamma(2.0 + rng.random() * 3, 30.0 + 100 * rng.random(), size=rng.integers(0, 300))
        h, _ = np.histogram(t, bins=64, range=(0, 2000))
        lab[b, 2:] = h * rng.uniform(0.5, 1.0)
        lab[b, 1] = lab[b, 2:].sum()
    return lab


def feats(rng, n):
    return np.clip(rng.normal(0, 1, (n, 8)), -3, 3).astype(np.float32)


def layer_slices(dims):
    o = 0
    for i in range(len(dims) - 1):
        nw = dims[i + 1] * dims[i]
        yield f"W{i}", slice(o, o + nw)
        o += nw
        yield f"b{i}", slice(o, o + dims[i + 1])
        o += dims[i + 1]


def test_mfma_forward_matches_oracle(S, O):
    rng = np.random.default_rng(1)
    with S.MLP(max_batch=4096, seed=3) as m:
        p = m.get_params()
        for n in (1, 63, 300, 4096):  # partial and full 64-row tiles
            x = feats(rng, n)
            y = m.forward(x)
            yo = O.mlp_forward(m.dims, p, x, quant=True)
            err = np.abs(y - yo)
            print(f"n={n}: max |dy| {err.max():.2e}")
            assert np.all(err <= 2e-3 + 2e-3 * np.abs(yo))


@pytest.mark.parametrize("n_in,n_out", [(8, 10), (24, 32), (1, 10)])
def test_fused_inference_equals_layered(S, O, n_in, n_out):
    """The all-layer kernel and the per-layer GEMM kernels give the same bits
    (same MFMA order, zero padding and epilogue), on partial and full 64-row
    tiles, and both agree with the oracle."""
    rng = np.random.default_rng(11 + n_in)
    with S.MLP(n_in=n_in, n_out=n_out, max_batch=4096, seed=5) as m:
        p = m.get_params()
        p[:] = rng.normal(0, 0.08, p.size)  # bias terms nonzero too
        m.set_params(p)
        for n in (1, 63, 65, 300, 4096):
            x = np.clip(rng.normal(0, 1, (n, n_in)), -3, 3).astype(np.float32)
            m.set_fused(True)
            yf = m.forward(x)
            m.set_fused(False)
            yl = m.forward(x)
            assert np.array_equal(yf.view(np.uint32), yl.view(np.uint32)), n
            yo = O.mlp_forward(m.dims, p, x, quant=True)
            assert np.all(np.abs(yf - yo) <= 2e-3 + 2e-3 * np.abs(yo))


def test_fused_pair_evaluation_equals_layered(S):
    rng = np.random.default_rng(13)
    ne, nm = 300, 500
    ev = np.zeros((ne, 8), np.float32)
    ev[:, :3] = rng.uniform(-300, 300, (ne, 3))
    d = rng.normal(size=(ne, 3))
    ev[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    ev[:, 6] = np.exp(rng.uniform(np.log(200), np.log(5e5), ne))
    ev[:, 7] = rng.uniform(15, 25, ne)
    mods = rng.uniform(-300, 300, (nm, 3)).astype(np.float32)
    dom = S.default_setup().domain()
    with S.MLP(max_batch=1 << 16, seed=10) as m:
        out = {}
        for fused in (False, True, False, True):  # second pass of each is the timed one
            m.set_fused(fused)
            out[fused] = m.eval_pairs(ev, mods, dom, mixture=True)
        (hl, ml, fl, tl), (hf, mf, ff, tf) = out[False], out[True]
        print(f"{ne * nm} pairs: layered {tl:.3f} ms, fused {tf:.3f} ms")
        assert np.array_equal(hf, hl) and np.array_equal(mf, ml) and np.array_equal(ff, fl)


def test_gradients_match_oracle(S, O):
    rng = np.random.default_rng(2)
    with S.MLP(max_batch=1024, seed=4) as m:
        p = m.get_params()
        x = feats(rng, 700)
        lab = labels_for(rng, 700)
        loss, g = m.loss_grad(x, lab)
        assert np.array_equal(m.get_params(), p)  # no parameter change
        lo, go = O.mlp_loss_grad(m.dims, p, x, lab, quant=True)
        print(f"loss gpu {loss:.6g} oracle {lo:.6g}")
        assert loss == pytest.approx(lo, rel=1e-3)
        cos = float(g.astype(np.float64) @ go / np.linalg.norm(g) / np.linalg.norm(go))
        assert cos > 0.999, cos
        for name, sl in layer_slices(m.dims):
            rel = np.linalg.norm(g[sl] - go[sl]) / max(np.linalg.norm(go[sl]), 1e-30)
            print(f"{name}: rel L2 {rel:.2e}")
            assert rel < 3e-2, name


def test_adam_step_matches_oracle(S, O):
    rng = np.random.default_rng(3)
    with S.MLP(max_batch=512, seed=5) as m:
        p = m.get_params()
        x, lab = feats(rng, 256), labels_for(rng, 256)
        _, g = m.loss_grad(x, lab)
        m.adam_step(lr=1e-3, weight_decay=1e-4)
        p1 = m.get_params()
        z = np.zeros_like(p)
        po, _, _ = O.adam(p, g, z, z, 0, 1e-3, wd=1e-4)
        np.testing.assert_allclose(p1, po, rtol=2e-6, atol=1e-9)
        # the second step uses the bias correction of step 2
        _, g2 = m.loss_grad(x, lab)
        m.adam_step(lr=1e-3)
        _, m1, v1 = O.adam(p, g, z, z, 0, 1e-3, wd=1e-4)
        po2, _, _ = O.adam(po, g2, m1, v1, 1, 1e-3)
        np.testing.assert_allclose(m.get_params(), po2, rtol=5e-6, atol=1e-8)


def test_batch_4096_gradient_and_adam_step_match_oracle(S, O):
    """At the training batch of the surrogate's configuration (4096 rows):
    the loss and every layer's gradient against the oracle, then one Adam
    step on the kernel's gradient against the oracle's Adam."""
    rng = np.random.default_rng(7)
    with S.MLP(max_batch=4096, seed=11) as m:
        p = m.get_params()
        x, lab = feats(rng, 4096), labels_for(rng, 4096)
        loss, g = m.loss_grad(x, lab)
        lo, go = O.mlp_loss_grad(m.dims, p, x, lab, quant=True)
        print(f"batch 4096: loss gpu {loss:.6g} oracle {lo:.6g}")
        assert loss == pytest.approx(lo, rel=1e-3)
        cos = float(g.astype(np.float64) @ go / np.linalg.norm(g) / np.linalg.norm(go))
        assert cos > 0.999, cos
        for name, sl in layer_slices(m.dims):
            rel = np.linalg.norm(g[sl] - go[sl]) / max(np.linalg.norm(go[sl]), 1e-30)
            print(f"{name}: rel L2 {rel:.2e}")
            assert rel < 3e-2, name
        m.adam_step(lr=2e-3)
        z = np.zeros_like(p)
        po, _, _ = O.adam(p, g, z, z, 0, 2e-3)
        np.testing.assert_allclose(m.get_params(), po, rtol=2e-6, atol=1e-9)


def _philox_indices(O, seed, step, rows, n):
    ctr = np.stack([np.arange(rows), np.full(rows, step), np.full(rows, 0x5EEDBA7C),
                    np.zeros(rows)], axis=1).astype(np.uint32)
    key = np.tile(np.array([seed & 0xFFFFFFFF, seed >> 32], np.uint32), (rows, 1))
    r = O.philox(ctr, key).astype(np.uint64)
    u = (r[:, 0] << np.uint64(32)) | r[:, 1]
    return np.array([(int(v) * n) >> 64 for v in u], np.int64)


def test_graph_training_equals_eager_steps(S, O):
    """pp_mlp_train replays one captured step as a hipGraph; every replay
    gives the parameters of the same steps run eagerly, bit for bit."""
    rng = np.random.default_rng(4)
    n, rows, steps, seed = 3000, 256, 4, 0x1234567890
    x, lab = feats(rng, n), labels_for(rng, n)
    with S.MLP(max_batch=512, seed=6) as mg, S.MLP(max_batch=512, seed=6) as me:
        curve = mg.train(x, lab, steps, batch=rows, seed=seed, lr=1e-3)
        for s in range(steps):
            idx = _philox_indices(O, seed, s, rows, n)
            le, _ = me.loss_grad(x[idx], lab[idx])
            me.adam_step(lr=1e-3)
            assert le == pytest.approx(curve[s], rel=1e-6)
        assert np.array_equal(mg.get_params(), me.get_params())
        # a second call continues from the same state
        mg.train(x, lab, 2, batch=rows, seed=seed, lr=1e-3)
        for s in range(steps, steps + 2):
            idx = _philox_indices(O, seed, s, rows, n)
            me.loss_grad(x[idx], lab[idx])
            me.adam_step(lr=1e-3)
        assert np.array_equal(mg.get_params(), me.get_params())


def test_training_reduces_loss_on_a_known_law(S, O):
    """Labels drawn from a smooth law of the features: training with the
    graph loop removes most of the held-out loss in excess of the loss of the
    true law (the histogram term keeps an irreducible entropy part) and
    learns the rate."""
    rng = np.random.default_rng(5)
    n = 40_000
    x = feats(rng, n)
    lam_per = np.exp(-9.0 + 0.8 * x[:, 0] - 0.5 * x[:, 1] ** 2)
    lab = np.zeros((n, 66), np.float32)
    lab[:, 0] = 1e5
    lab[:, 1] = rng.poisson(1e5 * lam_per)
    k = 2.0 + x[:, 2] ** 2
    for b in range(n):
        c = int(lab[b, 1])
        if c:
            h, _ = np.histogram(rng.gamma(k[b], 50.0, c), bins=64, range=(0, 2000))
            lab[b, 2:] = h
    with S.MLP(max_batch=8192, seed=8) as m:
        l0 = m.eval_loss(x[-5000:], lab[-5000:])
        curve = m.train(x[:-5000], lab[:-5000], 400, batch=2048, seed=9, lr=2e-3)
        l1 = m.eval_loss(x[-5000:], lab[-5000:])
        print(f"held-out loss {l0:.4g} -> {l1:.4g}; batch loss {curve[0]:.4g} -> {curve[-1]:.4g}")
        assert np.all(np.isfinite(curve))
        # loss of the true law: rate lam_per, one gamma(k, 50 ns) component
        yt = np.zeros((5000, 10))
        yt[:, 0] = np.log(lam_per[-5000:])
        yt[:, 1] = 30.0
        yt[:, 4] = np.log(k[-5000:])
        yt[:, 7] = np.log(50.0) - 4.0
        l_true = O.loss_head(yt, lab[-5000:])[0].mean()
        print(f"true-law loss {l_true:.4g}")
        assert l1 - l_true < 0.2 * (l0 - l_true)
        y = m.forward(x[:2000])
        rel = np.abs(np.exp(y[:, 0]) / lam_per[:2000] - 1)
        print(f"median rel error of the rate {np.median(rel):.3f}")
        assert np.median(rel) < 0.1


def test_features_and_pair_evaluation(S, O):
    rng = np.random.default_rng(6)
    ne, nm = 37, 101
    ev = np.zeros((ne, 8), np.float32)
    ev[:, :3] = rng.uniform(-200, 200, (ne, 3))
    d = rng.normal(size=(ne, 3))
    ev[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    ev[:, 6] = np.exp(rng.uniform(np.log(50), np.log(2e6), ne))
    ev[:, 7] = rng.uniform(15, 25, ne)
    mods = rng.uniform(-200, 200, (nm, 3)).astype(np.float32)
    mods[0] = ev[0, :3] + np.float32([1.0, 1.0, 1.0])  # a near pair
    st = S.default_setup()
    dom = st.domain()
    with S.MLP(max_batch=1000, seed=10) as m:  # several chunks of pairs
        ce = np.repeat(ev, nm, axis=0)
        cm = np.tile(mods, (ne, 1))
        f = m.features(ce, cm)
        np.testing.assert_allclose(f, O.sur_features(ce, cm), atol=2e-5, rtol=1e-5)
        hits, mix, fb, ms = m.eval_pairs(ev, mods, dom, mixture=True)
        y = np.concatenate([m.forward(f[i:i + 1000]) for i in range(0, len(f), 1000)])
        r = np.linalg.norm(cm - ce[:, :3], axis=1)
        expo = 2 * np.log(0.3 / (2 * np.maximum(r, 0.3)))  # (R / 2r)^2
        ref = np.exp(np.clip(y[:, 0], -60, 10) + ce[:, 7] + expo).reshape(ne, nm)
        np.testing.assert_allclose(hits, ref, rtol=2e-5)
        r = r.reshape(ne, nm)
        e = ce[:, 6].reshape(ne, nm)
        expect_fb = (r < 5) | (r > 800) | (e < 100) | (e > 1e6)
        assert np.array_equal(fb, expect_fb) and fb[0, 0]
        np.testing.assert_allclose(mix[..., :3].sum(-1), 1.0, rtol=1e-5)
        assert np.all(mix[..., 3:] > 0) and ms > 0


def test_surrogate_gain_over_mc(S, P):
    """End to end at test scale: Monte Carlo labels, training on the GPU, and
    the full Monte Carlo as judge on held-out events in the 1000-module array.
    The network's expected hits agree with the judge on well-measured pairs,
    it evaluates pairs >= 100x faster than the Monte Carlo, and the Monte
    Carlo alone, given the network's device time, is far less accurate."""
    rng = np.random.default_rng(12)
    st = S.default_setup()
    _, dev = P.make_hit_buffers(capacity=1 << 25)
    with dev:
        c, lab, gs = S.generate_training_data(st, 200_000, seed=2, dev=dev, log=None)
        mlp, info = S.train_surrogate(st, c, lab, steps=24_000, lr_stages=4, log=None)
        with mlp:
            model = S.SurrogateModel(mlp, st)
            targets = P.make_detector_grid(10, 10, 50.0, 10, 10.0, radius=0.3)
            mxyz = np.array([t.shape.position for t in targets], np.float32)
            ev = S.array_events(rng, 20, mxyz, setup=st)
            res, mc, sur = S.judge_events(model, ev, targets, max_photons=1e9, dev=dev)
            print("judge:", {k: v for k, v in res.items() if k != "strata"})
            for row in res["strata"]["by_distance"] + res["strata"]["by_mc_hits"]:
                print("stratum:", row)
            assert res["pairs"] >= 50
            assert res["mean_rel_err"] < 0.05
            # over every in-domain pair the summed expectations agree
            assert abs(res["strata"]["all"]["sum_ratio"] - 1.0) < 0.05
            # throughput: 2e6 pairs through the batched network vs the Monte Carlo
            big = S.array_events(rng, 2000, mxyz, setup=st)
            _, _, fb, ms = mlp.eval_pairs(big, mxyz, model.domain)
            sur_rate = big.shape[0] * len(targets) / (ms * 1e-3)
            mc_rate = ev.shape[0] * len(targets) / (mc["kernel_ms"] * 1e-3)
            print(f"pair evaluations/s: surrogate {sur_rate:.3g}, MC {mc_rate:.3g}")
            assert sur_rate > 100 * mc_rate
            # the Monte Carlo alone with the surrogate's device time for these events
            t_sur = res["surrogate_ms"]
            budget = mc["photons"] / mc["kernel_ms"] * t_sur / ev.shape[0]
            cheap = S.mc_expected_hits(st, ev, targets, max_photons=max(budget, 64), seed=99,
                                       dev=dev)
            sel = (~sur["fallback"]) & (mc["n_eff"] >= 2500)
            err_cheap = np.mean(np.abs(cheap["hits"][sel] / mc["hits"][sel] - 1))
            print(f"equal-time MC ({budget:.3g} photons/event): mean rel err {err_cheap:.3f}; "
                  f"surrogate {res['mean_rel_err']:.3f}")
            assert err_cheap > 3 * res["mean_rel_err"]


def test_device_labels_equal_host_labels(S, P):
    """pp_sur_hit_labels (hits reduced to labels on the device) gives the
    labels the host reduction of the copied hits gives, for the same samples
    and seeds. A batch that overflows the hit buffer is run as halves (other
    random streams: the sources' indices change) and gives the same labels
    within the Monte Carlo's statistics."""
    st = S.default_setup()
    _, dev = P.make_hit_buffers(capacity=1 << 22)
    with dev:
        c1, l1, g1 = S.generate_training_data(st, 600, seed=4, batch_sources=256, dev=dev,
                                              log=None, device_labels=True)
        c2, l2, g2 = S.generate_training_data(st, 600, seed=4, batch_sources=256, dev=dev,
                                              log=None, device_labels=False)
    assert np.array_equal(c1, c2) and g1["photons"] == g2["photons"]
    assert np.sum(l1[:, 1]) > 100
    np.testing.assert_allclose(l1, l2, rtol=2e-6, atol=1e-6)
    _, small = P.make_hit_buffers(capacity=1 << 14)
    with small:
        c3, l3, _ = S.generate_training_data(st, 600, seed=4, batch_sources=256, dev=small,
                                             log=None, device_labels=True)
    assert np.array_equal(c3, c1) and np.array_equal(l3[:, 0], l1[:, 0])
    w1, w3 = l1[:, 1].sum(), l3[:, 1].sum()
    print(f"weight sums: one batch {w1:.1f}, split batches {w3:.1f}")
    assert abs(w3 / w1 - 1) < 0.05
