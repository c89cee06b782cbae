"""PMT-level detection on the GPU (pp_pom_pmt_hits) against the oracle, and
end to end after a propagation with POM modules."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pom(P):
    from photonprop import pom
    return pom


def test_pmt_ids_match_oracle(O, pom):
    acc = pom.default_acceptance()
    rng = np.random.default_rng(3)
    n = 1_000_000
    rel = rng.normal(size=(n, 3))
    rel /= np.linalg.norm(rel, axis=1, keepdims=True)
    wl = rng.uniform(280, 820, n)
    pid = rng.integers(0, 1 << 40, n).astype(np.uint64)
    src = rng.integers(0, 500, n).astype(np.int32)
    orient = np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]])
    m = pom.POM(np.zeros(3), 1)
    g = pom._pmt_ids(rel, wl, acc, orient, pid, src, 77, 0)
    o, psum = O.pom_pmt_hits(rel, wl, pid, src, 77, acc.wavelengths, acc.acc_1, acc.acc_2,
                             acc.sigma_1, acc.sigma_2, orientation=orient)
    diff = np.nonzero(g != o)[0]
    print(f"{len(diff)} of {n} hits differ; detected {np.mean(g > 0):.4f}")
    # a decision may flip only where a uniform lies within float rounding of a
    # threshold (u_hit vs sum p, or a cumulative PMT weight)
    assert len(diff) <= 20
    np.testing.assert_allclose(pom.get_pmt_positions(m, orient), O.pom_pmt_dirs(orient),
                               atol=2e-7)
    # a photon's PMT does not depend on the batch it is in
    g2 = pom._pmt_ids(rel[::7], wl[::7], acc, orient, pid[::7], src[::7], 77, 0)
    assert np.array_equal(g2, g[::7])


def test_pmt_hits_after_propagation(P, pom):
    """README geometry with a POM: photons -> PMT hits; the detected fraction
    equals the mean summed PMT probability of the photons' positions and
    wavelengths, and hits concentrate on PMTs facing the source."""
    target = pom.POM(np.float32([0, 0, 10]), 1)
    src = P.PointlikeIsotropicEmitter(np.float32([0, 0, 0]), 0.0, 20_000_000)
    medium = P.CascadiaMediumProperties()
    spec = P.workloads.acceptance_weighted_spectrum(medium)
    setup = P.PhotonPropSetup([src], [target], medium, spec, 1)
    hbc, hbg = P.make_hit_buffers(capacity=1 << 22)
    with hbg:
        df = P.propagate_photons(setup, hbc, hbg, 15, vector_columns=False, with_ids=True)
    assert len(df) > 5000
    hits = pom.make_hits_from_photons(df, [target], seed=3)
    frac = len(hits) / len(df)
    acc = target.acceptance
    print(f"{len(df)} photons on the module, {len(hits)} PMT hits ({frac:.4f})")
    assert 0.005 < frac < 0.2
    assert hits["pmt_id"].between(1, 16).all()
    # PMT facing the source (-z side): the most hit PMTs point towards -z
    d = pom.get_pmt_positions(target)
    counts = np.bincount(hits["pmt_id"], minlength=17)[1:]
    top = np.argsort(-counts)[:3]
    assert np.all(d[top, 2] < 0), (counts, d[:, 2])
    pe = pom.calc_pe_weight(hits.copy(), [target])
    assert np.all(pe["qe_weight"] >= 0) and np.all(pe["qe_weight"] <= 1.0 + 1e-9)
    del acc


def test_pmt_count_table(P, tmp_path):
    """The surrogate-table writer: Arrow rows with 16 per-PMT weighted counts
    from propagation -> PMT hits -> QE weights."""
    import pyarrow as pa
    from photonprop import tables
    out = tmp_path / "tab.arrow"
    rows = tables.pmt_count_table(out, 3, mode="em_shower", energy=300.0, dist=15.0,
                                  randomize=False, seed=5, log=print)
    with pa.memory_map(str(out)) as f:
        t = pa.ipc.open_file(f).read_all()
    assert t.num_rows == 3
    assert set(t.column_names) == {"sim_ix", "pos", "dir", "pos_t", "pos_phi", "dir_t",
                                   "dir_phi", "hits", "energy", "dist"}
    hits = np.array(t.column("hits").to_pylist())
    assert hits.shape == (3, 16) and np.all(hits >= 0) and hits.sum() > 0
    np.testing.assert_allclose(np.linalg.norm(np.array(rows["pos"]), axis=1), 15.0, rtol=1e-6)


def test_photon_table_hdf5(P, tmp_path):
    """The per-photon table writer (photon_tables_photons.jl:30-148): two
    simulations appended as photons/dataset_1..2 (save_hdf!), 10 float64
    columns stored as HDF5.jl stores a Julia matrix, residuals t - t_geo - t0
    of photons that reached the module, the simulation's attributes."""
    from photonprop import hdf5, tables
    out = str(tmp_path / "photons.h5")
    for k, (mode, cos_t) in enumerate((("extended", 0.3), ("lightsabre_muon", -0.2))):
        name = tables.photon_table_photons(out, 1e4, 20.0, cos_t, 1.0, 100 + k, mode=mode,
                                           n_ph_limit=5000, log=print)
        assert name == f"dataset_{k + 1}"
    g = hdf5.read_hdf(out)["members"]["photons"]
    assert g["attrs"]["nsims"] == 2
    for k in (1, 2):
        d = g["members"][f"dataset_{k}"]
        t = d["data"].T
        assert t.shape[1] == 10 and 100 <= t.shape[0] <= 5000
        cols = dict(zip(tables.PHOTON_TABLE_COLUMNS, t.T))
        assert np.all(cols["module_id"] == 1)
        # photons hit the 0.3 m module at the origin; the reference's
        # discriminant puts the hit point of a direction of norm^2 1 + e,
        # arriving over a step of length d, at sqrt(r^2 + e d^2) from the
        # centre: photons scattered far away first come back a few cm off
        r = np.sqrt(cols["pos_x"] ** 2 + cols["pos_y"] ** 2 + cols["pos_z"] ** 2)
        assert np.all((r > 0.25) & (r < 0.4)) and np.median(np.abs(r - 0.3)) < 1e-3
        np.testing.assert_allclose(cols["dir_x"] ** 2 + cols["dir_y"] ** 2 + cols["dir_z"] ** 2,
                                   1.0, rtol=1e-5)
        assert np.all(cols["total_weight"] > 0) and np.all(cols["tres"] > -5.0)
        assert np.all((cols["wavelength"] >= 300) & (cols["wavelength"] <= 800))
        a = d["attrs"]
        assert a["mode"] == ("extended", "lightsabre_muon")[k - 1] and a["seed"] == 99 + k
        assert abs(a["distance"] - 20.0) < 1e-12 and "source_pos" in a


@pytest.mark.parametrize("order", [0, 1, 2])
def test_pmt_from_context_equals_host_entry(P, pom, order):
    """pp_pom_pmt_hits_ctx reads the context's hit buffer on the device; its
    PMT ids equal pp_pom_pmt_hits on the copied columns (relative positions
    formed in Float32 as the kernel forms them), row for row in the order the
    copy returns, for every copy order and for a sub-range."""
    targets = [pom.POM(np.float32([x, y, z]), 100 + i) for i, (x, y, z) in enumerate(
        [(0, 0, 10), (3, -2, 14), (-4, 1, 7), (1, 5, 12)])]
    src = P.ExtendedCherenkovEmitter(np.float32([0, 0, 0]), np.float32([0, 0, 1]), 0.0,
                                     4_000_000, P.LongitudinalParameterisation(7.837, 0.632, 0.356))
    medium = P.CascadiaMediumProperties()
    cfg = P._lib.pp_config()
    cfg.deterministic_order = order
    hbc, hbg = P.make_hit_buffers(capacity=1 << 22, config=cfg)
    with hbg:
        hits, _ = P.run_photon_prop_no_local_cache([src], targets, medium, P.Monochromatic(420.0),
                                                   5, hbc, hbg, with_ids=True)
        n = len(hits["wavelength"])
        assert n > 2000
        t = {}
        ctx_ids = pom.pmt_ids_in_context(hbg, seed=9, timing=t)
        sub = pom.pmt_ids_in_context(hbg, seed=9, first=n // 3, n=n // 4)
    c = np.array([t_.shape.position for t_ in targets], np.float32)[hits["module_id"] - 1]
    rel = (hits["position"] - c) / np.float32(0.3)
    host = pom._pmt_ids(rel.astype(np.float32), hits["wavelength"], pom.default_acceptance(),
                        None, hits["photon_id"], hits["source_ix"], 9, 0)
    assert np.array_equal(ctx_ids, host)
    assert np.array_equal(sub, host[n // 3:n // 3 + n // 4])
    assert t["n_detected"] == int(np.sum(host > 0)) > 0
    print(f"order {order}: {n} hits, {t['n_detected']} detected, kernel {t['kernel_ms']:.3f} ms")
