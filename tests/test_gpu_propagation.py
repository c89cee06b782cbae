"""Propagation on the GPU (through the C ABI) against the CPU oracle.

Both draw every photon from the same Philox stream, so the comparison is
photon by photon: the same photons must hit, at the same module and step,
with positions, times and lengths equal to float32 round-off. A photon may
differ only where round-off moves a grazing ray across a sphere boundary."""
import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu


def gpu_run(P, dev, wl, with_ids=True, photon_scaling=1.0):
    hbc = P.HostHitBuffer(np.float32)
    return P.run_photon_prop_no_local_cache(wl.sources, wl.targets, wl.medium, wl.spectrum,
                                            wl.seed, hbc, dev, n_steps=wl.n_steps,
                                            photon_scaling=photon_scaling, with_ids=with_ids)


def oracle_run(P, O, wl, photon_scaling=1.0, capacity=1 << 22):
    spec = wl.spectrum
    kw = dict(mono_wl=float(spec.wavelength)) if isinstance(spec, P.Monochromatic) else dict(
        table=spec.table, mapping=spec.mapping)
    return O.propagate(P.processing.sources_to_pod(wl.sources), wl.medium.to_pod(),
                       P.targets_to_pod(wl.targets), wl.seed, n_steps=wl.n_steps,
                       photon_scaling=photon_scaling, capacity=capacity, **kw)


def compare_photon_by_photon(g, r, max_diff_frac=0.005, tail_frac=0.02, module_diff_frac=0.0,
                             init_tail=0.0):
    """Same photons hit the same module at the same step; emission values
    agree to round-off for all of them. Transport values agree to round-off
    (5 mm + 2e-5 of the path) for all but `tail_frac` of them: HG(0.95)
    puts ~4e-4 of the scatterings per step within 1e-6 of cos = 1, where a
    1-ulp difference of the Float32 cosine becomes a >1e-5 rad direction
    difference that grows along the remaining path; those photons still
    agree within the module's diameter. With targets close to each other
    such a photon may reach another target instead: `module_diff_frac` of
    the common photons may then report a different module (and are left out
    of the value comparison). `init_tail`: the fraction of initial
    directions allowed beyond 2e-5 (all within 2e-3), for emitters that
    draw a smearing angle as a Float32 cosine next to 1 (the Cherenkov
    track: 1 - cos has density ~ x^-0.61 at 0, and one ulp of a cosine whose
    sine is s moves the direction by 2^-24 / s)."""
    gk = g["source_ix"].astype(np.uint64) << np.uint64(40) | g["photon_id"]
    rk = r["source_ix"].astype(np.uint64) << np.uint64(40) | r["photon_id"]
    assert np.all(np.diff(gk.astype(np.float64)) > 0), "GPU hits not in (source, photon) order"
    common, gi, ri = np.intersect1d(gk, rk, return_indices=True)
    n_diff = len(gk) + len(rk) - 2 * len(common)
    assert n_diff <= max(3, max_diff_frac * len(rk)), (len(gk), len(rk), len(common))
    same = g["module_id"][gi] == r["module_ix"][ri]
    if module_diff_frac > 0:
        print(f"{np.sum(~same)} of {len(common)} common photons report another module")
        assert np.sum(~same) <= module_diff_frac * len(common)
        gi, ri = gi[same], ri[same]
    assert np.all(g["module_id"][gi] == r["module_ix"][ri])
    assert np.all(g["n_steps"][gi] == r["n_steps"][ri])
    np.testing.assert_allclose(g["wavelength"][gi], r["wavelength"][ri], rtol=2e-6)
    d_init = np.max(np.abs(g["initial_direction"][gi].astype(np.float64) - r["init_dir"][ri]),
                    axis=1)
    assert np.mean(d_init > 2e-5) <= init_tail and np.all(d_init < 2e-3), (
        np.mean(d_init > 2e-5), d_init.max())
    dpos = np.max(np.abs(g["position"][gi] - r["pos"][ri]), axis=1)
    ddir = np.max(np.abs(g["direction"][gi] - r["dir"][ri]), axis=1)
    dt = np.abs(g["time"][gi] - r["time"][ri]) / np.maximum(r["time"][ri], 1.0)
    dd = np.abs(g["dist_travelled"][gi] - r["dist"][ri]) / np.maximum(r["dist"][ri], 1.0)
    tight = (dpos < 5e-3 + 2e-5 * r["dist"][ri]) & (ddir < 5e-3) & (dt < 1e-4) & (dd < 1e-4)
    print(f"tail {np.mean(~tight):.4f}; max dpos {dpos.max():.3g} m, ddir {ddir.max():.3g}, "
          f"dt {dt.max():.3g}, dd {dd.max():.3g}")
    assert np.mean(~tight) <= tail_frac, np.mean(~tight)
    assert np.all(dpos < 1.0) and np.all(dt < 1e-2) and np.all(dd < 1e-2)
    return len(common), n_diff


def test_readme_point_source_photon_parity(P, O, dev):
    wl = P.workloads.readme_point_source(photons=2_000_000)
    g, gsim = gpu_run(P, dev, wl)
    r, rsim = oracle_run(P, O, wl)
    assert gsim == rsim == 2_000_000
    n, d = compare_photon_by_photon(g, r)
    print(f"readme point source: {n} common hits, {d} differing")
    assert n > 300


def test_cascade_16_strings_photon_parity(P, O, dev):
    wl = P.workloads.cascade_16_strings(photons=300_000)
    g, gsim = gpu_run(P, dev, wl)
    r, rsim = oracle_run(P, O, wl)
    assert gsim == rsim
    n, d = compare_photon_by_photon(g, r)
    print(f"cascade 16 strings: {n} common hits, {d} differing")
    assert n > 20


EMITTERS = [
    lambda P: P.PointlikeIsotropicEmitter(np.float32([1, 2, 3]), 5.0, 300_000),
    lambda P: P.PointlikeTimeRangeEmitter(np.float32([1, 2, 3]), (0.0, 20.0), 300_000),
    lambda P: P.AxiconeEmitter(np.float32([0, 0, -5]), np.float32([0, 0, 1]), 0.0, 300_000,
                               np.deg2rad(45.0), np.deg2rad(0.5)),
    lambda P: P.CollimatedIsotropicEmitter(np.float32([0, 0, -5]), 0.0, 300_000,
                                           np.cos(np.deg2rad(60.0))),
    lambda P: P.PencilEmitter(np.float32([0, 0, -20]), np.float32([0.05, 0.0, 0.99875]), 0.02,
                              0.0, 300_000),
    lambda P: P.ExtendedCherenkovEmitter(np.float32([2, -3, -10]), np.float32([0, 0.6, 0.8]),
                                         0.0, 300_000,
                                         P.LongitudinalParameterisation(7.837, 0.632, 0.356)),
    lambda P: P.ExtendedCherenkovEmitterNoSmear(
        np.float32([2, -3, -10]), np.float32([0, 0.6, 0.8]), 0.0, 300_000,
        P.LongitudinalParameterisation(7.837, 0.632, 0.356)),
    lambda P: P.ExtendedCherenkovEmitterCustomSmear(
        np.float32([2, -3, -10]), np.float32([0, 0.6, 0.8]), 0.0, 300_000, 1.776, 0.164,
        P.LongitudinalParameterisation(7.837, 0.632, 0.356)),
    lambda P: P.PointlikeCherenkovEmitter(np.float32([2, -3, -10]), np.float32([0, 0, 1]), 0.0,
                                          300_000),
    lambda P: P.CherenkovTrackEmitter(np.float32([0, -40, -10]), np.float32([0, 1, 0]), 0.0,
                                      80.0, 300_000),
]


@pytest.mark.parametrize("k", range(len(EMITTERS)))
def test_every_emitter_photon_parity(P, O, dev, k):
    medium = P.CascadiaMediumProperties()
    targets = P.make_detector_grid(3, 3, 12.0, 5, 8.0, radius=0.5)
    src = EMITTERS[k](P)
    spec = P.workloads.acceptance_weighted_spectrum(medium)
    wl = P.workloads.Workload("emitter", [src], targets, medium, spec, 5 + k)
    g, gsim = gpu_run(P, dev, wl)
    r, rsim = oracle_run(P, O, wl)
    assert gsim == rsim == 300_000
    track = isinstance(src, P.CherenkovTrackEmitter)
    n, d = compare_photon_by_photon(g, r, init_tail=0.01 if track else 0.0)
    print(f"{type(src).__name__}: {n} common hits, {d} differing")
    assert n > 20


def test_multi_source_and_scaling(P, O, dev):
    medium = P.CascadiaMediumProperties()
    targets = P.make_detector_grid(2, 2, 15.0, 4, 8.0, radius=0.4)
    srcs = [P.PointlikeIsotropicEmitter(np.float32([0, 0, 0]), 0.0, 100_001),
            P.PointlikeIsotropicEmitter(np.float32([5, 5, 5]), 10.0, 0),
            P.CherenkovTrackEmitter(np.float32([-20, 0, 0]), np.float32([1, 0, 0]), 0.0, 40.0,
                                    77_777)]
    wl = P.workloads.Workload("multi", srcs, targets, medium, P.Monochromatic(400.0), 3)
    g, gsim = gpu_run(P, dev, wl, photon_scaling=0.5)
    r, rsim = oracle_run(P, O, wl, photon_scaling=0.5)
    assert gsim == rsim == 50_000 + 0 + 38_888
    compare_photon_by_photon(g, r)
    assert set(np.unique(g["source_ix"])) <= {0, 2}


def test_bit_exact_repeat_and_launch_shape_independence(P, dev):
    wl = P.workloads.cascade_16_strings(photons=2_000_000, seed=9)
    g1, n1 = gpu_run(P, dev, wl)
    g2, n2 = gpu_run(P, dev, wl)
    cfg = P._lib.pp_config()
    cfg.block_size = 64
    cfg.waves_per_simd = 1
    cfg.deterministic_order = 1
    _, dev2 = P.make_hit_buffers(capacity=1 << 20, config=cfg)
    with dev2:
        g3, n3 = gpu_run(P, dev2, wl)
    assert n1 == n2 == n3
    assert len(g1["wavelength"]) > 100
    for k in g1:
        assert g1[k].tobytes() == g2[k].tobytes(), k
        assert g1[k].tobytes() == g3[k].tobytes(), k


def test_ordered_copy_in_chunks(P, gpu_lib):
    """Ordered copies of sub-ranges, with an unordered copy in between, give
    the rows of one ordered copy of every hit; the unordered copy gives the
    same multiset of hits, and the host sort the device sort's rows."""
    import ctypes
    wl = P.workloads.uneven_batch_1000(n_sources=6, seed=5, min_photons=1e5, max_photons=2e5)
    cfg = P._lib.pp_config()
    cfg.deterministic_order = 1
    _, d = P.make_hit_buffers(capacity=1 << 20, config=cfg)
    with d:
        g, _ = gpu_run(P, d, wl)
        n = len(g["photon_id"])
        assert n > 100

        def copy(first, cnt):
            pid = np.zeros(cnt, np.uint64)
            src = np.zeros(cnt, np.int32)
            pos = np.zeros((cnt, 3), np.float32)
            assert gpu_lib.pp_copy_hits(d.handle, first, cnt, P._lib.fptr(pos), None, None, None,
                                        None, 0, None, None, None, P._lib.fptr(pid),
                                        P._lib.fptr(src)) == 0
            return pid, src, pos

        a = copy(0, n // 3)
        cfg.deterministic_order = 0
        assert gpu_lib.pp_set_config(d.handle, ctypes.byref(cfg)) == 0
        u = copy(0, n)
        cfg.deterministic_order = 1
        assert gpu_lib.pp_set_config(d.handle, ctypes.byref(cfg)) == 0
        b = copy(n // 3, n - n // 3)
        for k, part in enumerate((a, b)):
            lo = 0 if k == 0 else n // 3
            sl = slice(lo, lo + len(part[0]))
            assert part[0].tobytes() == g["photon_id"][sl].tobytes()
            assert part[1].tobytes() == g["source_ix"][sl].astype(np.int32).tobytes()
        keys = lambda p, s: sorted(zip(s.tolist(), p.tolist()))
        assert keys(u[0], u[1]) == keys(np.concatenate([a[0], b[0]]), np.concatenate([a[1], b[1]]))
        # the host sort (deterministic_order = 2) gives the device sort's rows
        cfg.deterministic_order = 2
        assert gpu_lib.pp_set_config(d.handle, ctypes.byref(cfg)) == 0
        h = copy(5, n - 7)
        assert h[0].tobytes() == g["photon_id"][5:n - 2].tobytes()
        assert h[1].tobytes() == g["source_ix"][5:n - 2].astype(np.int32).tobytes()
        assert h[2].tobytes() == g["position"][5:n - 2].astype(np.float32).tobytes()
        cfg.deterministic_order = 3
        assert gpu_lib.pp_set_config(d.handle, ctypes.byref(cfg)) == P._lib.PP_ERR_INVALID


def test_hit_buffer_overflow_raises(P, O, gpu_lib):
    wl = P.workloads.readme_point_source(photons=1_000_000)
    _, small = P.make_hit_buffers(capacity=8)
    with small:
        with pytest.raises(P.PhotonPropOOMException):
            gpu_run(P, small, wl)
        # the eight stored records are genuine hits
        import ctypes
        lib = gpu_lib
        pid = np.zeros(8, np.uint64)
        pos = np.zeros((8, 3), np.float32)
        assert lib.pp_copy_hits(small.handle, 0, 8, P._lib.fptr(pos), None, None, None, None, 0,
                                None, None, None, P._lib.fptr(pid), None) == 0
        # ranges beyond the stored records are refused, also where first + n overflows
        for first, n in ((0, 9), (8, 1), (9, 0), (1, 2**63 - 1), (-1, 1)):
            assert lib.pp_copy_hits(small.handle, first, n, P._lib.fptr(pos), None, None, None,
                                    None, 0, None, None, None, None, None) == P._lib.PP_ERR_INVALID
        r, _ = oracle_run(P, O, wl)
        idx = {int(p): i for i, p in enumerate(r["photon_id"])}
        hits = [idx.get(int(p)) for p in pid]
        assert sum(h is not None for h in hits) >= 7
        for p, h in zip(pos, hits):
            if h is not None:
                np.testing.assert_allclose(p, r["pos"][h], atol=5e-3)
        del ctypes


def test_photon_count_limit(P, dev):
    """A source of 2^40 or more photons after the Float32 scaling product (the
    photon id field's width) is refused, also where the product exceeds the
    int64 range, and the context stays usable."""
    medium = P.CascadiaMediumProperties()
    targets = P.make_detector_grid(2, 2, 15.0, 4, 8.0, radius=0.4)
    for photons, scaling in ((2**40, 1.0), (10**12, 1e10), (2**62, 1.0)):
        src = P.PointlikeIsotropicEmitter(np.float32([0, 0, 0]), 0.0, photons)
        wl = P.workloads.Workload("limit", [src], targets, medium, P.Monochromatic(400.0), 1)
        with pytest.raises(P._lib.PhotonPropLibraryError, match="2\\^40"):
            gpu_run(P, dev, wl, photon_scaling=scaling)
    src = P.PointlikeIsotropicEmitter(np.float32([0, 0, 0]), 0.0, 1000)
    wl = P.workloads.Workload("limit", [src], targets, medium, P.Monochromatic(400.0), 1)
    _, n = gpu_run(P, dev, wl)
    assert n == 1000


def test_zero_photons_and_zero_steps(P, dev):
    wl = P.workloads.readme_point_source(photons=0)
    g, n = gpu_run(P, dev, wl)
    assert n == 0 and len(g["wavelength"]) == 0
    wl = P.workloads.readme_point_source(photons=10_000)
    wl.n_steps = 0
    g, n = gpu_run(P, dev, wl)
    assert n == 10_000 and len(g["wavelength"]) == 0


def test_statistical_parity_cascade(P, O, dev):
    """Independent seeds: per-module hit counts (chi2) and the distributions
    of time, distance and hit z (KS) agree between GPU and oracle."""
    wl_g = P.workloads.cascade_16_strings(photons=40_000_000, seed=42)
    wl_g.seed = 1001
    g, gsim = gpu_run(P, dev, wl_g, with_ids=False)
    wl_r = P.workloads.cascade_16_strings(photons=2_000_000, seed=42)
    wl_r.seed = 2002
    r, rsim = oracle_run(P, O, wl_r)
    assert len(r) > 200
    mods = np.arange(1, 321)
    cg = np.bincount(g["module_id"], minlength=321)[1:]
    cr = np.bincount(r["module_ix"], minlength=321)[1:]
    keep = (cg + cr) >= 10
    table = np.stack([cg[keep], cr[keep]])
    other = np.array([[cg[~keep].sum()], [cr[~keep].sum()]])
    if other.sum() > 0:
        table = np.concatenate([table, other], axis=1)
    # scale: equal photon totals are not needed, contingency chi2 compares shapes
    chi2 = stats.chi2_contingency(table)
    print(f"modules={keep.sum()} chi2 p={chi2.pvalue:.3f}")
    assert chi2.pvalue > 0.01
    # total hit rate per photon
    rate = stats.binomtest(len(r), rsim, len(g["wavelength"]) / gsim)
    assert rate.pvalue > 0.01
    for name, gk, rk in (("time", "time", "time"), ("dist", "dist_travelled", "dist")):
        p = stats.ks_2samp(g[gk], r[rk]).pvalue
        print(f"KS {name} p={p:.3f}")
        assert p > 0.01
    p = stats.ks_2samp(g["position"][:, 2], r["pos"][:, 2]).pvalue
    assert p > 0.01
    del mods


@pytest.mark.parametrize("name", ["cascade_16_strings", "uneven_batch_1000"])
def test_full_size_workloads(P, name):
    """The benchmark's workloads at their full sizes (1e8 photons of one
    cascade; 256 cascades and tracks, ~3.7e8 photons, 1000 modules), checked
    through properties that hold at any size: every photon accounted for;
    the same hits, bit for bit, under two launch schedules (work split and
    chunking differ, the per-photon streams do not); and every hit physical:
    a module of the list, 1..n_steps steps, on its sphere, and arriving
    dist / c_grp(wavelength) after its emission at the source's time (plus
    the emission point's offset along a cascade or track, >= 0)."""
    wl = getattr(P.workloads, name)()
    runs = []
    for block, chunk in ((256, 256), (512, 2048)):
        cfg = P._lib.pp_config()
        cfg.block_size = block
        cfg.chunk_photons = chunk
        cfg.deterministic_order = 1
        _, dev = P.make_hit_buffers(capacity=1 << 24, config=cfg)
        with dev:
            runs.append(gpu_run(P, dev, wl))
    (g, n1), (g2, n2) = runs
    expected = sum(int(np.float32(s.photons) * np.float32(1.0)) for s in wl.sources)
    assert n1 == n2 == expected
    n_hits = len(g["wavelength"])
    print(f"{name}: {expected} photons, {n_hits} hits")
    assert n_hits > 10_000
    for k in g:
        assert g[k].tobytes() == g2[k].tobytes(), k
    mid = g["module_id"]
    assert mid.min() >= 1 and mid.max() <= len(wl.targets)
    assert g["n_steps"].min() >= 1 and g["n_steps"].max() <= wl.n_steps
    centre = np.array([t.shape.position for t in wl.targets], np.float64)[mid - 1]
    radius = np.array([t.shape.radius for t in wl.targets], np.float64)[mid - 1]
    # on the sphere, up to the reference's discriminant r^2 - |q|^2 + eps a^2
    # (eps = |dir|^2 - 1 of the Float32 direction), which puts the crossing at
    # |p - c|^2 = r^2 + eps d^2 for a crossing d along the step (d <= the
    # distance travelled), plus Float32 round-off of positions ~100 m out
    p2 = np.sum((g["position"].astype(np.float64) - centre) ** 2, axis=1)
    d = g["direction"].astype(np.float64)
    eps = np.abs(np.sum(d * d, axis=1) - 1.0)
    excess = np.abs(p2 - radius ** 2)
    bound = 1.05 * eps * g["dist_travelled"].astype(np.float64) ** 2 + 1e-3
    off = np.sqrt(p2) - radius
    print(f"distance from the sphere: median {np.median(np.abs(off)):.2e} m, max "
          f"{np.abs(off).max():.2e} m; max excess / bound {np.max(excess / bound):.3f}")
    assert np.median(np.abs(off)) < 1e-3
    assert np.all(excess <= bound)
    t_src = np.array([s.time for s in wl.sources], np.float64)[g["source_ix"]]
    lag = g["time"] - g["dist_travelled"] / wl.medium.group_velocity(g["wavelength"]) - t_src
    assert np.all(lag >= -(1e-3 + 1e-5 * g["time"])), lag.min()
    max_offset = max(float(getattr(s, "length", 0.0) or 0.0) for s in wl.sources) + 100.0
    assert np.all(lag <= max_offset / 0.299792458 + 1e-3 + 1e-5 * g["time"]), lag.max()


@pytest.mark.parametrize("cell,cell_z", [(0.0, 0.0), (4.0, 0.0), (30.0, 0.0), (0.0, -1.0),
                                         (20.0, 60.0)])
def test_grid_search_bit_exact_with_brute_force(P, gpu_lib, cell, cell_z):
    """The uniform-grid search returns the reference's first-in-index target:
    hits identical bit for bit to the brute-force scan."""
    import ctypes
    wl = P.workloads.cascade_16_strings(photons=3_000_000, seed=11)
    out = {}
    for search in (1, 2):
        cfg = P._lib.pp_config()
        cfg.search = search
        cfg.grid_cell = cell
        cfg.grid_cell_z = cell_z
        cfg.deterministic_order = 1
        _, d = P.make_hit_buffers(capacity=1 << 20, config=cfg)
        with d:
            out[search] = gpu_run(P, d, wl)
            info = ctypes.c_int32()
            h = ctypes.c_float()
            dims = (ctypes.c_int32 * 3)()
            ni = ctypes.c_int32()
            gpu_lib.pp_search_info(d.handle, ctypes.byref(info), ctypes.byref(h), dims,
                                   ctypes.byref(ni))
            assert info.value == search
            if search == 2:
                print(f"grid cell {h.value:.2f} m, dims {list(dims)}, {ni.value} entries")
    (gb, nb), (gg, ng) = out[1], out[2]
    assert nb == ng and len(gb["wavelength"]) > 300
    for k in gb:
        assert gb[k].tobytes() == gg[k].tobytes(), k


def test_lattice_of_4096_spheres(P, O, gpu_lib):
    """A 16 x 16 x 16 lattice of spheres (4096 targets: their records alone
    fill the grid's usual 64 KB of LDS, so the grid takes the whole LDS)
    around a point source: a grid of 4 m cubic cells (a 3-D walk) gives the
    brute-force scan's hits bit for bit, both agree photon by photon with the
    oracle, and the automatic search picks the grid, whose cells (the model's
    choice: one column per vertical line of spheres) give the same hits."""
    import ctypes
    import itertools
    medium = P.CascadiaMediumProperties()
    xs = np.arange(16) * 4.0 - 30.0
    targets = [P.spherical_target((x, y, z), 0.5, i + 1)
               for i, (x, y, z) in enumerate(itertools.product(xs, xs, xs))]
    src = P.PointlikeIsotropicEmitter(np.float32([0.3, -0.7, 1.1]), 0.0, 100_000)
    wl = P.workloads.Workload("lattice", [src], targets, medium, P.Monochromatic(420.0), 21)
    out = {}
    for search in (1, 2):
        cfg = P._lib.pp_config()
        cfg.search = search
        cfg.deterministic_order = 1
        if search == 2:
            cfg.grid_cell = 4.0  # cubic cells (cell_z 0 with a set edge)
        _, d = P.make_hit_buffers(capacity=1 << 20, config=cfg)
        with d:
            out[search] = gpu_run(P, d, wl)
            info, h, ni = ctypes.c_int32(), ctypes.c_float(), ctypes.c_int32()
            dims = (ctypes.c_int32 * 3)()
            gpu_lib.pp_search_info(d.handle, ctypes.byref(info), ctypes.byref(h), dims,
                                   ctypes.byref(ni))
            assert info.value == search
            if search == 2:
                print(f"grid cell {h.value:.2f} m, dims {list(dims)}, {ni.value} entries")
                assert dims[2] > 1  # a 3-D grid, not columns
    (gb, nb), (gg, ng) = out[1], out[2]
    cfg = P._lib.pp_config()
    cfg.deterministic_order = 1
    _, d = P.make_hit_buffers(capacity=1 << 20, config=cfg)
    with d:
        ga, na = gpu_run(P, d, wl)
        info, h = ctypes.c_int32(), ctypes.c_float()
        dims = (ctypes.c_int32 * 3)()
        gpu_lib.pp_search_info(d.handle, ctypes.byref(info), ctypes.byref(h), dims, None)
        assert info.value == 2
        print(f"automatic grid: cell {h.value:.2f} m, dims {list(dims)}")
    assert nb == ng == na == 100_000 and len(gb["wavelength"]) > 1000
    for k in gb:
        assert gb[k].tobytes() == gg[k].tobytes(), k
        assert gb[k].tobytes() == ga[k].tobytes(), k
    r, rsim = oracle_run(P, O, wl)
    assert rsim == 100_000
    n, d = compare_photon_by_photon(gg, r, module_diff_frac=0.01)
    print(f"lattice: {n} common hits, {d} differing")


def test_16000_spheres_from_device_memory(P, O, gpu_lib):
    """16000 spheres (a 20 x 20 x 40 lattice): records and grid tables exceed
    the LDS, so the automatic search walks a grid read from device memory. It
    agrees photon by photon with the oracle's every-target scan; the
    brute-force scan on the device, which needs the records in LDS, is
    refused with a message."""
    import ctypes
    import itertools
    medium = P.CascadiaMediumProperties()
    xs, zs = np.arange(20) * 5.0 - 47.5, np.arange(40) * 5.0 - 97.5
    targets = [P.spherical_target((x, y, z), 0.6, i + 1)
               for i, (x, y, z) in enumerate(itertools.product(xs, xs, zs))]
    src = P.PointlikeIsotropicEmitter(np.float32([1.3, -0.4, 2.1]), 0.0, 30_000)
    wl = P.workloads.Workload("lattice", [src], targets, medium, P.Monochromatic(420.0), 22)
    cfg = P._lib.pp_config()
    cfg.deterministic_order = 1
    _, d = P.make_hit_buffers(capacity=1 << 20, config=cfg)
    with d:
        g, n = gpu_run(P, d, wl)
        info, h, ni = ctypes.c_int32(), ctypes.c_float(), ctypes.c_int32()
        dims = (ctypes.c_int32 * 3)()
        gpu_lib.pp_search_info(d.handle, ctypes.byref(info), ctypes.byref(h), dims, ctypes.byref(ni))
        print(f"grid cell {h.value:.2f} m, dims {list(dims)}, {ni.value} entries")
        assert info.value == 2 and ni.value >= 16000
    assert n == 30_000 and len(g["wavelength"]) > 1000
    r, rsim = oracle_run(P, O, wl)
    cn, dn = compare_photon_by_photon(g, r, module_diff_frac=0.01)
    print(f"16000 spheres: {cn} common hits, {dn} differing")
    cfg.search = 1
    _, d2 = P.make_hit_buffers(capacity=1 << 20, config=cfg)
    with d2:
        with pytest.raises(P._lib.PhotonPropLibraryError, match="LDS"):
            gpu_run(P, d2, wl)


def test_grid_search_1000_modules_matches_oracle(P, O, dev):
    """Uneven multi-source batch over 1000 modules (grid search by default)
    photon by photon against the oracle's brute-force scan."""
    wl = P.workloads.uneven_batch_1000(n_sources=8, seed=3, min_photons=1e5, max_photons=3e5)
    g, gsim = gpu_run(P, dev, wl)
    r, rsim = oracle_run(P, O, wl)
    assert gsim == rsim
    # long track and far-source paths: more photons cross a module edge after
    # a round-off direction difference (see compare_photon_by_photon)
    n, d = compare_photon_by_photon(g, r, max_diff_frac=0.02)
    print(f"1000 modules, 6 sources: {n} common hits, {d} differing")
    assert n > 20


def _run_cfg(P, wl, seed=None, **cfg_fields):
    cfg = P._lib.pp_config()
    cfg.deterministic_order = 1
    for k, v in cfg_fields.items():
        setattr(cfg, k, v)
    _, d = P.make_hit_buffers(capacity=1 << 21, config=cfg)
    with d:
        out = gpu_run(P, d, wl)
        if seed is not None:  # a second call on the same context
            wl.seed, old = seed, wl.seed
            out = (out, gpu_run(P, d, wl))
            wl.seed = old
    return out


def test_batch_modes_bit_identical(P, dev):
    """One launch per source, the persistent work-stealing kernel (auto and
    64-photon chunks; 960-photon chunks, which a batch this small hands out
    entirely as pieces of 256, 256, 256 and 192 photons, some past a source's
    end; the automatic block size, 768 threads for these 1000 modules'
    tables, and 256) and its hipGraph give the same hits bit for bit,
    including on a second call that re-parameterises the graph."""
    wl = P.workloads.uneven_batch_1000(n_sources=40, seed=5, min_photons=1e3, max_photons=2e5)
    wl.sources[3].photons = 0
    wl.sources[17].photons = 37  # a partial wave
    ref = _run_cfg(P, wl, batch_mode=P._lib.BATCH_PER_SOURCE)
    ref2 = _run_cfg(P, wl, batch_mode=P._lib.BATCH_PER_SOURCE, seed=wl.seed + 1)[1]
    assert len(ref[0]["wavelength"]) > 100
    runs = {
        "persistent": _run_cfg(P, wl, batch_mode=P._lib.BATCH_PERSISTENT),
        "chunk64": _run_cfg(P, wl, batch_mode=P._lib.BATCH_PERSISTENT, chunk_photons=64,
                            block_size=128),
        "block256": _run_cfg(P, wl, batch_mode=P._lib.BATCH_PERSISTENT, block_size=256),
        "pieces": _run_cfg(P, wl, batch_mode=P._lib.BATCH_PERSISTENT, chunk_photons=960),
        "graph": _run_cfg(P, wl, batch_mode=P._lib.BATCH_GRAPH, seed=wl.seed + 1),
    }
    runs["graph"], graph2 = runs["graph"]
    for name, (g, n) in runs.items():
        assert n == ref[1], name
        for k in g:
            assert g[k].tobytes() == ref[0][k].tobytes(), (name, k)
    assert graph2[1] == ref2[1]
    for k in graph2[0]:
        assert graph2[0][k].tobytes() == ref2[0][k].tobytes(), ("graph second call", k)


def test_batch_mode_overflow_raises(P):
    wl = P.workloads.uneven_batch_1000(n_sources=6, seed=8, min_photons=1e5, max_photons=2e5)
    for mode in (P._lib.BATCH_PERSISTENT, P._lib.BATCH_GRAPH):
        cfg = P._lib.pp_config()
        cfg.batch_mode = mode
        _, small = P.make_hit_buffers(capacity=16, config=cfg)
        with small:
            with pytest.raises(P.PhotonPropOOMException):
                gpu_run(P, small, wl)


def test_autotune_picks_a_candidate_and_keeps_results(P, tmp_path):
    """The tuned launch is one of the searched candidates, is cached, and the
    hits after tuning equal those of the default launch bit for bit."""
    wl = P.workloads.uneven_batch_1000(n_sources=12, seed=4, min_photons=1e4, max_photons=3e5)
    ref = _run_cfg(P, wl)
    cache = str(tmp_path / "tune.json")
    cfg = P._lib.pp_config()
    cfg.deterministic_order = 1
    _, d = P.make_hit_buffers(capacity=1 << 21, config=cfg)
    with d:
        r = P.autotune.autotune(d, wl.sources, wl.targets, wl.medium, wl.spectrum, wl.seed,
                                n_steps=wl.n_steps, sample_photons=2e6, cache_path=cache)
        print("autotune:", r)
        c = r["config"]
        assert not r["cached"] and r["ms"] > 0
        assert c["block_size"] in (0, 64, 128, 256, 384, 512, 768, 1024)
        assert c["chunk_photons"] in (0, 64, 128, 256, 512, 1024)
        assert c["batch_mode"] == P._lib.BATCH_PERSISTENT
        g = gpu_run(P, d, wl)
    assert g[1] == ref[1]
    for k in g[0]:
        assert g[0][k].tobytes() == ref[0][k].tobytes(), k
    _, d2 = P.make_hit_buffers(capacity=1 << 21, config=cfg)
    with d2:
        r2 = P.autotune.autotune(d2, wl.sources, wl.targets, wl.medium, wl.spectrum, wl.seed,
                                 n_steps=wl.n_steps, sample_photons=2e6, cache_path=cache)
        assert r2["cached"] and r2["config"] == c
        got = P._lib.pp_config()
        P._lib.load_library().pp_get_config(d2.handle, __import__("ctypes").byref(got))
        assert P.autotune.config_to_dict(got) == c
    # a corrupt cache is tuned again and rewritten; an unwritable one only skips caching
    with open(cache, "w") as f:
        f.write("{not json")
    blocked = tmp_path / "file"
    blocked.write_text("")
    for path, cached_after in ((cache, True), (str(blocked / "sub" / "tune.json"), False)):
        _, d3 = P.make_hit_buffers(capacity=1 << 21, config=cfg)
        with d3:
            r3 = P.autotune.autotune(d3, wl.sources, wl.targets, wl.medium, wl.spectrum, wl.seed,
                                     n_steps=wl.n_steps, sample_photons=2e6, cache_path=path)
            assert not r3["cached"]
        import os
        assert os.path.exists(path) == cached_after


def test_ice_medium_photon_parity(P, O, dev):
    """Homogeneous clear ice (SPICE lengths, HG + simplified-Liu phase
    function): photon by photon against the oracle."""
    medium = P.make_homogenous_clearice_properties()
    targets = P.make_detector_grid(3, 3, 15.0, 4, 10.0, radius=0.3)
    src = P.PointlikeIsotropicEmitter(np.float32([2, 1, 3]), 0.0, 1_000_000)
    spec = P.make_cuda_spectral_dist(P.make_cherenkov_spectrum((300.0, 700.0), medium).spectral_dist,
                                     (300.0, 700.0), 600)
    wl = P.workloads.Workload("ice", [src], targets, medium, spec, 21)
    g, gsim = gpu_run(P, dev, wl)
    r, rsim = oracle_run(P, O, wl)
    assert gsim == rsim
    n, d = compare_photon_by_photon(g, r)
    print(f"ice: {n} common hits, {d} differing")
    assert n > 100


def _planar_workload(P, photons=1_000_000, seed=33):
    """Spheres, rectangles and discs in one list around an isotropic source in
    Cascadia water: the mixed (every-target) search."""
    targets = []
    mid = 1
    for k, z in enumerate((-8.0, -3.0, 4.0, 9.0)):
        targets.append(P.rectangular_target((1.5 * k - 2, 0.5, z), 3.0, 2.0, mid))
        targets.append(P.circular_target((-1.0, 2.0 - k, z + 1.0), 1.2, mid + 1))
        targets.append(P.spherical_target((3.0, -2.0, z), 0.5, mid + 2))
        mid += 3
    src = P.PointlikeIsotropicEmitter(np.float32([0.3, -0.2, 0.5]), 0.0, photons)
    medium = P.CascadiaMediumProperties()
    return P.workloads.Workload("planar", [src], targets, medium,
                                P.workloads.acceptance_weighted_spectrum(medium), seed)


def test_planar_targets_photon_parity(P, O, dev):
    """Rectangles and discs (check_intersection(::Rectangular / ::Circular))
    mixed with spheres: photon by photon against the oracle."""
    wl = _planar_workload(P)
    g, gsim = gpu_run(P, dev, wl)
    r, rsim = oracle_run(P, O, wl)
    assert gsim == rsim
    n, d = compare_photon_by_photon(g, r, module_diff_frac=0.005)
    mods = np.bincount(g["module_id"].astype(np.int64), minlength=13)
    print(f"planar: {n} common hits, {d} differing; per module {mods[1:].tolist()}")
    assert n > 1000
    # every kind of target is hit; planar hits lie in their plane
    kinds = np.array([type(t.shape).__name__ for t in wl.targets])
    hit_kinds = set(kinds[g["module_id"].astype(np.int64) - 1])
    assert hit_kinds == {"Rectangular", "Circular", "Spherical"}
    zt = np.array([t.shape.position[2] for t in wl.targets], np.float32)
    planar = kinds[g["module_id"].astype(np.int64) - 1] != "Spherical"
    np.testing.assert_allclose(g["position"][planar, 2],
                               zt[g["module_id"][planar].astype(np.int64) - 1], atol=2e-5)


def test_planar_targets_reject_grid_search_and_keep_batch_modes(P):
    wl = _planar_workload(P, photons=200_000)
    with pytest.raises(RuntimeError, match="spherical targets only"):
        _run_cfg(P, wl, search=P._lib.SEARCH_GRID)
    ref = _run_cfg(P, wl, batch_mode=P._lib.BATCH_PER_SOURCE)
    for mode in (P._lib.BATCH_PERSISTENT, P._lib.BATCH_GRAPH):
        g = _run_cfg(P, wl, batch_mode=mode)
        assert g[1] == ref[1]
        for k in g[0]:
            assert g[0][k].tobytes() == ref[0][k].tobytes(), (mode, k)


def test_microbench_issue_rates(P, gpu_lib):
    """The roofline's measured issue rates: FP32 FMA near the non-packed
    peak of MI355X (78.6 TFLOP/s) and never above it; the transcendental
    and 64-bit integer products slower."""
    import ctypes
    r = {}
    for k in (P._lib.MB_FMA_F32, P._lib.MB_TRANS_F32, P._lib.MB_MAD_U64):
        v = ctypes.c_double()
        assert gpu_lib.pp_microbench(0, k, ctypes.byref(v)) == 0
        r[k] = v.value
    print({k: f"{v / 1e12:.1f} T/s" for k, v in r.items()})
    assert 30e12 < r[P._lib.MB_FMA_F32] <= 1.05 * 78.6e12
    assert 0 < r[P._lib.MB_TRANS_F32] < r[P._lib.MB_FMA_F32] / 2
    assert 0 < r[P._lib.MB_MAD_U64] < r[P._lib.MB_FMA_F32] / 2
    assert gpu_lib.pp_microbench(0, 7, ctypes.byref(ctypes.c_double())) != 0


def test_copy_of_millions_of_hits_through_bounded_staging(P, gpu_lib):
    """More hits than one pinned staging chunk (2^20 records): the copies in
    buffer order, device order and host order hold the same hits, the two
    ordered copies agree row for row, and a copy after pp_release_scratch
    (which frees the sort buffer and the staging) equals the one before."""
    import ctypes
    medium = P.SimpleMediumProperties(30.0, 1e9, 1.35, 0.9, 0.2)
    src = P.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, 12_000_000)
    targets = [P.spherical_target((0.0, 0.0, 6.0), 5.0, 1)]
    cfg = P._lib.pp_config()
    cfg.deterministic_order = 1
    _, d = P.make_hit_buffers(capacity=1 << 22, config=cfg)
    with d:
        def copy(order):
            cfg.deterministic_order = order
            assert gpu_lib.pp_set_config(d.handle, ctypes.byref(cfg)) == 0
            hits, _ = P.run_photon_prop_no_local_cache([src], targets, medium,
                                                       P.Monochromatic(450.0), 3,
                                                       P.HostHitBuffer(np.float32), d,
                                                       with_ids=True)
            return hits
        h1 = copy(1)
        n = len(h1["photon_id"])
        print(f"{n} hits")
        assert n > (1 << 21)
        h0, h2 = copy(0), copy(2)
        d.release_scratch()
        h1b = copy(1)
        for k in ("photon_id", "position", "time", "module_id"):
            assert h1[k].tobytes() == h2[k].tobytes() == h1b[k].tobytes(), k
        o = np.argsort(h0["photon_id"], kind="stable")
        assert h0["photon_id"][o].tobytes() == h1["photon_id"].tobytes()
        assert h0["position"][o].tobytes() == h1["position"].tobytes()
