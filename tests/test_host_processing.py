"""Output assembly of propagate_photons (src/processing.jl:32-44, :67-72) on
host data: column order, module-id remap, base_weight, abs_weight and
total_weight, pinned by reference-held absorption lengths (26.775087 m at
488 nm printed by examples/axicone.ipynb; 27.7 m at 450 nm, the STRAW knot
that src/water_properties.jl:22-23 also lists) and by hand-computed values."""
import math
import types

import numpy as np
import pytest

REF_COLUMNS = ["position", "direction", "initial_direction", "wavelength", "time",
               "dist_travelled", "module_id", "n_steps", "base_weight", "abs_weight",
               "total_weight"]


def make_setup(P, abs_scale=1.0):
    targets = [P.spherical_target((0, 0, 0), 0.3, 7),
               P.spherical_target((0, 0, 10), 0.3, 3),
               P.spherical_target((0, 0, 20), 0.3, 11)]
    src = [P.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, 1000),
           P.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, 3000)]
    medium = P.CascadiaMediumProperties(0.95, 0.2, abs_scale, 1.0)
    return types.SimpleNamespace(sources=src, targets=targets, medium=medium)


def hits_dict(n, wl, dist, module_ix):
    rng = np.random.default_rng(5)
    return dict(position=rng.normal(size=(n, 3)).astype(np.float32),
                direction=rng.normal(size=(n, 3)).astype(np.float32),
                initial_direction=rng.normal(size=(n, 3)).astype(np.float32),
                wavelength=np.asarray(wl, np.float32), time=np.arange(n, dtype=np.float32),
                dist_travelled=np.asarray(dist, np.float32),
                module_id=np.asarray(module_ix, np.int64),
                n_steps=np.arange(1, n + 1, dtype=np.int32))


def test_column_order_remap_and_weights(P):
    setup = make_setup(P)
    d488 = 26.775087
    hits = hits_dict(4, [488.0, 450.0, 450.0, 488.0], [d488, 27.7, 0.0, 2 * d488], [1, 3, 2, 3])
    df = P.processing.assemble_hit_table(hits, 3999, setup)
    assert list(df.columns) == REF_COLUMNS
    # mod_ids[df.module_id]: 1-based target index -> the target's module_id
    assert df["module_id"].tolist() == [7, 11, 3, 11]
    # sum(photons) / n_ph_sim
    assert df["base_weight"].tolist() == [4000 / 3999] * 4
    # exp(-dist / abs_len(wl)) at reference-held absorption lengths
    np.testing.assert_allclose(df["abs_weight"], [math.exp(-1), math.exp(-1), 1.0, math.exp(-2)],
                               rtol=2e-6)
    np.testing.assert_allclose(df["total_weight"], df["abs_weight"] * (4000 / 3999), rtol=1e-15)
    assert df["abs_weight"].dtype == np.float64 and df["total_weight"].dtype == np.float64
    # untouched hit columns keep their values and Float32 type
    np.testing.assert_array_equal(np.stack(df["position"].to_numpy()), hits["position"])
    assert df["wavelength"].dtype == np.float32 and df["n_steps"].tolist() == [1, 2, 3, 4]


def test_abs_scale_enters_the_absorption_weight(P):
    setup = make_setup(P, abs_scale=2.0)
    hits = hits_dict(1, [450.0], [55.4], [1])
    df = P.processing.assemble_hit_table(hits, 4000, setup)
    assert df["base_weight"].tolist() == [1.0]
    np.testing.assert_allclose(df["abs_weight"], [math.exp(-1)], rtol=2e-6)


def test_expanded_vector_columns(P):
    setup = make_setup(P)
    hits = hits_dict(2, [450.0, 450.0], [1.0, 2.0], [2, 2])
    df = P.processing.assemble_hit_table(hits, 4000, setup, vector_columns=False)
    assert list(df.columns[:9]) == ["pos_x", "pos_y", "pos_z", "dir_x", "dir_y", "dir_z",
                                    "init_dir_x", "init_dir_y", "init_dir_z"]
    assert list(df.columns[9:]) == REF_COLUMNS[3:]
    assert df["module_id"].tolist() == [3, 3]


def test_empty_table_keeps_the_reference_columns(P):
    """calc_total_weight! returns an empty table unchanged (src/processing.jl:
    34-36): no abs_weight or total_weight columns."""
    df = P.processing.assemble_hit_table(hits_dict(0, [], [], []), 4000, make_setup(P))
    assert list(df.columns) == REF_COLUMNS[:9]
    assert len(df) == 0


def test_base_weight_division_semantics(P):
    """Int64 / Int64 in Julia is a Float64 division."""
    bw = P.processing.base_weight
    assert bw(4000, 3999) == 4000 / 3999
    assert bw(4000, 0) == math.inf
    assert math.isnan(bw(0, 0))
    assert bw(2**62, 2**61) == 2.0


def test_module_index_outside_the_target_list(P):
    with pytest.raises(IndexError):
        P.processing.assemble_hit_table(hits_dict(1, [450.0], [1.0], [4]), 10, make_setup(P))


def test_training_data_slices_follow_batch_boundaries(P):
    """generate_training_data labels a slice of the samples only along the
    batch boundaries of the whole run (so that a slice's random streams and
    labels are those of the whole run); other slices are refused before any
    device work."""
    from photonprop import surrogate as S
    st = S.default_setup()
    for first, count in ((1, 2048), (0, 1000), (2048, 9000), (0, 20000)):
        with pytest.raises(ValueError):
            S.generate_training_data(st, 10_000, batch_sources=2048, first=first, count=count,
                                     dev=None, log=None)


def test_label_slices_round_trip(P, tmp_path):
    """save_label_slice / load_label_slices: totals exact, time bins within
    2^-11 of their value (float16 fractions of the total) or 1e-7 of the
    total for the tiniest fractions, empty bins empty, rows of two slices
    placed at their own sample indices."""
    from photonprop import surrogate as S
    rng = np.random.default_rng(4)
    lab = np.zeros((6000, 66), np.float32)
    lab[:, 0] = 1.0
    lab[:, 2:] = rng.gamma(0.3, 50.0, (6000, 64)) * (rng.random((6000, 64)) < 0.4)
    lab[:, 2:] *= np.exp(-rng.uniform(0, 12, (6000, 1)))  # absorption weights down to 6e-6
    lab[:, 1] = lab[:, 2:].sum(1) + rng.uniform(0, 5, 6000)  # hits outside the bins
    stats = dict(photons=123, kernel_ms=1.5, seconds=2.5)
    S.save_label_slice(tmp_path / "a.npz", 0, lab[:4096], stats)
    S.save_label_slice(tmp_path / "b.npz", 4096, lab[4096:], stats)
    out = np.zeros_like(lab)
    cov, st = S.load_label_slices([tmp_path / "b.npz", tmp_path / "a.npz"], out)
    assert cov.all() and st["photons"] == 246 and sorted(st["slices"]) == [[0, 4096], [4096, 6000]]
    assert np.array_equal(out[:, 1], lab[:, 1])
    err = np.abs(out[:, 2:] - lab[:, 2:])
    assert np.all(err <= np.maximum(lab[:, 2:] * 2.0**-11, lab[:, 1:2] * 1e-7))
    assert np.all(out[:, 2:][lab[:, 2:] == 0] == 0)
