"""Autotuner host logic: cache keys and configuration round trips (no GPU)."""
import numpy as np


def test_workload_key_is_stable_and_sensitive(P):
    wl = P.workloads.uneven_batch_1000(n_sources=20, seed=1)
    key = lambda w, **kw: P.autotune.workload_key(  # noqa: E731
        kw.get("arch", "gfx950"), 256, w.sources, w.targets, w.medium, w.spectrum,
        kw.get("n_steps", 15))
    k0 = key(wl)
    assert k0 == key(P.workloads.uneven_batch_1000(n_sources=20, seed=1))
    assert k0 != key(wl, arch="gfx942")
    assert k0 != key(wl, n_steps=16)
    wl2 = P.workloads.uneven_batch_1000(n_sources=20, seed=1)
    wl2.sources[0].photons *= 16  # photon-count mix changes the key
    assert k0 != key(wl2)
    wl3 = P.workloads.uneven_batch_1000(n_sources=20, seed=1)
    wl3.targets = wl3.targets[:-1]
    assert k0 != key(wl3)


def test_config_dict_round_trip(P):
    c = P._lib.pp_config()
    c.block_size, c.chunk_photons, c.grid_cell, c.batch_mode = 512, 256, 37.5, 2
    d = P.autotune.config_to_dict(c)
    c2 = P.autotune.apply_dict(P._lib.pp_config(), d)
    assert bytes(c2) == bytes(c)
    assert np.isclose(d["grid_cell"], 37.5)
