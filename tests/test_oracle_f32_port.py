"""The Float32 port of the reference kernel (oracle/pp_ref_f32.c, the CPU
baseline) against the oracle: same random streams, so the same photons hit
the same modules; only Float32 rounding separates them. This pins the port as
a restatement of the same algorithm, so that its timing is a fair baseline."""
import numpy as np
import pytest
from scipy import stats


def both(O, P, wl, photons):
    args = (P.processing.sources_to_pod(wl.sources), wl.medium.to_pod(),
            P.targets_to_pod(wl.targets), wl.seed)
    spec = wl.spectrum
    kw = dict(n_steps=wl.n_steps)
    if isinstance(spec, P.Monochromatic):
        kw["mono_wl"] = spec.wavelength
    else:
        kw.update(table=spec.table, mapping=spec.mapping)
    ref, nref = O.propagate(*args, **kw)
    f32, nf, steps = O.propagate_f32(*args, **kw)
    assert nref == nf == photons
    return ref, f32, steps


def compare(ref, f32):
    ri = {int(p): i for i, p in enumerate(ref["photon_id"])}
    fi = {int(p): i for i, p in enumerate(f32["photon_id"])}
    common = sorted(set(ri) & set(fi))
    assert len(common) >= 0.97 * max(len(ri), len(fi)), (len(ri), len(fi), len(common))
    a = ref[[ri[p] for p in common]]
    b = f32[[fi[p] for p in common]]
    assert np.mean(a["module_ix"] == b["module_ix"]) > 0.999
    assert np.mean(a["n_steps"] == b["n_steps"]) > 0.999
    close = np.linalg.norm(a["pos"] - b["pos"], axis=1) < 5e-3 + 2e-5 * a["dist"]
    assert close.mean() > 0.97, close.mean()
    np.testing.assert_array_equal(a["wavelength"], b["wavelength"])
    return len(common)


def test_readme_point_source(O, P):
    wl = P.workloads.readme_point_source(photons=400_000)
    ref, f32, steps = both(O, P, wl, 400_000)
    assert len(ref) > 50
    compare(ref, f32)
    assert steps == 400_000 * 15 - int(np.sum(15 - f32["n_steps"]))


def test_cascade_event(O, P):
    wl = P.workloads.cascade_16_strings(photons=300_000)
    ref, f32, _ = both(O, P, wl, 300_000)
    assert len(ref) > 100
    compare(ref, f32)
    assert stats.ks_2samp(ref["time"], f32["time"]).pvalue > 1e-3


def test_uncovered_kinds_are_refused(O, P):
    src = [P.PencilEmitter(np.zeros(3), np.float32([0, 0, 1]), 0.0, 0.0, 100)]
    with pytest.raises(ValueError):
        O.propagate_f32(P.processing.sources_to_pod(src), P.CascadiaMediumProperties().to_pod(),
                        P.targets_to_pod([P.spherical_target((0, 0, 5), 0.3, 1)]), 1,
                        mono_wl=450.0)
