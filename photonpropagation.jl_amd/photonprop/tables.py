"""Training tables for surrogates: per-PMT photon counts
(scripts/photon_tables/photon_table_for_sr.jl:97-224), written as Arrow, and
per-photon tables (scripts/photon_tables/photon_tables_photons.jl:30-148),
written as HDF5 with save_hdf! semantics (photonprop/hdf5.py).

Each simulation places one source around a POM at the origin (distance and
energy log-uniform when randomised: 10-300 m, 100 GeV-1 PeV), propagates it
on the GPU, rescales the photon count and repeats when the hit buffer
overflows (x0.1) or fewer than 100 photons arrive (x10), turns the photons
into PMT hits (make_hits_from_photons), applies the QE weights
(calc_pe_weight) and stores the weighted counts per PMT.
"""
from __future__ import annotations

import numpy as np

from .lightyield import (ExtendedCherenkovEmitter, FastLightsabreMuonEmitter, Particle,
                         PEMinus, PHadronShower, PMuMinus, PointlikeCherenkovEmitter,
                         rescale_source)
from .medium import CascadiaMediumProperties
from .pom import POM, calc_pe_weight, make_hits_from_photons
from .processing import PhotonPropOOMException, make_hit_buffers, propagate_photons
from .propsetup import PhotonPropSetup
from .spectrum import make_cherenkov_spectrum

MODES = ("em_shower", "hadronic_shower", "pointlike_cherenkov", "lightsabre")


def rand_vec(rng):
    """Isotropic unit vector (photon_table_for_sr.jl:89-93)."""
    ct = rng.uniform(-1, 1)
    phi = rng.uniform(0, 2 * np.pi)
    st = np.sqrt(1 - ct * ct)
    return np.array([st * np.cos(phi), st * np.sin(phi), ct])


def random_rotation(rng):
    """Uniformly random rotation matrix (QR of a Gaussian matrix)."""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    return q


def make_source(mode, pos, direction, energy, medium, spectrum, rng):
    """make_particle of the script: the emitter of one simulation."""
    if mode in ("em_shower", "hadronic_shower"):
        ptype = PEMinus if mode == "em_shower" else PHadronShower
        p = Particle(np.float32(pos), np.float32(direction), 0.0, float(energy), 0.0, ptype)
        return ExtendedCherenkovEmitter.from_particle(p, medium, spectrum, rng=rng)
    if mode == "pointlike_cherenkov":
        p = Particle(np.float32(pos), np.float32(direction), 0.0, float(energy), 0.0, PEMinus)
        em = ExtendedCherenkovEmitter.from_particle(p, medium, spectrum, rng=rng)
        return PointlikeCherenkovEmitter(em.position, em.direction, 0.0, em.photons)
    if mode == "lightsabre":
        length = 400.0
        p = Particle(np.float32(pos - length / 2 * direction), np.float32(direction), 0.0,
                     float(energy), length, PMuMinus)
        return FastLightsabreMuonEmitter(p, medium, spectrum, rng=rng)
    raise ValueError(f"mode must be one of {MODES}")


def pmt_count_table(outfile, nsims=5, *, mode="em_shower", energy=None, dist=None,
                    randomize=True, seed=1, capacity=1 << 24, max_photons=1e13, log=None):
    """Write `nsims` rows (sim_ix, pos, dir, pos_t, pos_phi, dir_t, dir_phi,
    hits[16], energy, dist) to an Arrow file; returns the rows as a dict."""
    import pyarrow as pa
    if not randomize and (energy is None or dist is None):
        raise ValueError("Need to provide energy and distance when not randomizing.")
    rng = np.random.default_rng(seed)
    target = POM(np.float32([0, 0, 0]), 1)
    medium = CascadiaMediumProperties(0.95, 0.2, 1.0, 1.0)
    spectrum = make_cherenkov_spectrum((300.0, 800.0), medium)
    rows = {k: [] for k in ("sim_ix", "pos", "dir", "pos_t", "pos_phi", "dir_t", "dir_phi",
                            "hits", "energy", "dist")}
    hbc, hbg = make_hit_buffers(capacity=capacity)
    with hbg:
        for isim in range(1, nsims + 1):
            direction = rand_vec(rng).astype(np.float32)
            if randomize:
                d = 10 ** rng.uniform(1, np.log10(300))
                e = 10 ** rng.uniform(2, 6)
            else:
                d, e = float(dist), float(energy)
            pos = (d * rand_vec(rng)).astype(np.float32)
            if mode == "lightsabre":
                # closest approach at distance d (photon_table_for_sr.jl:140-147)
                r = direction[0] / direction[1] if direction[1] != 0 else 0.0
                pos_rot = np.array([d / np.sqrt(1 + r * r), -r * d / np.sqrt(1 + r * r), 0.0])
                R = random_rotation(rng)
                pos = (R @ pos_rot).astype(np.float32)
                direction = (R @ direction).astype(np.float32)
            source = make_source(mode, pos, direction, e, medium, spectrum, rng)
            setup = PhotonPropSetup([source], [target], medium, spectrum, int(seed) * 1000 + isim)
            base_weight, failure, photons = 1.0, -1, None
            while failure != 0:
                if setup.sources[0].photons > max_photons:
                    photons = None
                    break
                if failure == 2:
                    setup.sources[0] = rescale_source(setup.sources[0], 10)
                    base_weight /= 10.0
                elif failure == 1:
                    setup.sources[0] = rescale_source(setup.sources[0], 0.1)
                    base_weight *= 10.0
                try:
                    photons = propagate_photons(setup, hbc, hbg, 15, vector_columns=False,
                                                with_ids=True)
                except PhotonPropOOMException:
                    failure = 1
                    continue
                if len(photons) < 100:
                    failure = 2
                    continue
                failure = 0
            if photons is None:
                if log:
                    log(f"sim {isim}: more than {max_photons:g} photons, skipped")
                continue
            photons["total_weight"] *= base_weight
            hits = make_hits_from_photons(photons, [target], seed=int(seed) * 1000 + isim)
            calc_pe_weight(hits, [target])
            counts = np.bincount(hits["pmt_id"].to_numpy().astype(np.int64) - 1,
                                 weights=hits["total_weight"].to_numpy(), minlength=16)[:16]
            pn = pos / np.linalg.norm(pos)
            rows["sim_ix"].append(isim)
            rows["pos"].append(pos.tolist())
            rows["dir"].append(direction.tolist())
            rows["pos_t"].append(float(np.arccos(np.clip(pn[2], -1, 1))))
            rows["pos_phi"].append(float(np.arctan2(pn[1], pn[0])))
            rows["dir_t"].append(float(np.arccos(np.clip(direction[2], -1, 1))))
            rows["dir_phi"].append(float(np.arctan2(direction[1], direction[0])))
            rows["hits"].append(counts.tolist())
            rows["energy"].append(float(e))
            rows["dist"].append(float(d))
            if log:
                log(f"sim {isim}: E {e:.3g} GeV, d {d:.1f} m, {len(photons)} photons, "
                    f"{len(hits)} PMT hits")
    table = pa.table({k: v for k, v in rows.items()})
    with pa.OSFile(str(outfile), "wb") as sink:
        with pa.ipc.new_file(sink, table.schema) as w:
            w.write_table(table)
    return rows


PHOTON_TABLE_COLUMNS = ("tres", "pos_x", "pos_y", "pos_z", "dir_x", "dir_y", "dir_z",
                        "total_weight", "module_id", "wavelength")


def expected_det_photons_extended(source, target, medium):
    """expected_det_photons_extended (photon_tables_photons.jl:18-28):
    0.1 * 1e9 r^2 / (4 d^2) exp(-d / abs_len(450 nm))."""
    d = float(np.linalg.norm(np.asarray(target.shape.position, np.float64) -
                             np.asarray(source.position, np.float64)))
    att = float(medium.absorption_length(450.0))
    return 0.1 * 1e9 * float(target.shape.radius) ** 2 / (4 * d * d) * np.exp(-d / att)


def photon_table_photons(outfile, energy, distance, dir_costheta, dir_phi, seed, *,
                         mode="extended", g=0.95, abs_scale=1.0, sca_scale=1.0,
                         n_ph_limit=100_000, capacity=1 << 24, max_photons=1e13,
                         hit_buffers=None, log=None):
    """One simulation of run_sim (photon_tables_photons.jl:30-148), appended
    to `outfile` as photons/dataset_<n> (save_hdf!): a source at distance d
    from a module of radius 0.3 m at the origin - a cascade at (0, 0, d)
    ("extended"), or a track whose closest approach is d ("lightsabre_muon",
    "bare_infinite_track", 400 m long); photon counts rescaled until the hit
    buffer holds the result and >= 100 photons arrive; at most n_ph_limit
    photons kept, their weights scaled up; columns tres, position, direction,
    total_weight, module_id, wavelength in float64. The reference calls a
    make_setup that its sources do not define; the setup here is this
    module's (Cascadia water with the given g and scales, the Cherenkov
    spectrum 300-800 nm). Returns the dataset name, or None when more than
    max_photons would be needed."""
    import json

    from .calc import calc_time_residual
    from .detection import spherical_target
    from .lightyield import CherenkovTrackEmitter

    direction = np.float32([np.sqrt(max(0.0, 1 - dir_costheta ** 2)) * np.cos(dir_phi),
                            np.sqrt(max(0.0, 1 - dir_costheta ** 2)) * np.sin(dir_phi),
                            dir_costheta])
    if mode in ("bare_infinite_track", "lightsabre_muon"):
        r = direction[0] / direction[1] if direction[1] != 0 else 0.0
        ppos = np.float32([distance / np.sqrt(1 + r * r), -r * distance / np.sqrt(1 + r * r), 0])
    elif mode == "extended":
        ppos = np.float32([0, 0, distance])
    else:
        raise ValueError("mode must be extended, bare_infinite_track or lightsabre_muon")
    sim_attrs = {"energy": float(energy), "mode": str(mode), "distance": float(distance),
                 "dir_costheta": float(dir_costheta), "dir_phi": float(dir_phi),
                 "seed": int(seed), "source_pos": json.dumps([float(x) for x in ppos]),
                 "g": float(g), "abs_scale": float(abs_scale), "sca_scale": float(sca_scale)}
    medium = CascadiaMediumProperties(g, 0.2, abs_scale, sca_scale)
    spectrum = make_cherenkov_spectrum((300.0, 800.0), medium)
    target = spherical_target((0.0, 0.0, 0.0), 0.3, 1)
    rng = np.random.default_rng(seed)
    if mode == "extended":
        source = make_source("em_shower", ppos, direction, energy, medium, spectrum, rng)
    elif mode == "lightsabre_muon":
        source = make_source("lightsabre", ppos, direction, energy, medium, spectrum, rng)
    else:  # bare track: Cherenkov light of the muon alone, 400 m through ppos
        length = 400.0
        source = CherenkovTrackEmitter.from_particle(
            Particle(np.float32(ppos - length / 2 * direction), direction, 0.0, float(energy),
                     length, PMuMinus), medium, spectrum, rng=rng)
    setup = PhotonPropSetup([source], [target], medium, spectrum, int(seed))
    base_weight = 1.0
    if mode == "extended":
        exp_ph = expected_det_photons_extended(setup.sources[0], target, medium)
        if exp_ph > 1e5:
            factor = 1e5 / exp_ph
            setup.sources[0] = rescale_source(setup.sources[0], factor)
            base_weight /= factor
    own = hit_buffers is None
    hbc, hbg = make_hit_buffers(capacity=capacity) if own else hit_buffers
    try:
        failure, photons = -1, None
        while failure != 0:
            if setup.sources[0].photons > max_photons:
                if log:
                    log(f"more than {max_photons:g} photons, skipping")
                return None
            if failure == 2:
                setup.sources[0] = rescale_source(setup.sources[0], 10)
                base_weight /= 10.0
            elif failure == 1:
                setup.sources[0] = rescale_source(setup.sources[0], 0.1)
                base_weight *= 10.0
            try:
                photons = propagate_photons(setup, hbc, hbg, 15, vector_columns=False)
            except PhotonPropOOMException:
                failure = 1
                continue
            if len(photons) < 100:
                failure = 2
                continue
            failure = 0
    finally:
        if own:
            hbg.close()
    nph_sim = len(photons)
    if nph_sim > n_ph_limit:
        photons = photons.iloc[:n_ph_limit].copy()
        base_weight *= nph_sim / n_ph_limit
    calc_time_residual(photons, setup)
    photons["total_weight"] = photons["total_weight"].to_numpy() * base_weight
    table = np.column_stack([photons[c].to_numpy(np.float64) for c in PHOTON_TABLE_COLUMNS])
    from .hdf5 import save_hdf
    name = save_hdf(str(outfile), "photons", table, sim_attrs)
    if log:
        log(f"{name}: E {energy:.3g} GeV, d {distance:.1f} m, {nph_sim} photons")
    return name

