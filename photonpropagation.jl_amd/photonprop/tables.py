"""Per-PMT photon-count tables for surrogate training
(scripts/photon_tables/photon_table_for_sr.jl:97-224), written as Arrow.

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
