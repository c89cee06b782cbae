"""Synthetic workloads used by the benchmark and the parity tests.

readme_point_source  One PointlikeIsotropicEmitter at the origin, t = 0;
                     monochromatic 450 nm; one sphere r = 0.3 m at (0,0,10);
                     Cascadia water (g = 0.95); README.md:22-62 of the reference.
cascade_16_strings   4 x 4 strings at 50 m, 20 modules per string at 10 m
                     (320 spheres, r = 0.3 m); one ExtendedCherenkovEmitter of a
                     10 TeV electron at a random interior point with an
                     isotropic axis; Cherenkov spectrum 300-700 nm weighted by
                     the PMT acceptance (600-knot inverse CDF).
uneven_batch_1000    10 x 10 strings at 50 m, 10 modules per string (1000
                     spheres); 256 sources (70% cascades, 30% 100 m tracks),
                     photon counts log-uniform in [1e4, 1e7].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .detection import make_detector_grid, spherical_target
from .lightyield import (CherenkovTrackEmitter, ExtendedCherenkovEmitter,
                         LongitudinalParameterisation, PEMinus, PointlikeIsotropicEmitter)
from .medium import CascadiaMediumProperties
from .spectrum import (Monochromatic, make_biased_cherenkov_spectrum, make_cuda_spectral_dist,
                       pmt_acceptance_curve)


@dataclass
class Workload:
    name: str
    sources: list
    targets: list
    medium: object
    spectrum: object  # Monochromatic or DeviceSpectralDist
    seed: int
    n_steps: int = 15


def isotropic_unit(rng):
    ct = rng.uniform(-1, 1)
    phi = rng.uniform(0, 2 * np.pi)
    st = np.sqrt(1 - ct * ct)
    return np.float32([st * np.cos(phi), st * np.sin(phi), ct])


def readme_point_source(photons=100_000, seed=1):
    medium = CascadiaMediumProperties(0.95, 0.2, 1.0, 1.0)
    src = PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, int(photons))
    tgt = [spherical_target((0.0, 0.0, 10.0), 0.3, 1)]
    return Workload("readme_point_source", [src], tgt, medium, Monochromatic(450.0), seed)


_SPEC_CACHE = {}


def acceptance_weighted_spectrum(medium, wl_range=(300.0, 700.0), knots=600):
    key = (wl_range, knots)
    if key not in _SPEC_CACHE:
        spec = make_biased_cherenkov_spectrum(pmt_acceptance_curve(), wl_range, medium)
        _SPEC_CACHE[key] = make_cuda_spectral_dist(spec.spectral_dist, wl_range, knots)
    return _SPEC_CACHE[key]


def cascade_params(energy_gev, medium):
    return LongitudinalParameterisation.from_energy(energy_gev, medium, PEMinus)


def cascade_16_strings(photons=100_000_000, seed=42):
    rng = np.random.default_rng(seed)
    medium = CascadiaMediumProperties()
    targets = make_detector_grid(4, 4, 50.0, 20, 10.0, radius=0.3)
    pos = np.float32([rng.uniform(-75, 75), rng.uniform(-75, 75), rng.uniform(-95, 95)])
    src = ExtendedCherenkovEmitter(pos, isotropic_unit(rng), 0.0, int(photons),
                                   cascade_params(1e4, medium))
    return Workload("cascade_16_strings", [src], targets, medium,
                    acceptance_weighted_spectrum(medium), seed)


def uneven_batch_1000(n_sources=256, seed=7, min_photons=1e4, max_photons=1e7):
    rng = np.random.default_rng(seed)
    medium = CascadiaMediumProperties()
    targets = make_detector_grid(10, 10, 50.0, 10, 10.0, radius=0.3)
    half = np.float32([225.0 + 50.0, 225.0 + 50.0, 45.0 + 50.0])
    sources = []
    for _ in range(n_sources):
        pos = rng.uniform(-half, half).astype(np.float32)
        d = isotropic_unit(rng)
        n = int(10 ** rng.uniform(np.log10(min_photons), np.log10(max_photons)))
        if rng.random() < 0.7:
            e = 10 ** rng.uniform(3, 5)
            sources.append(ExtendedCherenkovEmitter(pos, d, 0.0, n, cascade_params(e, medium)))
        else:
            sources.append(CherenkovTrackEmitter(pos - 50.0 * d, d, 0.0, 100.0, n))
    return Workload("uneven_batch_1000", sources, targets, medium,
                    acceptance_weighted_spectrum(medium), seed)
