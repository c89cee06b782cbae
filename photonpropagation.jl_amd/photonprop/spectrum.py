"""Spectra: Monochromatic and inverse-CDF interpolated spectra.

Mirror of src/spectrum.jl. ``make_spectral_dist`` builds the CDF on a 1 nm
(biased: 0.5 nm) grid with 25-point Gauss-Legendre quadrature per knot
(:41-62); ``make_cuda_spectral_dist`` samples the inverse CDF at
``interp_steps`` equally spaced u in [0, 1] (:75-83) - the table the HIP
kernel keeps in LDS.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Callable

import numpy as np

from . import _lib

FINE_STRUCTURE = 7.2973525693e-3


def integrate_gauss_quad(f, a, b, n=25):
    """Gauss-Legendre quadrature of f over [a, b] with n nodes."""
    x, w = np.polynomial.legendre.leggauss(n)
    a = float(a)
    b = float(b)
    xs = 0.5 * (b - a) * x + 0.5 * (a + b)
    return 0.5 * (b - a) * float(np.sum(w * np.asarray(f(xs), dtype=np.float64)))


def frank_tamm(wavelength, ref_index):
    """Frank-Tamm photon yield density 2*pi*alpha*(1 - 1/n^2)/wl^2 (wl in nm);
    multiplied by 1e9 in the spectra this is photons per m of track per nm."""
    wl = np.asarray(wavelength, dtype=np.float64)
    n = np.asarray(ref_index, dtype=np.float64)
    return 2.0 * np.pi * FINE_STRUCTURE * (1.0 - 1.0 / n**2) / wl**2


class SpectralDist:
    """Abstract spectral distribution."""


@dataclass
class Monochromatic(SpectralDist):
    wavelength: float

    def sample(self, u):
        return np.full(np.shape(u), self.wavelength, dtype=np.float32)


@dataclass
class InterpolatedSpectralDist(SpectralDist):
    """Linear interpolant cdf -> wavelength. Beyond the last knot the
    interpolant is held at the last wavelength (the reference's zero
    extrapolation would yield 0 nm there, src/spectrum.jl:59)."""
    cdf_knots: np.ndarray
    wavelengths: np.ndarray
    normalization: float

    def __call__(self, u):
        u = np.asarray(u, dtype=np.float64)
        return np.interp(u, self.cdf_knots, self.wavelengths)

    def sample(self, u):
        return self(u).astype(np.float32)


@dataclass
class DeviceSpectralDist(SpectralDist):
    """Inverse-CDF knots at u_k = k/(n-1), as uploaded to the device."""
    table: np.ndarray
    normalization: float
    mapping: int = _lib.SPEC_KNOTS


@dataclass
class Spectrum:
    """Spectrum(spectral_dist, spectrum, wl_range)   src/spectrum.jl:85-89"""
    spectral_dist: SpectralDist
    spectrum: Callable
    wl_range: tuple


def _deduplicate_knots(knots: np.ndarray) -> np.ndarray:
    """Move repeated knots up by one float32 ulp so they become strictly
    increasing (Interpolations.deduplicate_knots!(move_knots=true))."""
    k = knots.astype(np.float32).copy()
    for i in range(1, len(k)):
        if k[i] <= k[i - 1]:
            k[i] = np.nextafter(k[i - 1], np.float32(np.inf))
    return k


def make_spectral_dist(spect_func, wl_range, step_size=1.0):
    """src/spectrum.jl:41-62"""
    lo, hi = float(wl_range[0]), float(wl_range[1])
    wl_steps = np.arange(lo, hi + 0.5 * step_size, step_size, dtype=np.float64)
    wl_steps = wl_steps[wl_steps <= hi + 1e-9]
    full_norm = np.float32(integrate_gauss_quad(spect_func, lo, hi, 25))
    norms = np.zeros(len(wl_steps), dtype=np.float32)
    for i in range(1, len(wl_steps)):
        norms[i] = integrate_gauss_quad(spect_func, lo, wl_steps[i], 25) / float(full_norm)
    norms = _deduplicate_knots(norms)
    order = np.argsort(norms, kind="stable")
    return InterpolatedSpectralDist(norms[order].astype(np.float64),
                                    wl_steps[order].astype(np.float32).astype(np.float64),
                                    float(full_norm))


def make_cuda_spectral_dist(spec: InterpolatedSpectralDist, wl_range=None, interp_steps=30,
                            mapping=_lib.SPEC_KNOTS):
    """src/spectrum.jl:75-83: evaluate the inverse CDF at interp_steps knots."""
    eval_knots = np.linspace(0.0, 1.0, int(interp_steps))
    table = spec(eval_knots).astype(np.float32)
    return DeviceSpectralDist(table, spec.normalization, mapping)


make_device_spectral_dist = make_cuda_spectral_dist


def make_cherenkov_spectrum(wl_range, medium):
    """src/spectrum.jl:91-95"""
    def sfunc(wl):
        return frank_tamm(wl, medium.phase_refractive_index(wl)) * 1e9
    return Spectrum(make_spectral_dist(sfunc, wl_range), sfunc, tuple(wl_range))


def make_biased_cherenkov_spectrum(bias_function, wl_range, medium):
    """src/spectrum.jl:97-101 (0.5 nm grid)"""
    def sfunc(wl):
        return frank_tamm(wl, medium.phase_refractive_index(wl)) * bias_function(wl) * 1e9
    return Spectrum(make_spectral_dist(sfunc, wl_range, 0.5), sfunc, tuple(wl_range))


def make_monochromatic_spectrum(wl):
    """src/spectrum.jl:103-105"""
    return Spectrum(Monochromatic(float(wl)), lambda w: np.where(np.asarray(w) == wl, 1.0, 0.0),
                    (0.0, 1.0))


def pmt_acceptance_curve():
    """Relative PMT wavelength acceptance (assets/PMTAcc.csv of the reference),
    linearly interpolated, zero outside the table."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "pmt_acc.csv")
    a = np.loadtxt(path, delimiter=",", skiprows=1)
    wl, acc = a[:, 0].copy(), a[:, 1].copy()

    def f(w):
        return np.interp(np.asarray(w, dtype=np.float64), wl, acc, left=0.0, right=0.0)
    return f
