"""Media: CascadiaMediumProperties and SimpleMediumProperties.

Host-side mirror of src/water_properties.jl and src/simple_medium.jl. The
per-wavelength functions (float64 numpy) serve the host side of the path:
absorption weights (src/processing.jl:32-44), spectra and light yields. On
the device the same quantities are recomputed per photon by the HIP kernel.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib

C_VAC_M_NS = 0.299792458  # src/medium.jl:10

# src/water_properties.jl:33-41
CASCADIA_ES_B = 0.835
CASCADIA_VOL_CONC_SMALL_PART = 7.5e-3
CASCADIA_VOL_CONC_LARGE_PART = 7.5e-3
CASCADIA_HG_G = 0.95
CASCADIA_SALINITY = 34.82  # permille
CASCADIA_TEMPERATURE = 1.8  # deg C
CASCADIA_PRESSURE = 265.917473476  # atm
CASCADIA_RADIATION_LENGTH = 36.08  # g/cm^2

# Quan & Fry (1995) coefficients n0..n9 and the linear pressure coefficient
QUAN_FRY_N = (1.31405, 1.779e-4, -1.05e-6, 1.6e-8, -2.02e-6, 15.868, 0.01155, -0.00423,
              -4382.0, 1.1455e6)
QUAN_FRY_PRESSURE = 1.45e-5  # per atm

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def _load_straw():
    a = np.loadtxt(os.path.join(_DATA, "straw_abs_len.csv"), delimiter=",", skiprows=1)
    return a[:, 0].copy(), a[:, 1].copy()


# STRAW attenuation fit, 101 knots 300-800 nm (src/water_properties.jl:29-31)
STRAW_WL, STRAW_ABS_LEN = _load_straw()


def dippr105(temperature_k, A=0.14395, B=0.0112, C=649.727, D=0.05107):
    """DIPPR equation 105 density of water (kg/m^3) at temperature in K."""
    return A / B ** (1.0 + (1.0 - temperature_k / C) ** D)


CASCADIA_DENSITY = dippr105(CASCADIA_TEMPERATURE + 273.15)


def quan_fry_n(wl, salinity, temperature, pressure):
    """Quan & Fry phase refractive index (wl in nm)."""
    n = QUAN_FRY_N
    wl = np.asarray(wl, dtype=np.float64)
    S, T = salinity, temperature
    return (n[0] + (n[1] + n[2] * T + n[3] * T * T) * S + n[4] * T * T
            + (n[5] + n[6] * S + n[7] * T) / wl + n[8] / wl**2 + n[9] / wl**3
            + QUAN_FRY_PRESSURE * pressure)


def quan_fry_group_velocity(wl, salinity, temperature, pressure):
    """c / (n - wl dn/dwl) in m/ns."""
    n = QUAN_FRY_N
    wl = np.asarray(wl, dtype=np.float64)
    S, T = salinity, temperature
    dndl = -(n[5] + n[6] * S + n[7] * T) / wl**2 - 2 * n[8] / wl**3 - 3 * n[9] / wl**4
    return C_VAC_M_NS / (quan_fry_n(wl, S, T, pressure) - wl * dndl)


def kopelevich_scattering_length(wl, vol_conc_small, vol_conc_large):
    """Kopelevich scattering length in m (wl in nm)."""
    x = np.asarray(wl, dtype=np.float64) / 550.0
    b = 0.0017 * x**-4.3 + 1.34 * vol_conc_small * x**-1.7 + 0.312 * vol_conc_large * x**-0.3
    return 1.0 / b


class MediumProperties:
    """Abstract medium (src/medium.jl)."""

    def to_pod(self) -> _lib.pp_medium:
        raise NotImplementedError


@dataclass
class CascadiaMediumProperties(MediumProperties):
    """CascadiaMediumProperties(hg_g, pure_water_fraction, abs_scale, sca_scale)
    (src/water_properties.jl:43-82; default constructor :82).

    ``CascadiaMediumProperties(0.95)`` is also accepted (README form)."""
    hg_g: float = 0.95
    pure_water_fraction: float = 0.2
    abs_scale: float = 1.0
    sca_scale: float = 1.0
    es_b: float = field(default=CASCADIA_ES_B)
    vol_conc_small: float = field(default=CASCADIA_VOL_CONC_SMALL_PART)
    vol_conc_large: float = field(default=CASCADIA_VOL_CONC_LARGE_PART)
    salinity: float = field(default=CASCADIA_SALINITY)
    temperature: float = field(default=CASCADIA_TEMPERATURE)
    pressure: float = field(default=CASCADIA_PRESSURE)

    @property
    def hg_fraction(self):
        return 1.0 - self.pure_water_fraction

    def scattering_length(self, wl):  # :95-100
        return kopelevich_scattering_length(wl, self.vol_conc_small,
                                            self.vol_conc_large) * self.sca_scale

    def absorption_length(self, wl):  # :88-93, linear in the STRAW table, clamped
        return np.interp(np.asarray(wl, dtype=np.float64), STRAW_WL,
                         STRAW_ABS_LEN) * self.abs_scale

    def phase_refractive_index(self, wl):
        return quan_fry_n(wl, self.salinity, self.temperature, self.pressure)

    def group_velocity(self, wl):
        return quan_fry_group_velocity(wl, self.salinity, self.temperature, self.pressure)

    def cherenkov_angle(self, wl):
        return np.arccos(1.0 / self.phase_refractive_index(wl))

    def mean_scattering_angle(self):
        return self.hg_g

    def radiation_length(self):
        return CASCADIA_RADIATION_LENGTH

    def material_density(self):
        return dippr105(self.temperature + 273.15)

    def to_pod(self):
        m = _lib.pp_medium()
        m.kind = _lib.MEDIUM_CASCADIA
        m.phase_kind = _lib.PHASE_MIXED_HG_ES
        m.hg_g = self.hg_g
        m.es_b = self.es_b
        m.hg_fraction = np.float32(1.0) - np.float32(self.pure_water_fraction)
        m.sca_scale = self.sca_scale
        m.abs_scale = self.abs_scale
        m.vol_conc_small = self.vol_conc_small
        m.vol_conc_large = self.vol_conc_large
        m.salinity = self.salinity
        m.temperature = self.temperature
        m.pressure = self.pressure
        return m


@dataclass
class SimpleMediumProperties(MediumProperties):
    """SimpleMediumProperties (src/simple_medium.jl:3-15): constant lengths,
    HG phase function with g = mean_scattering_angle."""
    absorption_length_: float
    scattering_length_: float
    phase_refractive_index_: float
    mean_scattering_angle_: float
    group_velocity_: float

    def scattering_length(self, wl):
        return np.full(np.shape(wl), self.scattering_length_, dtype=np.float64)

    def absorption_length(self, wl):
        return np.full(np.shape(wl), self.absorption_length_, dtype=np.float64)

    def phase_refractive_index(self, wl):
        return np.full(np.shape(wl), self.phase_refractive_index_, dtype=np.float64)

    def group_velocity(self, wl):
        return np.full(np.shape(wl), self.group_velocity_, dtype=np.float64)

    def cherenkov_angle(self, wl):
        return np.arccos(1.0 / self.phase_refractive_index(wl))

    def mean_scattering_angle(self):
        return self.mean_scattering_angle_

    def radiation_length(self):
        return CASCADIA_RADIATION_LENGTH

    def material_density(self):
        return 1000.0

    def to_pod(self):
        m = _lib.pp_medium()
        m.kind = _lib.MEDIUM_SIMPLE
        m.phase_kind = _lib.PHASE_HG
        m.hg_g = self.mean_scattering_angle_
        m.hg_fraction = 1.0
        m.sca_scale = 1.0
        m.abs_scale = 1.0
        m.sca_len = self.scattering_length_
        m.group_velocity = self.group_velocity_
        m.phase_refractive_index = self.phase_refractive_index_
        return m


def ice_phase_refractive_index(wl):
    """src/ice_properties.jl:105-108 (wl in nm)."""
    x = np.asarray(wl, np.float64) / 1000.0
    return 1.55749 - 1.57988 * x + 3.99993 * x**2 - 4.68271 * x**3 + 2.09354 * x**4


def ice_group_refractive_index(wl):
    """src/ice_properties.jl:117-121."""
    x = np.asarray(wl, np.float64) / 1000.0
    return ice_phase_refractive_index(wl) * (1 + 0.227106 - 0.954648 * x + 1.42568 * x**2
                                             - 0.711832 * x**3)


@dataclass
class HomogenousIceProperties(MediumProperties):
    """HomogenousIceProperties (src/ice_properties.jl:26-177): SPICE-model
    homogeneous ice; phase function: HG + simplified Liu (PPC)."""
    radiation_length_: float
    mean_scattering_angle_: float
    hg_fraction: float
    A_SPICE: float
    B_SPICE: float
    D_SPICE: float
    E_SPICE: float
    a_dust_400: float
    b_dust_400: float
    alpha_sca_dust: float
    kappa_abs_dust: float
    deltaTSPICE: float
    abs_scale: float = 1.0
    sca_scale: float = 1.0

    def phase_refractive_index(self, wl):
        return ice_phase_refractive_index(wl)

    def group_velocity(self, wl):
        return C_VAC_M_NS / ice_group_refractive_index(wl)

    def cherenkov_angle(self, wl):
        return np.arccos(1.0 / self.phase_refractive_index(wl))

    def absorption_length(self, wl):  # :137-157
        wl = np.asarray(wl, np.float64)
        adust = self.a_dust_400 * (wl / 400.0) ** (-self.kappa_abs_dust)
        atemp = self.A_SPICE * np.exp(-self.B_SPICE / wl) * (1 + 0.01 * self.deltaTSPICE)
        return 1.0 / (adust + atemp) * self.abs_scale

    def scattering_length(self, wl):  # :164-168
        eff = self.b_dust_400 * (np.asarray(wl, np.float64) / 400.0) ** (-self.alpha_sca_dust)
        return 1.0 / (eff / (1.0 - self.mean_scattering_angle_)) * self.sca_scale

    def mean_scattering_angle(self):
        return self.mean_scattering_angle_

    def radiation_length(self):
        return self.radiation_length_

    def material_density(self):
        return 0.910 * 1000

    def to_pod(self):
        m = _lib.pp_medium()
        m.kind = _lib.MEDIUM_ICE
        m.phase_kind = _lib.PHASE_MIXED_HG_SL
        m.hg_g = self.mean_scattering_angle_
        m.hg_fraction = self.hg_fraction
        m.sca_scale = self.sca_scale
        m.abs_scale = self.abs_scale
        m.ice_b_dust_400 = self.b_dust_400
        m.ice_alpha_sca_dust = self.alpha_sca_dust
        return m


def make_homogenous_clearice_properties(abs_scale=1.0, sca_scale=1.0):
    """The homogeneous clear-ice model (src/ice_properties.jl:50-65)."""
    return HomogenousIceProperties(
        radiation_length_=39.652 / 0.9216, mean_scattering_angle_=0.9, hg_fraction=0.45,
        A_SPICE=6954.090332031250, B_SPICE=6617.754394531250, D_SPICE=0.0, E_SPICE=0.0,
        a_dust_400=0.006350013, b_dust_400=0.02331206666666667,
        alpha_sca_dust=0.898608505726, kappa_abs_dust=1.084106802940,
        deltaTSPICE=12.115215000000001, abs_scale=abs_scale, sca_scale=sca_scale)


def make_simple_medium(absorption_length, scattering_length, phase_refractive_index,
                       mean_scattering_angle, group_velocity):
    """Keyword-style constructor mirroring Base.@kwdef SimpleMediumProperties."""
    return SimpleMediumProperties(absorption_length, scattering_length, phase_refractive_index,
                                  mean_scattering_angle, group_velocity)
