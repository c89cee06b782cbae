"""Emitters (PhotonSource subtypes) and the host-side light-yield physics.

Field layouts follow src/lightyield.jl:503-804; each emitter lowers to the
tagged ``pp_source`` POD the HIP kernel reads. Light yields follow
src/lightyield.jl:142-435 and :849-873.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np

from . import _lib

# ---- particles ---------------------------------------------------------------------


class ParticleType:
    name = "particle"


class PEMinus(ParticleType):
    name = "e-"


class PEPlus(ParticleType):
    name = "e+"


class PGamma(ParticleType):
    name = "gamma"


class PHadronShower(ParticleType):
    name = "hadron_shower"


class PMuMinus(ParticleType):
    name = "mu-"


class PMuPlus(ParticleType):
    name = "mu+"


CASCADE_TYPES = (PEMinus, PEPlus, PGamma, PHadronShower)
TRACK_TYPES = (PMuMinus, PMuPlus)


@dataclass
class Particle:
    position: np.ndarray
    direction: np.ndarray
    time: float
    energy: float
    length: float
    type: type

    def shape(self):
        return "cascade" if self.type in CASCADE_TYPES else "track"


# ---- longitudinal profile and track length (src/lightyield.jl:142-245) --------------

@dataclass(frozen=True)
class LongitudinalParametersBase:
    alpha: float
    beta: float
    b: float


LONG_PARAMS = {
    PEMinus: LongitudinalParametersBase(2.01849, 1.45469, 0.63207),
    PEPlus: LongitudinalParametersBase(2.00035, 1.45501, 0.63008),
    PGamma: LongitudinalParametersBase(2.83923, 1.34031, 0.64526),
    PHadronShower: LongitudinalParametersBase(1.58357, 0.96448, 0.33833),
}


@dataclass(frozen=True)
class CherenkovTrackLengthParameters:
    alpha: float
    beta: float
    alpha_dev: float
    beta_dev: float


TRACK_LEN_PARAMS = {
    PEMinus: CherenkovTrackLengthParameters(5.3207078881, 1.00000211, 0.0578170887, 0.5),
    PEPlus: CherenkovTrackLengthParameters(5.3211320598, 0.99999254, 0.0573419669, 0.5),
    PGamma: CherenkovTrackLengthParameters(5.3208540905, 0.99999877, 0.0578170887, 5.66586567),
    PHadronShower: CherenkovTrackLengthParameters(3.3355182722, 1.03662217, 1.1920455395,
                                                  0.80772057),
}


def get_longitudinal_params(ptype, energy):
    p = LONG_PARAMS[ptype]
    return p.alpha + p.beta * math.log10(energy), p.b


@dataclass
class LongitudinalParameterisation:
    a: float
    b: float
    lrad: float

    @classmethod
    def from_energy(cls, energy, medium, ptype):
        """src/lightyield.jl:187-197: a = alpha + beta log10 E, lrad = X0/rho*10."""
        a, b = get_longitudinal_params(ptype, energy)
        lrad = medium.radiation_length() / medium.material_density() * 10.0
        return cls(float(np.float32(a)), float(np.float32(b)), float(np.float32(lrad)))


def cascade_cherenkov_track_length(energy, ptype):
    p = TRACK_LEN_PARAMS[ptype]
    return p.alpha * energy ** p.beta


def rel_additional_track_length(ref_index, energy):
    """src/lightyield.jl:379-399"""
    xs = (1.30, 1.36)
    lam = (0.1842, 0.1916)
    kap = (0.0204, 0.0207)
    sl = (lam[1] - lam[0]) / (xs[1] - xs[0])
    sk = (kap[1] - kap[0]) / (xs[1] - xs[0])
    return (lam[0] + (ref_index - xs[0]) * sl) + (kap[0] + (ref_index - xs[0]) * sk) * np.log(
        energy)


def total_lightyield(particle, medium, spectrum):
    """src/lightyield.jl:401-435"""
    from .spectrum import integrate_gauss_quad
    if particle.shape() == "cascade":
        return spectrum.spectral_dist.normalization * cascade_cherenkov_track_length(
            particle.energy, particle.type)

    def integrand(wl):
        n = medium.phase_refractive_index(wl)
        return spectrum.spectrum(wl) * (1.0 + rel_additional_track_length(n, particle.energy))
    return integrate_gauss_quad(integrand, spectrum.wl_range[0], spectrum.wl_range[1]) * \
        particle.length


def _rng(rng):
    return rng if rng is not None else np.random.default_rng()


def pois_rand(lam, rng=None):
    return int(_rng(rng).poisson(lam))


# ---- emitters -------------------------------------------------------------------------

def _vec(v):
    return np.asarray(v, dtype=np.float32).reshape(3)


class PhotonSource:
    kind = -1

    def to_pod(self) -> _lib.pp_source:
        s = _lib.pp_source()
        s.kind = self.kind
        s.photons = int(self.photons)
        pos = getattr(self, "position", None)
        if pos is not None:
            s.position[:] = [float(x) for x in _vec(pos)]
        d = getattr(self, "direction", None)
        if d is not None:
            s.direction[:] = [float(x) for x in _vec(d)]
        else:
            s.direction[:] = [0.0, 0.0, 1.0]
        return s


@dataclass
class PointlikeIsotropicEmitter(PhotonSource):
    position: np.ndarray
    time: float
    photons: int
    kind = _lib.SRC_POINTLIKE_ISOTROPIC

    def to_pod(self):
        s = super().to_pod()
        s.time = self.time
        return s


@dataclass
class PointlikeTimeRangeEmitter(PhotonSource):
    position: np.ndarray
    time_range: tuple
    photons: int
    kind = _lib.SRC_POINTLIKE_TIME_RANGE

    def to_pod(self):
        s = super().to_pod()
        s.time, s.time_end = float(self.time_range[0]), float(self.time_range[1])
        return s


@dataclass
class AxiconeEmitter(PhotonSource):
    position: np.ndarray
    direction: np.ndarray
    time: float
    photons: int
    angle: float
    beam_divergence: float
    kind = _lib.SRC_AXICONE

    def to_pod(self):
        s = super().to_pod()
        s.time, s.angle, s.beam_divergence = self.time, self.angle, self.beam_divergence
        return s


@dataclass
class CollimatedIsotropicEmitter(PhotonSource):
    position: np.ndarray
    time: float
    photons: int
    cut_angle_cos: float
    kind = _lib.SRC_COLLIMATED_ISOTROPIC

    def to_pod(self):
        s = super().to_pod()
        s.time, s.cut_angle_cos = self.time, self.cut_angle_cos
        return s


@dataclass
class PencilEmitter(PhotonSource):
    position: np.ndarray
    direction: np.ndarray
    beam_divergence: float
    time: float
    photons: int
    kind = _lib.SRC_PENCIL

    def to_pod(self):
        s = super().to_pod()
        s.time, s.beam_divergence = self.time, self.beam_divergence
        return s


class CherenkovEmitter(PhotonSource):
    pass


def _cascade_pod(self, s):
    s.time = self.time
    s.long_a = self.long_param.a
    s.long_b = self.long_param.b
    s.long_lrad = self.long_param.lrad
    return s


@dataclass
class ExtendedCherenkovEmitter(CherenkovEmitter):
    position: np.ndarray
    direction: np.ndarray
    time: float
    photons: int
    long_param: LongitudinalParameterisation
    kind = _lib.SRC_EXTENDED_CHERENKOV

    def to_pod(self):
        return _cascade_pod(self, super().to_pod())

    @classmethod
    def from_particle(cls, particle, medium, spectrum, oversample=1.0, rng=None):
        """src/lightyield.jl:642-653"""
        lp = LongitudinalParameterisation.from_energy(particle.energy, medium, particle.type)
        photons = pois_rand(total_lightyield(particle, medium, spectrum) * oversample, rng)
        return cls(_vec(particle.position), _vec(particle.direction), float(particle.time),
                   photons, lp)


@dataclass
class ExtendedCherenkovEmitterNoSmear(CherenkovEmitter):
    position: np.ndarray
    direction: np.ndarray
    time: float
    photons: int
    long_param: LongitudinalParameterisation
    kind = _lib.SRC_EXTENDED_CHERENKOV_NO_SMEAR

    def to_pod(self):
        return _cascade_pod(self, super().to_pod())

    @classmethod
    def from_particle(cls, particle, medium, spectrum, oversample=1.0, rng=None):
        lp = LongitudinalParameterisation.from_energy(particle.energy, medium, particle.type)
        photons = pois_rand(total_lightyield(particle, medium, spectrum) * oversample, rng)
        return cls(_vec(particle.position), _vec(particle.direction), float(particle.time),
                   photons, lp)


@dataclass
class ExtendedCherenkovEmitterCustomSmear(CherenkovEmitter):
    position: np.ndarray
    direction: np.ndarray
    time: float
    photons: int
    cherenkov_beta_a: float
    cherenkov_beta_b: float
    long_param: LongitudinalParameterisation
    kind = _lib.SRC_EXTENDED_CHERENKOV_CUSTOM_SMEAR

    def to_pod(self):
        s = _cascade_pod(self, super().to_pod())
        s.beta_a, s.beta_b = self.cherenkov_beta_a, self.cherenkov_beta_b
        return s

    @classmethod
    def from_particle(cls, particle, medium, spectrum, cherenkov_beta_a=1.776,
                      cherenkov_beta_b=0.164, oversample=1.0, rng=None):
        """src/lightyield.jl:755-769"""
        lp = LongitudinalParameterisation.from_energy(particle.energy, medium, particle.type)
        photons = pois_rand(total_lightyield(particle, medium, spectrum) * oversample, rng)
        return cls(_vec(particle.position), _vec(particle.direction), float(particle.time),
                   photons, cherenkov_beta_a, cherenkov_beta_b, lp)


@dataclass
class PointlikeCherenkovEmitter(CherenkovEmitter):
    position: np.ndarray
    direction: np.ndarray
    time: float
    photons: int
    kind = _lib.SRC_POINTLIKE_CHERENKOV

    def to_pod(self):
        s = super().to_pod()
        s.time = self.time
        return s


@dataclass
class CherenkovTrackEmitter(CherenkovEmitter):
    position: np.ndarray
    direction: np.ndarray
    time: float
    length: float
    photons: int
    kind = _lib.SRC_CHERENKOV_TRACK

    def to_pod(self):
        s = super().to_pod()
        s.time, s.length = self.time, self.length
        return s

    @classmethod
    def from_particle(cls, particle, medium, spectrum, rng=None):
        n = pois_rand(total_lightyield(particle, medium, spectrum), rng)
        return cls(_vec(particle.position), _vec(particle.direction), float(particle.time),
                   float(particle.length), n)


_LIGHTSABRE_COEFFS = (4.9489616410707695, -2.4858046180252362, 1.1885034976827853,
                      -0.2015848374875856, 0.01626011917463439, -0.0004947317263294414)


def FastLightsabreMuonEmitter(particle, medium, spectrum, rng=None):
    """src/lightyield.jl:849-873: bare track yield plus a polynomial
    stochastic-loss yield in log10(E), as one CherenkovTrackEmitter."""
    if particle.energy > 1e8 or particle.energy < 1e1:
        raise ValueError("particle energy outside parametrization range")
    le = math.log10(particle.energy)
    ly_stoch = 10 ** sum(c * le**i for i, c in enumerate(_LIGHTSABRE_COEFFS)) * particle.length
    ly_bare = total_lightyield(particle, medium, spectrum)
    n = pois_rand(ly_bare + ly_stoch, rng)
    return CherenkovTrackEmitter(_vec(particle.position), _vec(particle.direction),
                                 float(particle.time), float(particle.length), n)


def split_source(source, parts):
    """src/lightyield.jl:440-470"""
    if source.photons < parts:
        raise ValueError("Cannot split source. Fewer photons than parts")
    q, r = divmod(int(source.photons), parts)
    return [replace(source, photons=q + (r if i == parts - 1 else 0)) for i in range(parts)]


def rescale_source(source, factor):
    """src/lightyield.jl:479-488"""
    return replace(source, photons=int(math.ceil(source.photons * factor)))


