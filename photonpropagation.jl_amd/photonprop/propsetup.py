"""PhotonPropSetup (src/photon_propagation_setup.jl:9-25)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Sequence

from .spectrum import (DeviceSpectralDist, InterpolatedSpectralDist, Monochromatic, Spectrum,
                       SpectralDist, make_cuda_spectral_dist)


@dataclass
class PhotonPropSetup:
    sources: Sequence[Any]
    targets: Sequence[Any]
    medium: Any
    spec_dist: SpectralDist
    seed: int
    photon_scaling: float = 1.0

    def __init__(self, sources, targets, medium, spectrum, seed, photon_scaling=1.0,
                 spectrum_interp_steps=600):
        """Interpolated spectra become a device table of spectrum_interp_steps
        knots (:18-21); other spectra are used as they are (:23-25). A bare
        SpectralDist such as Monochromatic(450) is accepted (README form)."""
        self.sources = list(sources) if isinstance(sources, (list, tuple)) else [sources]
        self.targets = list(targets) if isinstance(targets, (list, tuple)) else [targets]
        self.medium = medium
        if isinstance(spectrum, Spectrum):
            dist = spectrum.spectral_dist
            if isinstance(dist, InterpolatedSpectralDist):
                dist = make_cuda_spectral_dist(dist, spectrum.wl_range, spectrum_interp_steps)
        elif isinstance(spectrum, (Monochromatic, DeviceSpectralDist)):
            dist = spectrum
        elif isinstance(spectrum, InterpolatedSpectralDist):
            dist = make_cuda_spectral_dist(spectrum, None, spectrum_interp_steps)
        else:
            raise TypeError(f"unsupported spectrum {type(spectrum).__name__}")
        self.spec_dist = dist
        self.seed = int(seed)
        self.photon_scaling = float(photon_scaling)
