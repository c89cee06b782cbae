"""photonprop - MI355X (gfx950) photon propagation behind PhotonPropagation.jl's API.

The hot path (per-photon Monte Carlo transport) runs in libphotonprop_hip.so,
hand-written HIP for gfx950, reached through its C ABI
(include/photonprop_hip.h). This package is the host side: the reference's
setup, emitter, target, medium and spectrum types and ``propagate_photons``
with the reference's hit-table layout.
"""
from . import _lib, autotune, workloads
from .detection import (Circular, DetectorLine, DetectorLines, HomogeneousDetector, PhotonTarget,
                        Rectangular, Spherical, circular_target, flatten_targets,
                        make_detector_grid, rectangular_target, spherical_target,
                        targets_to_pod)
from .lightyield import (AxiconeEmitter, CherenkovEmitter, CherenkovTrackEmitter,
                         CollimatedIsotropicEmitter, ExtendedCherenkovEmitter,
                         ExtendedCherenkovEmitterCustomSmear, ExtendedCherenkovEmitterNoSmear,
                         FastLightsabreMuonEmitter, LongitudinalParameterisation, Particle,
                         PEMinus, PEPlus, PGamma, PHadronShower, PhotonSource, PMuMinus, PMuPlus,
                         PencilEmitter, PointlikeCherenkovEmitter, PointlikeIsotropicEmitter,
                         PointlikeTimeRangeEmitter, cascade_cherenkov_track_length,
                         rescale_source, split_source, total_lightyield)
from .medium import (CascadiaMediumProperties, HomogenousIceProperties, MediumProperties,
                     SimpleMediumProperties, make_homogenous_clearice_properties,
                     make_simple_medium)
from .processing import (DeviceHitBuffer, HostHitBuffer, PhotonPropOOMException,
                         assemble_hit_table, calc_total_weight, hits_to_dataframe,
                         make_hit_buffers, propagate_photons, run_photon_prop_no_local_cache)
from .propsetup import PhotonPropSetup
from .spectrum import (DeviceSpectralDist, InterpolatedSpectralDist, Monochromatic, Spectrum,
                       frank_tamm, integrate_gauss_quad, make_biased_cherenkov_spectrum,
                       make_cherenkov_spectrum, make_cuda_spectral_dist,
                       make_monochromatic_spectrum, make_spectral_dist, pmt_acceptance_curve)

__all__ = [n for n in dir() if not n.startswith("_")]
