"""Learned surrogate of the Monte Carlo: per (cascade, module) pair, the expected
number of photon hits and a 3-component gamma mixture of their arrival-time
residuals (include/photonprop_surrogate.h).

It has no counterpart in the reference. It is trained only on this Monte
Carlo's output (generate_training_data), runs on the gfx950 matrix cores
(bf16 MFMA, fp32 accumulation, fp32 master weights, Adam), and the Monte Carlo
stays the fallback for pairs outside its domain (SurrogateModel.expected_hits)
and the judge of its error (judge_events).

Label of a sample (one cascade, one module): the cascade is placed at distance
r from a single module and propagated with n_mc photons; w = sum of the hits'
absorption weights exp(-d / abs_len(wl)) (src/processing.jl:32-44), and the
weighted histogram of x = t - t_geo + 20 ns, t_geo = (r - R) / c_ref, over
64 bins of [0, 2000) ns. Expected hits of a cascade of energy E at that
geometry = yield(E) * w / n_mc, yield(E) = the photon count of an
ExtendedCherenkovEmitter of energy E (src/lightyield.jl:642-653) without
Poisson fluctuation.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .detection import HomogeneousDetector, Spherical
from .lightyield import (ExtendedCherenkovEmitter, LongitudinalParameterisation, Particle,
                         PEMinus, total_lightyield)
from .medium import CascadiaMediumProperties
from .processing import (HostHitBuffer, make_hit_buffers, run_photon_prop_no_local_cache)

N_IN = _lib.SUR_N_FEATURES
N_OUT = _lib.SUR_N_OUT
BINS = _lib.SUR_BINS
TMAX = _lib.SUR_TMAX
SHIFT = _lib.SUR_TIME_SHIFT
LABEL = _lib.SUR_LABEL_FLOATS


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class MLP:
    """A network on the device (pp_mlp_*)."""

    def __init__(self, n_in=N_IN, n_hidden=256, n_layers=4, n_out=N_OUT, max_batch=1 << 16,
                 seed=1, device=0):
        self.lib = _lib.load_library()
        h = ctypes.c_void_p()
        _lib.check(self.lib.pp_mlp_create(ctypes.byref(h), device, n_in, n_hidden, n_layers,
                                          n_out, max_batch, ctypes.c_uint64(seed)),
                   "pp_mlp_create")
        self.handle = h
        self.dims = [n_in] + [n_hidden] * n_layers + [n_out]
        self.max_batch = max_batch

    def close(self):
        if self.handle:
            self.lib.pp_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def n_params(self):
        return int(self.lib.pp_mlp_param_count(self.handle))

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        _lib.check(self.lib.pp_mlp_get_params(self.handle, _p(p)), "pp_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _lib.check(self.lib.pp_mlp_set_params(self.handle, _p(p)), "pp_mlp_set_params")

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.dims[-1]), np.float32)
        _lib.check(self.lib.pp_mlp_forward(self.handle, _p(x), x.shape[0], _p(y)),
                   "pp_mlp_forward")
        return y

    def loss_grad(self, x, labels):
        x = np.ascontiguousarray(x, np.float32)
        labels = np.ascontiguousarray(labels, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_float()
        _lib.check(self.lib.pp_mlp_loss_grad(self.handle, _p(x), _p(labels), x.shape[0],
                                             ctypes.byref(loss), _p(g)), "pp_mlp_loss_grad")
        return float(loss.value), g

    @staticmethod
    def adam_config(lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        o = _lib.pp_adam()
        o.lr, o.beta1, o.beta2, o.eps, o.weight_decay = lr, beta1, beta2, eps, weight_decay
        return o

    def adam_step(self, **kw):
        o = self.adam_config(**kw)
        _lib.check(self.lib.pp_mlp_adam_step(self.handle, ctypes.byref(o)), "pp_mlp_adam_step")

    def train(self, x, labels, n_steps, batch=4096, seed=7, **adam):
        x = np.ascontiguousarray(x, np.float32)
        labels = np.ascontiguousarray(labels, np.float32)
        o = self.adam_config(**adam)
        curve = np.zeros(max(n_steps, 1), np.float32)
        _lib.check(self.lib.pp_mlp_train(self.handle, _p(x), _p(labels), x.shape[0], batch,
                                         n_steps, ctypes.byref(o), ctypes.c_uint64(seed),
                                         _p(curve)), "pp_mlp_train")
        return curve[:n_steps]

    def eval_loss(self, x, labels):
        x = np.ascontiguousarray(x, np.float32)
        labels = np.ascontiguousarray(labels, np.float32)
        loss = ctypes.c_float()
        _lib.check(self.lib.pp_mlp_eval_loss(self.handle, _p(x), _p(labels), x.shape[0],
                                             ctypes.byref(loss)), "pp_mlp_eval_loss")
        return float(loss.value)

    def features(self, cascades, modules):
        """cascades: (n, 8) rows x, y, z, dx, dy, dz, energy, log_yield; modules (n, 3)."""
        c = np.ascontiguousarray(cascades, np.float32)
        m = np.ascontiguousarray(modules, np.float32)
        f = np.zeros((c.shape[0], self.dims[0]), np.float32)
        _lib.check(self.lib.pp_sur_features(self.handle, _p(c), _p(m), c.shape[0], _p(f)),
                   "pp_sur_features")
        return f

    def eval_pairs(self, events, modules, domain, mixture=False):
        """events (n_e, 8) cascade rows, modules (n_m, 3): expected hits (n_e, n_m),
        mixture (n_e, n_m, 9) or None, fallback mask (n_e, n_m) and device ms."""
        ev = np.ascontiguousarray(events, np.float32)
        mo = np.ascontiguousarray(modules, np.float32)
        ne, nm = ev.shape[0], mo.shape[0]
        hits = np.zeros((ne, nm), np.float32)
        mix = np.zeros((ne, nm, 9), np.float32) if mixture else None
        fb = np.zeros((ne, nm), np.uint8)
        ms = ctypes.c_float()
        _lib.check(self.lib.pp_sur_eval_pairs(self.handle, _p(ev), ne, _p(mo), nm,
                                              ctypes.byref(domain), _p(hits),
                                              _p(mix) if mixture else None, _p(fb),
                                              ctypes.byref(ms)), "pp_sur_eval_pairs")
        return hits, mix, fb.astype(bool), float(ms.value)


# ---------------------------------------------------------------------------
# physics helpers
# ---------------------------------------------------------------------------
@dataclass
class SurrogateSetup:
    """Medium, spectrum and module of the surrogate's world."""
    medium: object
    spectrum: object        # a Spectrum (for the yield)
    device_spectrum: object  # its device table (for the Monte Carlo)
    module_radius: float = 0.3
    n_steps: int = 15

    @property
    def c_ref(self):
        """Fastest group velocity of the spectrum's range (t_geo reference)."""
        lo, hi = self.spectrum.wl_range
        wl = np.linspace(lo, hi, 201)
        return float(np.max(self.medium.group_velocity(wl)))

    def log_yield(self, energy):
        e = np.atleast_1d(np.asarray(energy, np.float64))
        ys = [total_lightyield(Particle(np.zeros(3), np.array([0, 0, 1.0]), 0.0, float(x), 0.0,
                                        PEMinus), self.medium, self.spectrum) for x in e]
        return np.log(np.asarray(ys, np.float64))

    def domain(self, min_dist=5.0, max_dist=300.0, min_energy=100.0, max_energy=1e6):
        d = _lib.pp_sur_domain()
        d.min_dist, d.max_dist = min_dist, max_dist
        d.min_energy, d.max_energy = min_energy, max_energy
        d.module_radius = self.module_radius
        d.c_ref = self.c_ref
        return d


def default_setup():
    from .spectrum import make_biased_cherenkov_spectrum, make_cuda_spectral_dist, \
        pmt_acceptance_curve
    medium = CascadiaMediumProperties()
    spec = make_biased_cherenkov_spectrum(pmt_acceptance_curve(), (300.0, 700.0), medium)
    dev = make_cuda_spectral_dist(spec.spectral_dist, (300.0, 700.0), 600)
    return SurrogateSetup(medium, spec, dev)


def sample_geometry(rng, n, r_range=(1.0, 300.0), e_range=(100.0, 1e6)):
    """Cascade vertices around a module at the origin: distance log-uniform,
    direction to the module uniform on the sphere, cascade axis isotropic,
    energy log-uniform."""
    r = np.exp(rng.uniform(np.log(r_range[0]), np.log(r_range[1]), n))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    e = np.exp(rng.uniform(np.log(e_range[0]), np.log(e_range[1]), n))
    pos = -r[:, None] * u  # module at the origin, vertex at -r u
    return pos.astype(np.float32), d.astype(np.float32), e


def photons_for_distance(r, lo=2e4, hi=1e7, scale=2e4):
    """Photons simulated per sample: about constant expected hits up to r ~ 22 m."""
    return np.clip(scale * np.asarray(r) ** 2, lo, hi).astype(np.int64)


def generate_training_data(setup: SurrogateSetup, n_samples, seed=1, batch_sources=2048,
                           photons_fn=photons_for_distance, dev=None, log=print):
    """Labels from the Monte Carlo: every sample is a cascade (vertex, axis,
    energy) around one module at the origin; batches of samples run as the
    sources of one persistent propagation launch. Returns (cascades (n, 8),
    labels (n, 66), stats)."""
    rng = np.random.default_rng(seed)
    pos, d, e = sample_geometry(rng, n_samples)
    r = np.linalg.norm(pos, axis=1)
    nmc = photons_fn(r)
    logy = setup.log_yield(e)
    target = HomogeneousDetector(Spherical(np.float32([0, 0, 0]), setup.module_radius),
                                 np.pi * setup.module_radius ** 2, 1)
    labels = np.zeros((n_samples, LABEL), np.float32)
    labels[:, 0] = nmc
    own = dev is None
    if own:
        _, dev = make_hit_buffers(capacity=1 << 25)
    c_ref = setup.c_ref
    t0 = time.perf_counter()
    photons = 0
    kernel_ms = 0.0
    try:
        for b0 in range(0, n_samples, batch_sources):
            b1 = min(n_samples, b0 + batch_sources)
            srcs = [ExtendedCherenkovEmitter(pos[i], d[i], 0.0, int(nmc[i]),
                                             LongitudinalParameterisation.from_energy(
                                                 float(e[i]), setup.medium, PEMinus))
                    for i in range(b0, b1)]
            hits, nsim = run_photon_prop_no_local_cache(
                srcs, [target], setup.medium, setup.device_spectrum, seed * 1000003 + b0,
                HostHitBuffer(np.float32), dev, n_steps=setup.n_steps, with_ids=True)
            kernel_ms += dev.last_timing()[0]
            photons += nsim
            s = hits["source_ix"].astype(np.int64) + b0
            w = np.exp(-hits["dist_travelled"].astype(np.float64) /
                       setup.medium.absorption_length(hits["wavelength"].astype(np.float64)))
            labels[:, 1] += np.bincount(s, weights=w, minlength=n_samples).astype(np.float32)
            tgeo = (r[s] - setup.module_radius) / c_ref
            x = hits["time"].astype(np.float64) - tgeo + SHIFT
            binx = np.floor(x / (TMAX / BINS)).astype(np.int64)
            ok = (binx >= 0) & (binx < BINS)
            np.add.at(labels, (s[ok], 2 + binx[ok]), w[ok].astype(np.float32))
            if log and (b0 // batch_sources) % 10 == 0:
                log(f"  labels {b1}/{n_samples}: {photons:.3g} photons, "
                    f"{time.perf_counter() - t0:.1f} s")
    finally:
        if own:
            dev.close()
    cascades = np.concatenate([pos, d, e[:, None].astype(np.float32),
                               logy[:, None].astype(np.float32)], axis=1)
    stats = dict(photons=photons, seconds=time.perf_counter() - t0, kernel_ms=kernel_ms,
                 samples=n_samples)
    return cascades.astype(np.float32), labels, stats


def module_at_origin(n):
    return np.zeros((n, 3), np.float32)
