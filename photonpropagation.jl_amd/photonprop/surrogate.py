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
from .medium import STRAW_ABS_LEN, STRAW_WL, CascadiaMediumProperties
from .processing import (HostHitBuffer, make_hit_buffers, run_photon_prop_no_local_cache,
                         sources_to_pod, stage)

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

    def set_fused(self, fused=True):
        """Inference through the fused all-layer kernel (default) or one GEMM
        kernel per layer; the outputs are bitwise identical."""
        _lib.check(self.lib.pp_mlp_set_fused(self.handle, int(bool(fused))), "pp_mlp_set_fused")

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

    def domain(self, min_dist=5.0, max_dist=800.0, min_energy=100.0, max_energy=1e6):
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


def exposure(n_photons, r, radius):
    """Photons x (R / 2r)^2: the label normaliser of the expected hits."""
    r = np.maximum(np.asarray(r, np.float64), radius)
    return np.asarray(n_photons, np.float64) * (radius / (2.0 * r)) ** 2


def sample_geometry(rng, n, r_range=(1.0, 800.0), e_range=(100.0, 1e6)):
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


def photons_for_distance(r, lo=2e4, hi=4e6, scale=2e4):
    """Photons simulated per sample: about constant expected hits up to r ~ 22 m."""
    return np.clip(scale * np.asarray(r) ** 2, lo, hi).astype(np.int64)


# The surrogate's training configuration at its stated size: 2e6 cascades
# around one module, distance log-uniform 1-300 m, energy log-uniform
# 100 GeV - 1 PeV, labels from 1e7-photon runs, Adam at batch 4096.
CONFIG_SAMPLES = 2_000_000
CONFIG_PHOTONS = 10_000_000
CONFIG_R_RANGE = (1.0, 300.0)
CONFIG_BATCH = 4096
CONFIG_SEED = 2026


def save_label_slice(path, first, labels, stats):
    """Labels of samples [first, first + len(labels)) (rows as returned by
    generate_training_data) to a compressed .npz: the weighted hit total in
    float32 and each time bin as a float16 fraction of it (relative error at
    most 2^-11 of the bin, far below the Monte Carlo's own), with the run's
    photon count and times."""
    tot = labels[:, 1].astype(np.float32)
    frac = np.where(tot[:, None] > 0, labels[:, 2:] / np.maximum(tot, 1e-30)[:, None], 0.0)
    np.savez_compressed(path, first=first, count=len(labels), total=tot,
                        frac=frac.astype(np.float16), photons=stats["photons"],
                        kernel_ms=stats["kernel_ms"], seconds=stats["seconds"])


def load_label_slices(paths, labels):
    """Fill labels[:, 1:] from save_label_slice files; returns (rows covered
    (bool mask), summed photons, kernel ms and seconds, [first, end) list)."""
    covered = np.zeros(len(labels), bool)
    out = dict(photons=0, kernel_ms=0.0, seconds=0.0, slices=[])
    for f in paths:
        z = np.load(f)
        a, k = int(z["first"]), int(z["count"])
        labels[a:a + k, 1] = z["total"]
        labels[a:a + k, 2:] = z["frac"].astype(np.float32) * z["total"][:, None]
        covered[a:a + k] = True
        out["photons"] += int(z["photons"])
        out["kernel_ms"] += float(z["kernel_ms"])
        out["seconds"] += float(z["seconds"])
        out["slices"].append([a, a + k])
    return covered, out


def constant_photons(n):
    """photons_fn of generate_training_data: n photons for every sample."""
    return lambda r: np.full(np.shape(r), n, np.int64)


def generate_training_data(setup: SurrogateSetup, n_samples, seed=1, batch_sources=2048,
                           photons_fn=photons_for_distance, dev=None, log=print,
                           r_range=(1.0, 800.0), device_labels=True, first=0, count=None):
    """Labels from the Monte Carlo: every sample is a cascade (vertex, axis,
    energy) around one module at the origin; batches of samples run as the
    sources of one persistent propagation launch. With device_labels the
    hits are reduced to labels on the device (pp_sur_hit_labels: only the
    labels come back); otherwise the hits are copied and reduced here.
    first/count select a slice of the n_samples (first a multiple of
    batch_sources; count too unless the slice ends at n_samples): the slice's
    labels are those of the whole run (same geometry, same batches, same
    random streams), the other rows stay zero apart from the exposure.
    Returns (cascades (n, 8), labels (n, 66), stats)."""
    if count is None:
        count = n_samples - first
    end = first + count
    if not (0 <= first <= end <= n_samples) or count and (
            first % batch_sources or (count % batch_sources and end != n_samples)):
        raise ValueError("generate_training_data: the slice must follow the batch boundaries")
    lib = _lib.load_library()
    rng = np.random.default_rng(seed)
    pos, d, e = sample_geometry(rng, n_samples, r_range=r_range)
    r = np.linalg.norm(pos, axis=1)
    nmc = photons_fn(r)
    logy = setup.log_yield(e)
    target = HomogeneousDetector(Spherical(np.float32([0, 0, 0]), setup.module_radius),
                                 np.pi * setup.module_radius ** 2, 1)
    labels = np.zeros((n_samples, LABEL), np.float32)
    labels[:, 0] = exposure(nmc, r, setup.module_radius)
    own = dev is None
    if own:
        _, dev = make_hit_buffers(capacity=1 << 26)
    c_ref = setup.c_ref
    medium = setup.medium
    abs_wl = np.ascontiguousarray(STRAW_WL, np.float64)
    abs_len = np.ascontiguousarray(STRAW_ABS_LEN, np.float64)
    t0 = time.perf_counter()
    photons = 0
    kernel_ms = 0.0

    def sources(b0, b1):
        return [ExtendedCherenkovEmitter(pos[i], d[i], 0.0, int(nmc[i]),
                                         LongitudinalParameterisation.from_energy(
                                             float(e[i]), medium, PEMinus))
                for i in range(b0, b1)]

    def device_batch(b0, b1, key):
        """Propagate samples [b0, b1) and reduce their hits on the device; a
        batch whose hits overflow the buffer is run as two halves."""
        nonlocal photons, kernel_ms
        pods = sources_to_pod(sources(b0, b1))
        nh, ns = ctypes.c_int64(), ctypes.c_int64()
        rc = lib.pp_propagate(dev.handle, pods, b1 - b0, ctypes.c_uint64(key), setup.n_steps, 1.0,
                              ctypes.byref(nh), ctypes.byref(ns))
        if rc == _lib.PP_ERR_HIT_OVERFLOW and b1 - b0 > 1:
            mid = (b0 + b1) // 2
            device_batch(b0, mid, key)
            device_batch(mid, b1, key + 1)
            return
        _lib.check(rc, "pp_propagate")
        kernel_ms += dev.last_timing()[0]
        photons += ns.value
        tg = np.ascontiguousarray((r[b0:b1] - setup.module_radius) / c_ref, np.float64)
        out = np.zeros((b1 - b0, 1 + BINS), np.float32)
        _lib.check(lib.pp_sur_hit_labels(dev.handle, _p(abs_wl), _p(abs_len), len(abs_wl),
                                         float(medium.abs_scale), _p(tg), b1 - b0, _p(out)),
                   "pp_sur_hit_labels")
        labels[b0:b1, 1:] = out

    try:
        if device_labels:
            stage(dev, [target], medium, setup.device_spectrum)
        for b0 in range(first, end, batch_sources):
            b1 = min(end, b0 + batch_sources)
            if device_labels:
                device_batch(b0, b1, seed * 1000003 + b0)
            else:
                hits, nsim = run_photon_prop_no_local_cache(
                    sources(b0, b1), [target], medium, setup.device_spectrum,
                    seed * 1000003 + b0, HostHitBuffer(np.float32), dev, n_steps=setup.n_steps,
                    with_ids=True)
                kernel_ms += dev.last_timing()[0]
                photons += nsim
                s = hits["source_ix"].astype(np.int64) + b0
                w = np.exp(-hits["dist_travelled"].astype(np.float64) /
                           medium.absorption_length(hits["wavelength"].astype(np.float64)))
                labels[:, 1] += np.bincount(s, weights=w, minlength=n_samples).astype(np.float32)
                tgeo = (r[s] - setup.module_radius) / c_ref
                x = hits["time"].astype(np.float64) - tgeo + SHIFT
                binx = np.floor(x / (TMAX / BINS)).astype(np.int64)
                ok = (binx >= 0) & (binx < BINS)
                flat = np.bincount((s[ok] - b0) * BINS + binx[ok], weights=w[ok],
                                   minlength=(b1 - b0) * BINS)
                labels[b0:b1, 2:] += flat.reshape(b1 - b0, BINS).astype(np.float32)
            if log and (b0 // batch_sources) % 10 == 0:
                log(f"  labels {b1 - first}/{count}: {photons:.3g} photons, "
                    f"{time.perf_counter() - t0:.1f} s")
    finally:
        if own:
            dev.close()
    cascades = np.concatenate([pos, d, e[:, None].astype(np.float32),
                               logy[:, None].astype(np.float32)], axis=1)
    stats = dict(photons=photons, seconds=time.perf_counter() - t0, kernel_ms=kernel_ms,
                 samples=n_samples, first=first, count=count)
    return cascades.astype(np.float32), labels, stats


def module_at_origin(n):
    return np.zeros((n, 3), np.float32)


# ---------------------------------------------------------------------------
# training, Monte Carlo judge, fallback
# ---------------------------------------------------------------------------
def train_surrogate(setup: SurrogateSetup, cascades, labels, *, steps=8000, batch=4096,
                    lr=2e-3, lr_stages=4, seed=3, n_hidden=256, n_layers=4, val_frac=0.05,
                    max_batch=1 << 16, log=print):
    """Train a fresh network on (cascades, labels) from generate_training_data.
    The learning rate halves over lr_stages equal stages. Returns (mlp, info)."""
    n = cascades.shape[0]
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    n_val = max(1, int(val_frac * n))
    val, tr = perm[:n_val], perm[n_val:]
    mlp = MLP(n_hidden=n_hidden, n_layers=n_layers, max_batch=max_batch, seed=seed)
    feats = mlp.features(cascades, module_at_origin(n))
    t0 = time.perf_counter()
    curves = []
    per = max(1, steps // lr_stages)
    for s in range(lr_stages):
        curves.append(mlp.train(feats[tr], labels[tr], per, batch=batch, seed=seed * 7919 + s,
                                lr=lr * 0.5 ** s))
        if log:
            log(f"  stage {s}: lr {lr * 0.5 ** s:.2g}, batch loss {curves[-1][-1]:.5g}, "
                f"val loss {mlp.eval_loss(feats[val], labels[val]):.5g}")
    secs = time.perf_counter() - t0
    info = dict(train_seconds=secs, steps=per * lr_stages, batch=batch,
                steps_per_s=per * lr_stages / secs, samples=int(len(tr)),
                val_loss=mlp.eval_loss(feats[val], labels[val]),
                loss_curve=np.concatenate(curves))
    return mlp, info


def array_events(rng, n, module_xyz, e_range=(1e3, 1e6), setup=None):
    """Cascades at uniform positions inside the modules' bounding box, isotropic
    axes, log-uniform energies: rows x, y, z, dx, dy, dz, energy, log_yield."""
    lo, hi = module_xyz.min(axis=0), module_xyz.max(axis=0)
    pos = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    e = np.exp(rng.uniform(np.log(e_range[0]), np.log(e_range[1]), n))
    ly = (setup or default_setup()).log_yield(e)
    return np.concatenate([pos, d, e[:, None], ly[:, None]], axis=1).astype(np.float32)


def mc_expected_hits(setup: SurrogateSetup, events, targets, *, max_photons=1e9, seed=11,
                     dev=None, events_per_call=16, with_times=False):
    """The Monte Carlo's expected hits per (event, module): every event is an
    ExtendedCherenkovEmitter with min(yield, max_photons) photons, each hit
    weighted by its absorption weight and yield / photons. Returns dict with
    'hits' (n_e, n_m), 'raw' hit counts, 'seconds', 'kernel_ms', 'photons' and,
    with_times, per-hit (event, module, x = t - t_geo + 20 ns, weight)."""
    ne, nm = events.shape[0], len(targets)
    mxyz = np.array([t.shape.position for t in targets], np.float64)
    own = dev is None
    if own:
        _, dev = make_hit_buffers(capacity=1 << 26)
    exp_hits = np.zeros((ne, nm), np.float64)
    w2 = np.zeros((ne, nm), np.float64)
    raw = np.zeros((ne, nm), np.int64)
    times = []
    photons = 0
    kernel_ms = 0.0
    c_ref = setup.c_ref
    t0 = time.perf_counter()
    try:
        for e0 in range(0, ne, events_per_call):
            e1 = min(ne, e0 + events_per_call)
            srcs, factor = [], []
            for i in range(e0, e1):
                y = float(np.exp(events[i, 7]))
                n_ph = int(min(y, max_photons))
                factor.append(y / n_ph)
                srcs.append(ExtendedCherenkovEmitter(
                    events[i, :3].copy(), events[i, 3:6].copy(), 0.0, n_ph,
                    LongitudinalParameterisation.from_energy(float(events[i, 6]), setup.medium,
                                                             PEMinus)))
            hits, nsim = run_photon_prop_no_local_cache(
                srcs, targets, setup.medium, setup.device_spectrum, seed + e0,
                HostHitBuffer(np.float32), dev, n_steps=setup.n_steps, with_ids=True)
            kernel_ms += dev.last_timing()[0]
            photons += nsim
            ev = hits["source_ix"].astype(np.int64) + e0
            mod = hits["module_id"].astype(np.int64) - 1
            w = np.exp(-hits["dist_travelled"].astype(np.float64) /
                       setup.medium.absorption_length(hits["wavelength"].astype(np.float64)))
            w *= np.asarray(factor)[ev - e0]
            flat = ev * nm + mod
            exp_hits += np.bincount(flat, weights=w, minlength=ne * nm).reshape(ne, nm)
            w2 += np.bincount(flat, weights=w * w, minlength=ne * nm).reshape(ne, nm)
            raw += np.bincount(flat, minlength=ne * nm).reshape(ne, nm)
            if with_times:
                r = np.linalg.norm(mxyz[mod] - events[ev, :3], axis=1)
                x = hits["time"].astype(np.float64) - (r - setup.module_radius) / c_ref + SHIFT
                times.append((ev, mod, x, w))
    finally:
        if own:
            dev.close()
    n_eff = np.where(w2 > 0, exp_hits ** 2 / np.maximum(w2, 1e-300), 0.0)
    out = dict(hits=exp_hits, raw=raw, n_eff=n_eff, w2=w2, seconds=time.perf_counter() - t0,
               kernel_ms=kernel_ms, photons=photons)
    if with_times:
        out["times"] = tuple(np.concatenate(c) for c in zip(*times)) if times else None
    return out


def gamma_mixture_logpdf(x, mix):
    """log density of x (n,) under mixtures mix (n, 9): weights, shapes, scales."""
    from scipy.special import gammaln
    x = np.maximum(np.asarray(x, np.float64), 1e-6)[:, None]
    w, k, th = mix[:, :3].astype(np.float64), mix[:, 3:6].astype(np.float64), \
        mix[:, 6:9].astype(np.float64)
    lp = np.log(np.maximum(w, 1e-300)) + (k - 1) * np.log(x) - x / th - k * np.log(th) - gammaln(k)
    m = lp.max(axis=1, keepdims=True)
    return (m + np.log(np.exp(lp - m).sum(axis=1, keepdims=True)))[:, 0]


class SurrogateModel:
    """Trained network + its domain; pairs outside the domain fall back to the
    Monte Carlo."""

    def __init__(self, mlp: MLP, setup: SurrogateSetup, domain=None):
        self.mlp = mlp
        self.setup = setup
        self.domain = domain or setup.domain()

    def expected_hits(self, events, targets, *, mc_fallback=True, max_photons=1e9, seed=5,
                      dev=None, mixture=False):
        """Expected hits for every (event, module) pair. Returns dict with
        'hits', 'mixture', 'fallback' (mask), 'surrogate_ms' (device time of
        the batched network evaluation) and 'mc_seconds' (fallback)."""
        mxyz = np.array([t.shape.position for t in targets], np.float32)
        hits, mix, fb, ms = self.mlp.eval_pairs(events, mxyz, self.domain, mixture=mixture)
        hits = hits.astype(np.float64)
        out = dict(hits=hits, mixture=mix, fallback=fb, surrogate_ms=ms, mc_seconds=0.0,
                   mc_events=0)
        if mc_fallback and fb.any():
            rows = np.nonzero(fb.any(axis=1))[0]
            mc = mc_expected_hits(self.setup, events[rows], targets, max_photons=max_photons,
                                  seed=seed, dev=dev)
            for k, e in enumerate(rows):
                hits[e, fb[e]] = mc["hits"][k, fb[e]]
            out["mc_seconds"] = mc["seconds"]
            out["mc_events"] = int(len(rows))
        return out


DIST_EDGES = (5.0, 20.0, 50.0, 100.0, 200.0, 400.0, 800.0)
HIT_EDGES = (0.0, 1.0, 10.0, 100.0, 1e3, 1e4, np.inf)


def judge_strata(sur_hits, mc, fallback, dist, min_eff=100.0, min_raw=10):
    """Every in-domain (event, module) pair, stratified by distance and by the
    Monte Carlo's expected hits. Per stratum: pairs, pairs with MC hits, the
    ratio of summed surrogate to summed MC expected hits (a bias), the pulls
    (sur - mc) / sigma_mc over pairs with >= min_raw MC hits (sigma_mc =
    sqrt(sum w^2), the MC's own statistical error: a pull RMS near 1 means the
    surrogate is as close as the MC's statistics can tell), the mean relative
    error over pairs whose MC estimate has >= min_eff effective hits, and the
    surrogate's summed expectation on pairs the MC never hit."""
    ind = ~fallback
    mh, raw, w2, n_eff = mc["hits"], mc["raw"], mc["w2"], mc["n_eff"]
    sigma = np.sqrt(w2)

    def stats(sel):
        d = dict(pairs=int(sel.sum()), pairs_with_hits=int((sel & (raw > 0)).sum()))
        if not sel.any():
            return d
        tot_mc = float(mh[sel].sum())
        d["sum_ratio"] = float(sur_hits[sel].sum() / tot_mc) if tot_mc > 0 else None
        pull_sel = sel & (raw >= min_raw)
        if pull_sel.any():
            pull = (sur_hits[pull_sel] - mh[pull_sel]) / sigma[pull_sel]
            d.update(pull_pairs=int(pull_sel.sum()), pull_mean=float(pull.mean()),
                     pull_rms=float(np.sqrt(np.mean(pull ** 2))))
        rel_sel = sel & (n_eff >= min_eff)
        if rel_sel.any():
            d.update(rel_pairs=int(rel_sel.sum()),
                     mean_rel_err=float(np.mean(np.abs(sur_hits[rel_sel] / mh[rel_sel] - 1))))
        zero = sel & (raw == 0)
        if zero.any():
            d["zero_hit_pairs"] = int(zero.sum())
            d["sur_hits_on_zero_pairs"] = float(sur_hits[zero].sum())
        return d

    out = {"by_distance": [], "by_mc_hits": []}
    for lo, hi in zip(DIST_EDGES[:-1], DIST_EDGES[1:]):
        out["by_distance"].append(dict(range_m=[lo, hi], **stats(ind & (dist >= lo) & (dist < hi))))
    for lo, hi in zip(HIT_EDGES[:-1], HIT_EDGES[1:]):
        sel = ind & (mh >= lo) & (mh < hi) & (raw > 0)
        out["by_mc_hits"].append(dict(range=[lo, None if np.isinf(hi) else hi], **stats(sel)))
    out["all"] = stats(ind)
    return out


def judge_events(model: SurrogateModel, events, targets, *, max_photons=1e9, min_eff_hits=2500,
                 seed=21, dev=None):
    """The full Monte Carlo judges the surrogate on held-out events: mean
    relative error of the expected hits over the in-domain pairs whose MC
    estimate has at least min_eff_hits effective hits (relative statistical
    error <= 1/sqrt(min_eff_hits), 2% by default), and the mean per-hit
    log-likelihood of the MC arrival-time residuals under the surrogate
    mixture vs under the MC's own histogram."""
    mc = mc_expected_hits(model.setup, events, targets, max_photons=max_photons, seed=seed,
                          dev=dev, with_times=True)
    sur = model.expected_hits(events, targets, mc_fallback=False, mixture=True)
    sel = (~sur["fallback"]) & (mc["n_eff"] >= min_eff_hits)
    rel = np.abs(sur["hits"][sel] / mc["hits"][sel] - 1.0)
    stat = 1.0 / np.sqrt(mc["n_eff"][sel])
    res = dict(pairs=int(sel.sum()), mean_rel_err=float(rel.mean()) if sel.any() else None,
               median_rel_err=float(np.median(rel)) if sel.any() else None,
               mean_signed_err=float(np.mean(sur["hits"][sel] / mc["hits"][sel] - 1.0))
               if sel.any() else None,
               mean_mc_stat_err=float(stat.mean()) if sel.any() else None,
               mc_seconds=mc["seconds"], mc_kernel_ms=mc["kernel_ms"], mc_photons=mc["photons"],
               surrogate_ms=sur["surrogate_ms"], n_events=int(events.shape[0]),
               n_modules=len(targets))
    mxyz = np.array([t.shape.position for t in targets], np.float64)
    dist = np.linalg.norm(mxyz[None, :, :] - events[:, None, :3].astype(np.float64), axis=2)
    res["strata"] = judge_strata(sur["hits"], mc, sur["fallback"], dist)
    if mc.get("times") is not None:
        ev, mod, x, w = mc["times"]
        keep = sel[ev, mod] & (x > 0) & (x < TMAX)
        if keep.any():
            ll_sur = gamma_mixture_logpdf(x[keep], sur["mixture"][ev[keep], mod[keep]])
            # MC's own binned density per pair (same binning as the labels)
            b = np.floor(x[keep] / (TMAX / BINS)).astype(np.int64)
            pair = ev[keep] * len(targets) + mod[keep]
            key = pair * BINS + b
            uniq, inv = np.unique(key, return_inverse=True)
            wb = np.bincount(inv, weights=w[keep])
            up, pinv = np.unique(pair, return_inverse=True)
            wp = np.bincount(pinv, weights=w[keep])
            dens = wb[inv] / wp[pinv] / (TMAX / BINS)
            ww = w[keep] / w[keep].sum()
            res["time_ll_surrogate"] = float(np.sum(ww * ll_sur))
            res["time_ll_mc_hist"] = float(np.sum(ww * np.log(dens)))
    return res, mc, sur
