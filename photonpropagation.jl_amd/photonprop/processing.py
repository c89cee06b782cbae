"""Propagation driver and hit table.

Drop-in mirror of the reference path:
  make_hit_buffers                 src/photon_prop_cuda.jl:878-886
  run_photon_prop_no_local_cache   src/photon_prop_cuda.jl:889-959
  propagate_photons                src/processing.jl:47-75
  calc_total_weight                src/processing.jl:32-44
Propagation runs in libphotonprop_hip.so (HIP, gfx950); this module lowers
the setup to the C ABI's PODs and builds the hit table.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import pandas as pd

from . import _lib
from .detection import flatten_targets, targets_to_pod
from .spectrum import DeviceSpectralDist, Monochromatic

HIT_COLUMNS = ("position", "direction", "initial_direction", "wavelength", "time",
               "dist_travelled", "module_id", "n_steps")


class PhotonPropOOMException(Exception):
    """The hit buffer overflowed (src/photon_prop_cuda.jl:40)."""


def device_memory(device=0):
    lib = _lib.load_library()
    total = ctypes.c_int64()
    free = ctypes.c_int64()
    _lib.check(lib.pp_device_memory(device, ctypes.byref(total), ctypes.byref(free)),
               "pp_device_memory")
    return total.value, free.value


class DeviceHitBuffer:
    """Device side of make_hit_buffers: a library context owning the device
    hit buffer (capacity records), the staged tables and a HIP stream."""

    def __init__(self, capacity, device=0, config=None):
        lib = _lib.load_library()
        self.capacity = int(capacity)
        self.device = device
        cfg = config if config is not None else _lib.pp_config()
        if config is None:
            cfg.deterministic_order = 1
        h = ctypes.c_void_p()
        _lib.check(lib.pp_create(ctypes.byref(h), device, self.capacity, ctypes.byref(cfg)),
                   "pp_create")
        self.handle = h
        self.config = cfg
        self.n_hits = 0  # hits of the last propagation

    def close(self):
        if self.handle:
            _lib.load_library().pp_destroy(self.handle)
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

    def last_timing(self):
        ms = ctypes.c_float()
        n = ctypes.c_int32()
        _lib.load_library().pp_last_timing(self.handle, ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def release_scratch(self):
        """Free the buffers the context grows on demand (ordered-copy sort
        buffer, pinned copy staging, PMT detection buffers)."""
        _lib.check(_lib.load_library().pp_release_scratch(self.handle), "pp_release_scratch")


@dataclass
class HostHitBuffer:
    """Host side of make_hit_buffers: the time column type. Host columns are
    allocated per call for the used prefix only (the reference keeps a host
    copy of the full device capacity, src/photon_prop_cuda.jl:882)."""
    time_type: type = np.float32


def make_hit_buffers(time_type=np.float32, reserved_memory_fraction=0.7, *, capacity=None,
                     device=0, config=None):
    """(host, device) hit buffers. Capacity = reserved_memory_fraction of the
    device memory / bytes per record, unless `capacity` is given."""
    if capacity is None:
        total, _ = device_memory(device)
        capacity = int(reserved_memory_fraction * total) // int(
            _lib.load_library().pp_hit_record_bytes())
    return HostHitBuffer(time_type), DeviceHitBuffer(capacity, device, config)


def _spectrum_args(spec):
    if isinstance(spec, Monochromatic):
        return float(spec.wavelength), None, 0, 0
    if isinstance(spec, DeviceSpectralDist):
        tab = np.ascontiguousarray(spec.table, dtype=np.float32)
        return 0.0, tab, len(tab), int(spec.mapping)
    raise TypeError(f"unsupported spectral distribution {type(spec).__name__}")


def stage(dev: DeviceHitBuffer, targets, medium, spectrum):
    """Upload medium, spectrum and targets to the context."""
    lib = _lib.load_library()
    m = medium.to_pod()
    _lib.check(lib.pp_set_medium(dev.handle, ctypes.byref(m)), "pp_set_medium")
    mono, tab, n, mapping = _spectrum_args(spectrum)
    tp = None if tab is None else tab.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    _lib.check(lib.pp_set_spectrum(dev.handle, mono, tp, n, mapping), "pp_set_spectrum")
    flat = flatten_targets(targets)
    pods = targets_to_pod(flat)
    _lib.check(lib.pp_set_targets(dev.handle, pods, len(flat)), "pp_set_targets")


def sources_to_pod(sources):
    arr = (_lib.pp_source * max(len(sources), 1))()
    for i, s in enumerate(sources):
        arr[i] = s.to_pod()
    return arr


def run_photon_prop_no_local_cache(sources, targets, medium, spectrum, seed,
                                   hit_buffer_cpu: HostHitBuffer,
                                   hit_buffer_gpu: DeviceHitBuffer, n_steps=15,
                                   photon_scaling=1.0, with_ids=False):
    """src/photon_prop_cuda.jl:889-959. Returns (hits, n_ph_sim): hits is a
    dict of numpy columns (the PhotonHit fields; module_id is the 1-based
    target index). Raises PhotonPropOOMException on hit-buffer overflow."""
    lib = _lib.load_library()
    dev = hit_buffer_gpu
    stage(dev, targets, medium, spectrum)
    pods = sources_to_pod(sources)
    n_hits = ctypes.c_int64()
    n_sim = ctypes.c_int64()
    rc = lib.pp_propagate(dev.handle, pods, len(sources), ctypes.c_uint64(seed & (2**64 - 1)),
                          int(n_steps), float(photon_scaling), ctypes.byref(n_hits),
                          ctypes.byref(n_sim))
    if rc == _lib.PP_ERR_HIT_OVERFLOW:
        raise PhotonPropOOMException()
    _lib.check(rc, "pp_propagate")
    n = n_hits.value
    dev.n_hits = n
    ttype = np.float64 if hit_buffer_cpu.time_type in (np.float64, float) else np.float32
    cols = dict(position=np.empty((n, 3), np.float32), direction=np.empty((n, 3), np.float32),
                initial_direction=np.empty((n, 3), np.float32),
                wavelength=np.empty(n, np.float32), time=np.empty(n, ttype),
                dist_travelled=np.empty(n, np.float32), module_id=np.empty(n, np.int64),
                n_steps=np.empty(n, np.int32))
    ids = dict(photon_id=np.empty(n, np.uint64), source_ix=np.empty(n, np.int32))
    f = _lib.fptr
    _lib.check(lib.pp_copy_hits(dev.handle, 0, n, f(cols["position"]), f(cols["direction"]),
                                f(cols["initial_direction"]), f(cols["wavelength"]),
                                f(cols["time"]), int(ttype is np.float64),
                                f(cols["dist_travelled"]), f(cols["module_id"]),
                                f(cols["n_steps"]),
                                f(ids["photon_id"]) if with_ids else None,
                                f(ids["source_ix"]) if with_ids else None), "pp_copy_hits")
    if with_ids:
        cols.update(ids)
    return cols, n_sim.value


def calc_total_weight(df: pd.DataFrame, setup) -> pd.DataFrame:
    """src/processing.jl:32-44: abs_weight = exp(-Float64(dist ./ abs_len(wl))),
    total_weight = base_weight * abs_weight. The medium is Float32, so the
    absorption length and the ratio dist ./ abs_len are Float32 values and
    only the exponential is taken in Float64, as in the reference. An empty
    table is returned unchanged (no weight columns are added)."""
    if len(df) == 0:
        return df
    abs_len = setup.medium.absorption_length(df["wavelength"].to_numpy(np.float64))
    ratio = df["dist_travelled"].to_numpy(np.float32) / abs_len.astype(np.float32)
    df["abs_weight"] = np.exp(-ratio.astype(np.float64))
    df["total_weight"] = df["base_weight"].to_numpy() * df["abs_weight"].to_numpy()
    return df


def hits_to_dataframe(cols, vector_columns=True):
    """DataFrame in the reference's column order. With vector_columns the
    position/direction/initial_direction cells are length-3 arrays (as the
    SVector cells of the reference); otherwise they are expanded to
    pos_x..dir_z columns."""
    n = len(cols["wavelength"])
    data = {}
    for name in HIT_COLUMNS:
        v = cols[name]
        if name in ("position", "direction", "initial_direction"):
            if vector_columns:
                s = np.empty(n, dtype=object)
                s[:] = list(v)
                data[name] = s
            else:
                pre = {"position": "pos", "direction": "dir", "initial_direction": "init_dir"}[name]
                for k, ax in enumerate("xyz"):
                    data[f"{pre}_{ax}"] = v[:, k]
        else:
            data[name] = v
    for extra in ("photon_id", "source_ix"):
        if extra in cols:
            data[extra] = cols[extra]
    return pd.DataFrame(data)


def propagate_photons(setup, hit_buffer_cpu=None, hit_buffer_gpu=None, steps=15, *,
                      reserved_memory_fraction=0.7, copy_output=False, vector_columns=True,
                      with_ids=False):
    """src/processing.jl:47-75. ``propagate_photons(setup)`` allocates hit
    buffers (reserved_memory_fraction of device memory);
    ``propagate_photons(setup, hbc, hbg, steps)`` reuses the caller's.
    Returns the hit DataFrame with module ids remapped to the targets'
    module_id, base_weight = sum(photons)/n_ph_sim and absorption weights."""
    if isinstance(hit_buffer_cpu, (int, np.integer)) and hit_buffer_gpu is None:
        steps, hit_buffer_cpu = int(hit_buffer_cpu), None
    own = hit_buffer_gpu is None
    if own:
        hit_buffer_cpu, hit_buffer_gpu = make_hit_buffers(
            np.float32, reserved_memory_fraction)
    try:
        hits, n_ph_sim = run_photon_prop_no_local_cache(
            setup.sources, setup.targets, setup.medium, setup.spec_dist, setup.seed,
            hit_buffer_cpu, hit_buffer_gpu, n_steps=steps,
            photon_scaling=setup.photon_scaling, with_ids=with_ids)
    finally:
        if own:
            hit_buffer_gpu.close()
    return assemble_hit_table(hits, n_ph_sim, setup, vector_columns)


def base_weight(total_photons: int, n_ph_sim: int) -> float:
    """sum(source.photons) / n_ph_sim of src/processing.jl:70: Julia's Int64 /
    Int64 is a Float64 division, so a run that simulated no photon gives
    Inf (NaN if the sources hold none)."""
    num, den = float(total_photons), float(n_ph_sim)
    if den == 0.0:
        return float("nan") if num == 0.0 else float("inf")
    return num / den


def assemble_hit_table(hits, n_ph_sim, setup, vector_columns=True) -> pd.DataFrame:
    """The host part of propagate_photons after the propagation
    (src/processing.jl:67-72, :32-44): the hit columns in the PhotonHit
    order, module_id remapped from the 1-based target index to the target's
    module_id (mod_ids[df.module_id]), base_weight = sum(photons) / n_ph_sim,
    then calc_total_weight (abs_weight, total_weight)."""
    df = hits_to_dataframe(hits, vector_columns)
    # bare shapes (and the shapes of a DetectorLines) keep their 1-based index
    mod_ids = np.array([getattr(t, "module_id", i + 1)
                        for i, t in enumerate(flatten_targets(setup.targets))], dtype=np.int64)
    if len(df):
        ix = df["module_id"].to_numpy()
        if ix.min() < 1 or ix.max() > len(mod_ids):
            raise IndexError("hit module index outside the target list")
        df["module_id"] = mod_ids[ix - 1]
    df["base_weight"] = base_weight(sum(int(s.photons) for s in setup.sources), n_ph_sim)
    calc_total_weight(df, setup)
    return df
