"""Autotuner of the propagation launch (block size, photons per lane per work
chunk, grid cell edge) with a JSON cache.

The reference takes its launch shape from CUDA's occupancy calculator
(src/photon_prop_cuda.jl:914-920). Here pp_autotune (C ABI) times candidate
configurations on a sample of the workload; this module keys the result by
device architecture and workload shape and caches it, so that later runs of
the same setup skip the search. Results never depend on the configuration.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os

import numpy as np

from . import _lib
from .detection import targets_to_pod
from .processing import DeviceHitBuffer, sources_to_pod, stage

CONFIG_FIELDS = ("block_size", "waves_per_simd", "search", "grid_cell", "grid_cell_z",
                 "batch_mode", "chunk_photons")


def default_cache_path() -> str:
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "photonprop_hip", "autotune.json")


def workload_key(arch: str, n_cu: int, sources, targets, medium, spectrum, n_steps: int) -> str:
    """Stable key of what the best launch depends on: device, target layout,
    medium kind, spectrum size, step count and the mix of emitter kinds and
    photon counts (log2 bins)."""
    h = hashlib.sha256()
    h.update(f"{arch}|{n_cu}|{n_steps}|{type(medium).__name__}|".encode())
    tp = targets_to_pod(targets)
    h.update(bytes(tp))
    table = getattr(spectrum, "table", None)
    h.update(f"spec{0 if table is None else len(table)}|".encode())
    pods = sources_to_pod(sources)
    kinds = np.array([p.kind for p in pods], np.int64)
    nph = np.array([max(int(p.photons), 1) for p in pods], np.int64)
    hist = np.zeros((16, 48), np.int64)
    np.add.at(hist, (kinds, np.clip(np.log2(nph).astype(np.int64), 0, 47)), 1)
    h.update(hist.tobytes())
    return h.hexdigest()[:32]


def config_to_dict(cfg) -> dict:
    return {k: getattr(cfg, k) for k in CONFIG_FIELDS}


def apply_dict(cfg, d: dict):
    for k in CONFIG_FIELDS:
        if k in d:
            setattr(cfg, k, d[k])
    return cfg


def device_arch(device: int = 0) -> tuple[str, int]:
    """(architecture name, CU count) of a HIP device (pp_device_info)."""
    buf = ctypes.create_string_buffer(64)
    n_cu = ctypes.c_int32()
    _lib.check(_lib.load_library().pp_device_info(device, buf, 64, ctypes.byref(n_cu)),
               "pp_device_info")
    return buf.value.decode(), int(n_cu.value)


def autotune(dev: DeviceHitBuffer, sources, targets, medium, spectrum, seed: int, *,
             n_steps: int = 15, photon_scaling: float = 1.0, sample_photons: float = 2e7,
             cache_path: str | None = None, use_cache: bool = True) -> dict:
    """Tune `dev`'s context for this workload (or load the cached result) and
    return {"config": {...}, "ms": t, "cached": bool, "key": key}. The tuned
    configuration stays active on the context."""
    lib = _lib.load_library()
    stage(dev, targets, medium, spectrum)
    arch, n_cu = device_arch(dev.device)
    key = workload_key(arch, n_cu, sources, targets, medium, spectrum, n_steps)
    path = cache_path or default_cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
            if not isinstance(cache, dict):
                cache = {}
        except (OSError, ValueError):
            cache = {}  # unreadable or corrupt: tune again and rewrite it
    if use_cache and key in cache:
        ent = cache[key]
        cfg = _lib.pp_config()
        _lib.check(lib.pp_get_config(dev.handle, ctypes.byref(cfg)), "pp_get_config")
        _lib.check(lib.pp_set_config(dev.handle, ctypes.byref(apply_dict(cfg, ent["config"]))),
                   "pp_set_config")
        return {**ent, "cached": True, "key": key}
    pods = sources_to_pod(sources)
    best = _lib.pp_config()
    ms = ctypes.c_float()
    _lib.check(lib.pp_autotune(dev.handle, pods, len(pods), ctypes.c_uint64(seed), n_steps,
                               float(photon_scaling), int(sample_photons), ctypes.byref(best),
                               ctypes.byref(ms)), "pp_autotune")
    ent = {"config": config_to_dict(best), "ms": float(ms.value),
           "sample_photons": float(sample_photons), "arch": arch}
    if use_cache:
        cache[key] = ent
        tmp = f"{path}.{os.getpid()}.tmp"  # one writer per process; the rename is atomic
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(tmp, "w") as f:
                json.dump(cache, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass  # an unwritable cache only costs a tuning run next time
    return {**ent, "cached": False, "key": key}
