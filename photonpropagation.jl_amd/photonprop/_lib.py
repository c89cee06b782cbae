"""ctypes binding of libphotonprop_hip.so (C ABI: include/photonprop_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``photonprop.build.build_library()``). Loading fails loudly when the shared
object is missing: there is no CPU fallback for the propagation path.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_float, c_int, c_int32, c_int64,
                    c_uint64, c_void_p)

import numpy as np

LIB_NAME = "libphotonprop_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

PP_OK = 0
PP_ERR_HIT_OVERFLOW = 1
PP_ERR_INVALID = -1
PP_ERR_HIP = -2
PP_ERR_NOMEM = -3
PP_ERR_STATE = -4

# enum pp_source_kind
SRC_POINTLIKE_ISOTROPIC = 0
SRC_POINTLIKE_TIME_RANGE = 1
SRC_AXICONE = 2
SRC_COLLIMATED_ISOTROPIC = 3
SRC_PENCIL = 4
SRC_EXTENDED_CHERENKOV = 5
SRC_EXTENDED_CHERENKOV_NO_SMEAR = 6
SRC_EXTENDED_CHERENKOV_CUSTOM_SMEAR = 7
SRC_POINTLIKE_CHERENKOV = 8
SRC_CHERENKOV_TRACK = 9

MEDIUM_SIMPLE = 0
MEDIUM_CASCADIA = 1
MEDIUM_ICE = 2
PHASE_HG = 0
PHASE_MIXED_HG_ES = 1
PHASE_MIXED_HG_SL = 2
SHAPE_SPHERE = 0
SHAPE_RECTANGLE = 1
SHAPE_DISC = 2
SPEC_KNOTS = 0
SPEC_TEXTURE = 1
SEARCH_AUTO = 0
SEARCH_BRUTE = 1
SEARCH_GRID = 2

MB_FMA_F32 = 0
MB_TRANS_F32 = 1
MB_MAD_U64 = 2

BATCH_AUTO = 0
BATCH_PER_SOURCE = 1
BATCH_PERSISTENT = 2
BATCH_GRAPH = 3

# enum pp_fn_id
FN = dict(ISOTROPIC_DIR=0, ROT_FROM_EZ=1, HG_COS=2, ES_COS=3, UPDATE_DIR=4, SPHERE_ISEC=5,
          STEP_LEN=6, CASCADIA=7, TRACK_DIR=8, PHILOX=9, GAMMA=10, SPECTRUM=11, SL_COS=12,
          PLANE_ISEC=13, MIXED_HGES=14)
FN_ARITY = {0: (2, 3), 1: (6, 3), 2: (2, 1), 3: (2, 1), 4: (5, 3), 5: (11, 2), 6: (2, 1),
            7: (1, 3), 8: (2, 3), 9: (6, 4), 10: (7, 2), 11: (1, 1), 12: (2, 1),
            13: (13, 2), 14: (5, 1)}


class pp_source(Structure):
    _fields_ = [("kind", c_int32), ("reserved0", c_int32), ("photons", c_int64),
                ("position", c_float * 3), ("direction", c_float * 3), ("time", c_float),
                ("time_end", c_float), ("angle", c_float), ("beam_divergence", c_float),
                ("cut_angle_cos", c_float), ("length", c_float), ("long_a", c_float),
                ("long_b", c_float), ("long_lrad", c_float), ("beta_a", c_float),
                ("beta_b", c_float), ("reserved1", c_float * 3)]


class pp_medium(Structure):
    _fields_ = [("kind", c_int32), ("phase_kind", c_int32), ("hg_g", c_float), ("es_b", c_float),
                ("hg_fraction", c_float), ("sca_scale", c_float), ("abs_scale", c_float),
                ("sca_len", c_float), ("group_velocity", c_float),
                ("phase_refractive_index", c_float), ("vol_conc_small", c_float),
                ("vol_conc_large", c_float), ("salinity", c_float), ("temperature", c_float),
                ("pressure", c_float), ("ice_b_dust_400", c_float),
                ("ice_alpha_sca_dust", c_float), ("reserved", c_float * 7)]


class pp_target(Structure):
    _fields_ = [("position", c_float * 3), ("radius", c_float), ("shape_kind", c_int32),
                ("module_id", c_int32), ("length", c_float * 2)]


class pp_config(Structure):
    _fields_ = [("block_size", c_int32), ("waves_per_simd", c_int32), ("search", c_int32),
                ("grid_cell", c_float), ("deterministic_order", c_int32),
                ("count_work", c_int32), ("batch_mode", c_int32),
                ("chunk_photons", c_int32), ("grid_cell_z", c_float),
                ("reserved", c_int32 * 7)]


assert ctypes.sizeof(pp_source) == 96
assert ctypes.sizeof(pp_medium) == 96
assert ctypes.sizeof(pp_target) == 32
assert ctypes.sizeof(pp_config) == 64


# ---- surrogate (include/photonprop_surrogate.h) ----
SUR_N_FEATURES = 8
SUR_N_OUT = 10
SUR_BINS = 64
SUR_TMAX = 2000.0
SUR_TIME_SHIFT = 20.0
SUR_LABEL_FLOATS = 2 + SUR_BINS


class pp_pom_model(Structure):
    _fields_ = [("sigma_grp1", c_float), ("sigma_grp2", c_float), ("orientation", c_float * 9),
                ("n_wl", c_int32), ("reserved", c_float * 4)]


assert ctypes.sizeof(pp_pom_model) == 64


class pp_cascade(Structure):
    _fields_ = [("position", c_float * 3), ("direction", c_float * 3), ("energy", c_float),
                ("log_yield", c_float)]


class pp_sur_domain(Structure):
    _fields_ = [("min_dist", c_float), ("max_dist", c_float), ("min_energy", c_float),
                ("max_energy", c_float), ("module_radius", c_float), ("c_ref", c_float),
                ("reserved", c_float * 2)]


class pp_adam(Structure):
    _fields_ = [("lr", c_float), ("beta1", c_float), ("beta2", c_float), ("eps", c_float),
                ("weight_decay", c_float), ("reserved", c_float * 3)]


assert ctypes.sizeof(pp_cascade) == 32 and ctypes.sizeof(pp_sur_domain) == 32
assert ctypes.sizeof(pp_adam) == 32

# (name, restype, argtypes) of every exported entry point of include/photonprop_hip.h
SIGNATURES = [
    ("pp_abi_version", c_int, []),
    ("pp_last_error", c_char_p, []),
    ("pp_device_count", c_int, []),
    ("pp_device_memory", c_int, [c_int, POINTER(c_int64), POINTER(c_int64)]),
    ("pp_create", c_int, [POINTER(c_void_p), c_int, c_int64, POINTER(pp_config)]),
    ("pp_destroy", c_int, [c_void_p]),
    ("pp_hit_record_bytes", c_int64, []),
    ("pp_set_medium", c_int, [c_void_p, POINTER(pp_medium)]),
    ("pp_set_spectrum", c_int, [c_void_p, c_float, POINTER(c_float), c_int, c_int]),
    ("pp_set_targets", c_int, [c_void_p, POINTER(pp_target), c_int]),
    ("pp_propagate", c_int, [c_void_p, POINTER(pp_source), c_int, c_uint64, c_int32, c_float,
                             POINTER(c_int64), POINTER(c_int64)]),
    ("pp_copy_hits", c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p]),
    ("pp_last_timing", c_int, [c_void_p, POINTER(c_float), POINTER(c_int32)]),
    ("pp_release_scratch", c_int, [c_void_p]),
    ("pp_stream", c_void_p, [c_void_p]),
    ("pp_search_info", c_int, [c_void_p, POINTER(c_int32), POINTER(c_float), POINTER(c_int32),
                               POINTER(c_int32)]),
    ("pp_work_counts", c_int, [c_void_p, POINTER(c_int64), POINTER(c_int64), POINTER(c_int64),
                               POINTER(c_int64)]),
    ("pp_pom_pmt_directions", c_int, [POINTER(pp_pom_model), c_void_p]),
    ("pp_pom_pmt_hits", c_int, [c_int, POINTER(pp_pom_model), c_void_p, c_void_p, c_void_p,
                                c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_uint64,
                                c_void_p, POINTER(c_int64), POINTER(c_float)]),
    ("pp_pom_pmt_hits_ctx", c_int, [c_void_p, POINTER(pp_pom_model), c_void_p, c_void_p,
                                    c_void_p, c_uint64, c_int64, c_int64, c_void_p,
                                    POINTER(c_int64), POINTER(c_float)]),
    ("pp_pom_pmt_hits_noalloc", c_int, [c_int, POINTER(pp_pom_model), c_void_p, c_void_p,
                                        c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_uint64, c_void_p, POINTER(c_int64),
                                        c_void_p]),
    ("pp_device_info", c_int, [c_int, ctypes.c_char_p, c_int, POINTER(c_int32)]),
    ("pp_microbench", c_int, [c_int, c_int, POINTER(ctypes.c_double)]),
    ("pp_get_config", c_int, [c_void_p, POINTER(pp_config)]),
    ("pp_set_config", c_int, [c_void_p, POINTER(pp_config)]),
    ("pp_mlp_create", c_int, [POINTER(c_void_p), c_int, c_int, c_int, c_int, c_int, c_int,
                              c_uint64]),
    ("pp_mlp_destroy", c_int, [c_void_p]),
    ("pp_mlp_param_count", c_int64, [c_void_p]),
    ("pp_mlp_get_params", c_int, [c_void_p, c_void_p]),
    ("pp_mlp_set_params", c_int, [c_void_p, c_void_p]),
    ("pp_mlp_forward", c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    ("pp_mlp_set_fused", c_int, [c_void_p, c_int]),
    ("pp_mlp_loss_grad", c_int, [c_void_p, c_void_p, c_void_p, c_int, POINTER(c_float),
                                 c_void_p]),
    ("pp_mlp_adam_step", c_int, [c_void_p, POINTER(pp_adam)]),
    ("pp_mlp_train", c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_int,
                             POINTER(pp_adam), c_uint64, c_void_p]),
    ("pp_mlp_eval_loss", c_int, [c_void_p, c_void_p, c_void_p, c_int64, POINTER(c_float)]),
    ("pp_sur_features", c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    ("pp_sur_eval_pairs", c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int,
                                  POINTER(pp_sur_domain), c_void_p, c_void_p, c_void_p,
                                  POINTER(c_float)]),
    ("pp_sur_hit_labels", c_int, [c_void_p, c_void_p, c_void_p, c_int, ctypes.c_double,
                                  c_void_p, c_int, c_void_p]),
    ("pp_autotune", c_int, [c_void_p, c_void_p, c_int, c_uint64, c_int32, c_float, c_int64,
                            POINTER(pp_config), POINTER(c_float)]),
    ("pp_eval_fn", c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_int64]),
]

_lib = None


class PhotonPropLibraryError(RuntimeError):
    """The HIP library is missing or returned a runtime error."""


def library_path() -> str:
    return LIB_PATH


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load libphotonprop_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("PHOTONPROP_HIP_LIB") or LIB_PATH
    if not os.path.exists(p):
        raise PhotonPropLibraryError(
            f"{p} not found: build it with __graft_entry__.build() "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = ctypes.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.pp_abi_version() != 1:
        raise PhotonPropLibraryError("ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    return load_library().pp_last_error().decode(errors="replace")


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise PhotonPropLibraryError(f"{what} failed ({rc}): {last_error()}")
    return rc


def fptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(c_void_p)


def eval_fn(ctx, fn: str | int, inputs: np.ndarray) -> np.ndarray:
    """Evaluate one device step function on rows of `inputs` (GPU)."""
    lib = load_library()
    fid = FN[fn] if isinstance(fn, str) else int(fn)
    n_in, n_out = FN_ARITY[fid]
    x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, n_in)
    y = np.zeros((x.shape[0], n_out), dtype=np.float32)
    check(lib.pp_eval_fn(ctx.handle, fid, fptr(x), n_in, fptr(y), n_out, x.shape[0]),
          "pp_eval_fn")
    return y
