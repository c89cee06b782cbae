"""ctypes wrapper of the CPU oracle (oracle/libpp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference values the GPU results are
compared against. The product path
(photonprop, libphotonprop_hip.so) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, Structure, c_float, c_int, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpp_oracle.so")

FN = dict(ISOTROPIC_DIR=0, ROT_FROM_EZ=1, HG_COS=2, ES_COS=3, UPDATE_DIR=4, SPHERE_ISEC=5,
          STEP_LEN=6, CASCADIA=7, TRACK_DIR=8, PHILOX=9, GAMMA=10, SPECTRUM=11, SL_COS=12,
          PLANE_ISEC=13, MIXED_HGES=14)
FN_ARITY = {0: (2, 3), 1: (6, 3), 2: (2, 1), 3: (2, 1), 4: (5, 3), 5: (11, 2), 6: (2, 1),
            7: (1, 3), 8: (2, 3), 9: (6, 4), 10: (7, 2), 11: (1, 1), 12: (2, 1),
            13: (13, 2), 14: (5, 1)}

HIT_DTYPE = np.dtype([("pos", np.float32, 3), ("dir", np.float32, 3),
                      ("init_dir", np.float32, 3), ("wavelength", np.float32),
                      ("time", np.float32), ("dist", np.float32), ("module_ix", np.int64),
                      ("n_steps", np.int32), ("source_ix", np.int32), ("photon_id", np.uint64)])
assert HIT_DTYPE.itemsize == 72


class ppo_spectrum(Structure):
    _fields_ = [("mono_wl", c_float), ("table", POINTER(c_float)), ("n", c_int),
                ("mapping", c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ppo_eval_fn.restype = c_int
        L.ppo_eval_fn.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int,
                                  c_int64]
        L.ppo_propagate.restype = c_int64
        L.ppo_propagate.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int,
                                    c_uint64, c_int, c_float, c_void_p, c_int64,
                                    POINTER(c_int64), c_int]
        L.ppf_propagate.restype = c_int64
        L.ppf_propagate.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int,
                                    ctypes.c_uint64, c_int, c_float, c_void_p, c_int64,
                                    c_void_p, c_void_p, c_int]
        L.ppo_photon_count.restype = c_int64
        L.ppo_photon_count.argtypes = [c_int64, c_float]
        L.ppo_quan_fry_n.restype = ctypes.c_double
        L.ppo_quan_fry_n.argtypes = [ctypes.c_double] * 4
        L.ppo_quan_fry_group_velocity.restype = ctypes.c_double
        L.ppo_quan_fry_group_velocity.argtypes = [ctypes.c_double] * 4
        L.ppo_kopelevich_sca_len.restype = c_float
        L.ppo_kopelevich_sca_len.argtypes = [c_float] * 4
        L.ppo_beta_sample.restype = c_float
        L.ppo_beta_sample.argtypes = [c_float, c_float, c_uint64, c_uint64, c_uint32]
        L.ppo_gamma_sample.restype = c_float
        L.ppo_gamma_sample.argtypes = [c_float, c_float, c_uint64, c_uint64, c_uint32]
        L.ppo_bf16.restype = c_float
        L.ppo_bf16.argtypes = [c_float]
        L.ppo_digamma.restype = ctypes.c_double
        L.ppo_digamma.argtypes = [ctypes.c_double]
        L.ppo_loss_head.restype = None
        L.ppo_loss_head.argtypes = [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]
        L.ppo_mlp_forward.restype = c_int
        L.ppo_mlp_forward.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_void_p]
        L.ppo_mlp_loss_grad.restype = c_int
        L.ppo_mlp_loss_grad.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int,
                                        c_int, POINTER(ctypes.c_double), c_void_p]
        L.ppo_adam.restype = None
        L.ppo_adam.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int] + \
            [ctypes.c_double] * 5
        L.ppo_pom_pmt_dirs.restype = None
        L.ppo_pom_pmt_dirs.argtypes = [c_void_p, c_void_p]
        L.ppo_pom_pmt_hits.restype = None
        L.ppo_pom_pmt_hits.argtypes = [c_void_p, c_float, c_float, c_int, c_void_p, c_void_p,
                                       c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_uint64, c_void_p, c_void_p]
        L.ppo_sur_features.restype = None
        L.ppo_sur_features.argtypes = [c_void_p, c_void_p, c_int64, c_void_p]
        _lib = L
    return _lib


def _spec_struct(mono_wl=0.0, table=None, mapping=0):
    s = ppo_spectrum()
    s.mono_wl = mono_wl
    keep = None
    if table is not None:
        keep = np.ascontiguousarray(table, dtype=np.float32)
        s.table = keep.ctypes.data_as(POINTER(c_float))
        s.n = len(keep)
        s.mapping = mapping
    return s, keep


def eval_fn(fn, inputs, medium=None, spectrum=None):
    """Evaluate one oracle step function over rows of `inputs`.
    medium: a pp_medium POD; spectrum: (mono_wl, table, mapping)."""
    fid = FN[fn] if isinstance(fn, str) else int(fn)
    n_in, n_out = FN_ARITY[fid]
    x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, n_in)
    y = np.zeros((x.shape[0], n_out), dtype=np.float32)
    mp = ctypes.byref(medium) if medium is not None else None
    sp, keep = (None, None)
    spp = None
    if spectrum is not None:
        sp, keep = _spec_struct(*spectrum)
        spp = ctypes.byref(sp)
    rc = lib().ppo_eval_fn(fid, mp, spp, x.ctypes.data_as(c_void_p), n_in,
                           y.ctypes.data_as(c_void_p), n_out, x.shape[0])
    if rc != 0:
        raise RuntimeError(f"ppo_eval_fn({fn}) failed")
    return y


def philox(ctr, key):
    """Raw Philox4x32-10: ctr (n,4) uint32, key (n,2) uint32 -> (n,4) uint32."""
    ctr = np.asarray(ctr, dtype=np.uint32).reshape(-1, 4)
    key = np.asarray(key, dtype=np.uint32).reshape(-1, 2)
    x = np.concatenate([ctr, key], axis=1).view(np.float32)
    return eval_fn("PHILOX", x).view(np.uint32)


def photon_count(photons, scaling):
    return lib().ppo_photon_count(int(photons), float(scaling))


def propagate(sources, medium, targets, seed, n_steps=15, photon_scaling=1.0, mono_wl=0.0,
              table=None, mapping=0, capacity=1 << 20, n_threads=0):
    """Full oracle propagation. sources/targets: ctypes arrays of pp_source /
    pp_target; medium: pp_medium. Returns (hits structured array sorted by
    (source, photon id), n_ph_sim). Raises OverflowError past capacity."""
    hits = np.zeros(capacity, dtype=HIT_DTYPE)
    sp, keep = _spec_struct(mono_wl, table, mapping)
    nsim = c_int64()
    n = lib().ppo_propagate(ctypes.cast(sources, c_void_p), len(sources), ctypes.byref(medium),
                            ctypes.byref(sp), ctypes.cast(targets, c_void_p), len(targets),
                            ctypes.c_uint64(seed), int(n_steps), float(photon_scaling),
                            hits.ctypes.data_as(c_void_p), capacity, ctypes.byref(nsim),
                            int(n_threads))
    del keep
    if n < 0:
        raise OverflowError("oracle hit capacity exceeded")
    return hits[:n].copy(), nsim.value


def propagate_f32(sources, medium, targets, seed, n_steps=15, photon_scaling=1.0, mono_wl=0.0,
                  table=None, mapping=0, capacity=1 << 20, n_threads=0):
    """Float32 port of the reference kernel (pp_ref_f32.c, the CPU
    baseline). Returns (hits sorted by (source, photon id), n_ph_sim,
    photon steps executed)."""
    hits = np.zeros(capacity, dtype=HIT_DTYPE)
    sp, keep = _spec_struct(mono_wl, table, mapping)
    nsim, nstep = c_int64(), c_int64()
    n = lib().ppf_propagate(ctypes.cast(sources, c_void_p), len(sources), ctypes.byref(medium),
                            ctypes.byref(sp), ctypes.cast(targets, c_void_p), len(targets),
                            ctypes.c_uint64(seed), int(n_steps), float(photon_scaling),
                            hits.ctypes.data_as(c_void_p), capacity, ctypes.byref(nsim),
                            ctypes.byref(nstep), int(n_threads))
    del keep
    if n == -2:
        raise ValueError("the Float32 port does not cover this source, medium or target kind")
    if n < 0:
        raise OverflowError("hit capacity exceeded")
    h = hits[:n]
    h = h[np.lexsort((h["photon_id"], h["source_ix"]))].copy()
    return h, nsim.value, nstep.value


# ---- surrogate MLP (pp_mlp_oracle.c) ----

def _p(a):
    return a.ctypes.data_as(c_void_p)


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    special = (x.view(np.uint32) & 0x7F800000) == 0x7F800000
    return np.where(special, x, r.view(np.float32))


def mlp_forward(dims, params, x, quant=True):
    dims = np.ascontiguousarray(dims, np.int32)
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    n = x.shape[0]
    y = np.zeros((n, dims[-1]), np.float64)
    assert lib().ppo_mlp_forward(_p(dims), len(dims) - 1, _p(params), _p(x), n, int(quant),
                                 _p(y)) == 0
    return y


def mlp_loss_grad(dims, params, x, labels, quant=True):
    dims = np.ascontiguousarray(dims, np.int32)
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    labels = np.ascontiguousarray(labels, np.float32)
    g = np.zeros(params.size, np.float64)
    loss = ctypes.c_double()
    assert lib().ppo_mlp_loss_grad(_p(dims), len(dims) - 1, _p(params), _p(x), _p(labels),
                                   x.shape[0], int(quant), ctypes.byref(loss), _p(g)) == 0
    return loss.value, g


def loss_head(y, labels):
    y = np.ascontiguousarray(y, np.float64)
    labels = np.ascontiguousarray(labels, np.float32)
    n, n_out = y.shape
    ls = np.zeros(n, np.float64)
    dy = np.zeros((n, n_out), np.float64)
    lib().ppo_loss_head(_p(y), n_out, _p(labels), n, _p(ls), _p(dy))
    return ls, dy


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    p, g, m, v = (np.array(a, np.float32) for a in (p, g, m, v))
    lib().ppo_adam(_p(p), _p(g), _p(m), _p(v), p.size, int(t), lr, b1, b2, eps, wd)
    return p, m, v


def sur_features(cascades, modules):
    """cascades (n, 8): x, y, z, dx, dy, dz, energy, log_yield; modules (n, 3)."""
    c = np.ascontiguousarray(cascades, np.float32)
    m = np.ascontiguousarray(modules, np.float32)
    out = np.zeros((c.shape[0], 8), np.float32)
    lib().ppo_sur_features(_p(c), _p(m), c.shape[0], _p(out))
    return out


# ---- POM PMT detection (pp_pmt_oracle.c) ----

def pom_pmt_dirs(orientation=None):
    o = np.ascontiguousarray(np.eye(3) if orientation is None else orientation, np.float32)
    d = np.zeros((16, 3), np.float64)
    lib().ppo_pom_pmt_dirs(_p(o), _p(d))
    return d


def pom_pmt_hits(rel_pos, wl, pid, src, seed, wl_grid, acc1, acc2, sigma1, sigma2,
                 orientation=None):
    o = np.ascontiguousarray(np.eye(3) if orientation is None else orientation, np.float32)
    rp = np.ascontiguousarray(rel_pos, np.float32)
    w = np.ascontiguousarray(wl, np.float32)
    pid = np.ascontiguousarray(pid, np.uint64)
    src = np.ascontiguousarray(src, np.int32)
    g, a1, a2 = (np.ascontiguousarray(v, np.float32) for v in (wl_grid, acc1, acc2))
    out = np.zeros(len(w), np.int32)
    ps = np.zeros(len(w), np.float64)
    lib().ppo_pom_pmt_hits(_p(o), sigma1, sigma2, len(g), _p(g), _p(a1), _p(a2), len(w), _p(rp),
                           _p(w), _p(pid), _p(src), ctypes.c_uint64(seed), _p(out), _p(ps))
    return out, ps
