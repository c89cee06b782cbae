"""PMT-level detection of the POM module (src/pom.jl, src/processing.jl:88-205).

A photon that reached a module's sphere is turned into a PMT hit (1..16) or
dropped by pp_pom_pmt_hits on the GPU: per-PMT Rayleigh angular acceptance
around each PMT's direction times the wavelength acceptance of its PMT group
(assets/pmt_acc.hd5 of the reference, converted to data/pom_acceptance.csv by
tests/golden/convert_pom_acceptance.py). Its random draws come from the photon's own
Philox stream, so a photon's PMT does not depend on how hits are batched.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pandas as pd

from . import _lib
from .detection import PhotonTarget, Spherical
from .spectrum import pmt_acceptance_curve

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
N_PMT = 16
PMT_GROUP = np.array([1, 1, 1, 1, 2, 2, 2, 2, 1, 1, 1, 1, 2, 2, 2, 2])  # src/pom.jl:172-174


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class POMAcceptance:
    """POMAcceptance (src/pom.jl:17-35): Rayleigh widths and total wavelength
    acceptances of the two PMT groups."""

    def __init__(self, sigma_1, sigma_2, wavelengths, acc_1, acc_2):
        self.sigma_1, self.sigma_2 = float(sigma_1), float(sigma_2)
        self.wavelengths = np.asarray(wavelengths, np.float64)
        self.acc_1 = np.asarray(acc_1, np.float64)
        self.acc_2 = np.asarray(acc_2, np.float64)

    @classmethod
    def from_file(cls, path=None):
        path = path or os.path.join(DATA, "pom_acceptance.csv")
        with open(path) as f:
            head = f.readline()
        meta = dict(kv.strip().split("=") for kv in head.lstrip("# ").split(";")[1:])
        tab = np.loadtxt(path, delimiter=",", skiprows=2)
        return cls(float(meta["sigma_grp_1"]), float(meta["sigma_grp_2"]), tab[:, 0], tab[:, 1],
                   tab[:, 2])

    def pos_wl_acc(self, group, wl):
        acc = self.acc_1 if group == 1 else self.acc_2
        return np.interp(np.asarray(wl, np.float64), self.wavelengths, acc, left=0.0, right=0.0)


_DEFAULT_ACC = None


def default_acceptance():
    global _DEFAULT_ACC
    if _DEFAULT_ACC is None:
        _DEFAULT_ACC = POMAcceptance.from_file()
    return _DEFAULT_ACC


class POM(PhotonTarget):
    """POM(position, module_id) (src/pom.jl:69-104): a sphere of radius 0.3 m
    with 16 PMTs of 75 mm diameter."""

    def __init__(self, position, module_id, acceptance: POMAcceptance | None = None):
        self.shape = Spherical(np.asarray(position, np.float32), 0.3)
        self.pmt_area = (75e-3 / 2) ** 2 * np.pi
        self.module_id = int(module_id)
        self.acceptance = acceptance or default_acceptance()
        self.quantum_eff = pmt_acceptance_curve()  # InterpQuantumEff(PMTAcc.csv, relative)

    @staticmethod
    def pmt_count():
        return N_PMT


def make_pom_pmt_coordinates():
    """(theta, phi) of the 16 PMTs, shape (2, 16), after the reference's
    rotation of e_z onto e_x (src/pom.jl:108-141)."""
    d = get_pmt_positions(None)
    return np.stack([np.arccos(np.clip(d[:, 2], -1, 1)), np.arctan2(d[:, 1], d[:, 0])])


def _model(acc: POMAcceptance, orientation):
    m = _lib.pp_pom_model()
    m.sigma_grp1, m.sigma_grp2 = acc.sigma_1, acc.sigma_2
    o = np.eye(3) if orientation is None else np.asarray(orientation, np.float64)
    for i, v in enumerate(o.reshape(9)):
        m.orientation[i] = float(v)
    m.n_wl = len(acc.wavelengths)
    return m


def get_pmt_positions(target, orientation=None):
    """Unit directions of the 16 PMTs (16 x 3) of a POM with `orientation`."""
    acc = target.acceptance if target is not None else default_acceptance()
    m = _model(acc, orientation)
    out = np.zeros((N_PMT, 3), np.float32)
    _lib.check(_lib.load_library().pp_pom_pmt_directions(ctypes.byref(m), _p(out)),
               "pp_pom_pmt_directions")
    return out


def check_pmt_hit(hit_positions, hit_directions, hit_wavelengths, target: POM, orientation=None,
                  *, photon_id=None, source_ix=None, seed=0, device=0):
    """check_pmt_hit(::POM{<:POMAcceptance}) (src/pom.jl:246-305) on the GPU:
    PMT id (1..16) per hit, 0 if no PMT detects it. Positions may be of hits
    on one module (target) or rows with a 'center' per hit (see
    make_hits_from_photons)."""
    pos = np.asarray(hit_positions, np.float64).reshape(-1, 3)
    centers = np.broadcast_to(np.asarray(target.shape.position, np.float64), pos.shape)
    return _pmt_ids((pos - centers) / target.shape.radius, hit_wavelengths, target.acceptance,
                    orientation, photon_id, source_ix, seed, device)


def _pmt_ids(rel_pos, wl, acc, orientation, photon_id, source_ix, seed, device, timing=None):
    n = len(rel_pos)
    rp = np.ascontiguousarray(rel_pos, np.float32)
    w = np.ascontiguousarray(wl, np.float32)
    pid = np.ascontiguousarray(np.arange(n) if photon_id is None else photon_id, np.uint64)
    src = np.ascontiguousarray(np.zeros(n) if source_ix is None else source_ix, np.int32)
    grid = np.ascontiguousarray(acc.wavelengths, np.float32)
    a1 = np.ascontiguousarray(acc.acc_1, np.float32)
    a2 = np.ascontiguousarray(acc.acc_2, np.float32)
    out = np.zeros(n, np.int32)
    nd = ctypes.c_int64()
    ms = ctypes.c_float()
    m = _model(acc, orientation)
    _lib.check(_lib.load_library().pp_pom_pmt_hits(
        device, ctypes.byref(m), _p(grid), _p(a1), _p(a2), n, _p(rp), _p(w), _p(pid), _p(src),
        ctypes.c_uint64(seed), _p(out), ctypes.byref(nd), ctypes.byref(ms)), "pp_pom_pmt_hits")
    if timing is not None:
        timing["kernel_ms"] = float(ms.value)
        timing["n_detected"] = int(nd.value)
    return out


def pmt_ids_in_context(dev, acceptance: POMAcceptance | None = None, orientation=None, *,
                       seed=0, first=0, n=None, timing=None):
    """PMT ids of the hits of the context's last propagation, computed on the
    device from its hit buffer (pp_pom_pmt_hits_ctx): rows [first, first + n)
    in the order of run_photon_prop_no_local_cache's columns."""
    lib = _lib.load_library()
    acc = acceptance or default_acceptance()
    count = int(dev.n_hits) - int(first) if n is None else int(n)
    grid = np.ascontiguousarray(acc.wavelengths, np.float32)
    a1 = np.ascontiguousarray(acc.acc_1, np.float32)
    a2 = np.ascontiguousarray(acc.acc_2, np.float32)
    out = np.zeros(max(count, 0), np.int32)
    nd = ctypes.c_int64()
    ms = ctypes.c_float()
    m = _model(acc, orientation)
    _lib.check(lib.pp_pom_pmt_hits_ctx(dev.handle, ctypes.byref(m), _p(grid), _p(a1), _p(a2),
                                       ctypes.c_uint64(seed), int(first), len(out), _p(out),
                                       ctypes.byref(nd), ctypes.byref(ms)), "pp_pom_pmt_hits_ctx")
    if timing is not None:
        timing["kernel_ms"] = float(ms.value)
        timing["n_detected"] = int(nd.value)
    return out


def make_hits_from_photons(df: pd.DataFrame, targets, orientation=None, apply_wl_acc=True, *,
                           seed=0, device=0):
    """make_hits_from_photons (src/processing.jl:88-144): the photon rows of
    `df` (propagate_photons output) that a PMT of their POM detects, with a
    pmt_id column. All modules in one GPU call."""
    if len(df) == 0:
        out = df.copy()
        out["pmt_id"] = np.zeros(0, np.int32)
        return out
    by_id = {t.module_id: t for t in targets}
    mid = df["module_id"].to_numpy()
    if "pos_x" in df.columns:
        pos = df[["pos_x", "pos_y", "pos_z"]].to_numpy(np.float64)
    else:
        pos = np.stack(df["position"].to_numpy()).astype(np.float64)
    centers = np.stack([np.asarray(by_id[int(k)].shape.position, np.float64) for k in mid])
    radii = np.array([by_id[int(k)].shape.radius for k in mid], np.float64)[:, None]
    acc = by_id[int(mid[0])].acceptance
    pid = df["photon_id"].to_numpy() if "photon_id" in df.columns else None
    src = df["source_ix"].to_numpy() if "source_ix" in df.columns else None
    ids = _pmt_ids((pos - centers) / radii, df["wavelength"].to_numpy(), acc, orientation, pid,
                   src, seed, device)
    # apply_wl_acceptance(::POM) accepts everything: the wavelength acceptance
    # is part of check_pmt_hit (src/pom.jl:186-195)
    keep = ids > 0
    out = df[keep].copy()
    out["pmt_id"] = ids[keep]
    return out.reset_index(drop=True)


def apply_qe(wavelength, target):
    """Relative quantum efficiency (InterpQuantumEff(PMTAcc.csv, relative),
    src/detection.jl:172-188)."""
    return target.quantum_eff(np.asarray(wavelength, np.float64))


def calc_pe_weight(photons: pd.DataFrame, targets):
    """calc_pe_weight! (src/processing.jl:179-197): qe_weight per row and
    total_weight multiplied by it."""
    by_id = {t.module_id: t for t in targets}
    qe = np.zeros(len(photons), np.float64)
    mid = photons["module_id"].to_numpy()
    wl = photons["wavelength"].to_numpy()
    for k in np.unique(mid):
        sel = mid == k
        qe[sel] = apply_qe(wl[sel], by_id[int(k)])
    photons["qe_weight"] = qe
    photons["total_weight"] = photons["total_weight"].to_numpy() * qe
    return photons
