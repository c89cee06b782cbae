"""Geometric arrival times and time residuals (src/calc.jl).

t_geo of a cascade is the straight-line distance to the module surface over
the group velocity at 800 nm; of a track, the Cherenkov-cone arrival time at
the module (src/calc.jl:37-67, :137-167). The time residual of a hit is
t - t_geo - t0 (src/calc.jl:283-323); the surrogate's arrival-time labels
use the same definition with the spectrum's fastest group velocity.
"""
from __future__ import annotations

import numpy as np

from .detection import PhotonTarget
from .lightyield import Particle

C_VAC_M_NS = 0.299792458  # src/calc.jl:15


def calc_tgeo(distance, c_n_or_medium, target: PhotonTarget | None = None):
    """calc_tgeo(distance, c_n) = distance / c_n; with a medium, c_n is its
    group velocity at 800 nm; with a spherical target the radius is
    subtracted first (src/calc.jl:37-67)."""
    d = np.asarray(distance, np.float64)
    if target is not None:
        d = d - float(target.shape.radius)
    c = c_n_or_medium
    if not np.isscalar(c) and hasattr(c, "group_velocity"):
        c = float(c.group_velocity(800.0))
    return d / c


def closest_approach_distance(p0, direction, pos):
    """|(pos - p0) x dir| (src/calc.jl:82-84)."""
    a = np.asarray(pos, np.float64) - np.asarray(p0, np.float64)
    return np.linalg.norm(np.cross(a, np.asarray(direction, np.float64)), axis=-1)


def closest_approach_param(p0, direction, pos):
    """(pos - p0) . dir (src/calc.jl:205-211)."""
    a = np.asarray(pos, np.float64) - np.asarray(p0, np.float64)
    return np.sum(a * np.asarray(direction, np.float64), axis=-1)


def calc_tgeo_tracks(p0, direction, pos, n_ph_or_medium, n_grp=None, wavelength=800.0):
    """Cherenkov arrival time of a track at pos (src/calc.jl:137-167):
    (dir . (pos - p0) + d (n_grp n_ph - 1) / sqrt(n_ph^2 - 1)) / c, d the
    closest-approach distance; with a medium, n_ph and n_grp at `wavelength`."""
    if n_grp is None:
        m = n_ph_or_medium
        n_ph = float(m.phase_refractive_index(wavelength))
        n_grp = C_VAC_M_NS / float(m.group_velocity(wavelength))
    else:
        n_ph = float(n_ph_or_medium)
    d = closest_approach_distance(p0, direction, pos)
    along = closest_approach_param(p0, direction, pos)
    return (along + d * (n_grp * n_ph - 1.0) / np.sqrt(n_ph**2 - 1.0)) / C_VAC_M_NS


def calc_tgeo_particle(particle: Particle, target: PhotonTarget, medium):
    """calc_tgeo(particle, target, medium) (src/calc.jl:182-188)."""
    if particle.shape() == "track":
        return calc_tgeo_tracks(particle.position, particle.direction, target.shape.position,
                                medium)
    dist = np.linalg.norm(np.asarray(particle.position, np.float64) -
                          np.asarray(target.shape.position, np.float64))
    return calc_tgeo(dist, medium, target)


def shift_particle(particle: Particle, param):
    """Move a particle along its direction by param metres, its time by
    param / c (src/calc.jl:242-250)."""
    from dataclasses import replace
    pos = np.asarray(particle.position, np.float64) + param * np.asarray(particle.direction,
                                                                          np.float64)
    return replace(particle, position=pos, time=particle.time + param / C_VAC_M_NS)


def shift_to_closest_approach(particle: Particle, pos):
    """shift_to_closest_approach (src/calc.jl:262-266)."""
    return shift_particle(particle, closest_approach_param(particle.position,
                                                           particle.direction, pos))


def calc_time_residual(df, setup):
    """calc_time_residual! (src/calc.jl:283-323): adds a `tres` column =
    time - t_geo(module) - t0 of the setup's first source; tracks use the
    Cherenkov-cone t_geo, other sources the straight distance."""
    src = setup.sources[0]
    by_id = {t.module_id: t for t in setup.targets}
    t0 = float(src.time)
    tres = np.empty(len(df), np.float64)
    mid = df["module_id"].to_numpy()
    is_track = hasattr(src, "length") and type(src).__name__ == "CherenkovTrackEmitter"
    for k in np.unique(mid):
        t = by_id[int(k)]
        if is_track:
            tg = calc_tgeo_tracks(src.position, src.direction, t.shape.position, setup.medium)
        else:
            dist = np.linalg.norm(np.asarray(src.position, np.float64) -
                                  np.asarray(t.shape.position, np.float64))
            tg = calc_tgeo(dist, setup.medium, t)
        sel = mid == k
        tres[sel] = df["time"].to_numpy()[sel] - tg - t0
    df["tres"] = tres
    return df
