"""Targets: spherical optical modules and planar (rectangular, circular)
detectors.

Mirror of the NeutrinoTelescopeBase target types the reference passes to the
kernel: ``Spherical(position, radius)``, ``Rectangular(position, length_x,
length_y)`` and ``Circular(position, radius)`` shapes wrapped in targets that
carry a ``module_id`` (src/processing.jl:57-70 reads ``targ.shape`` and
``targ.module_id``). Planar shapes lie in the horizontal plane through their
position (check_intersection, src/photon_prop_cuda.jl:353-404).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class Spherical:
    position: np.ndarray
    radius: float

    def surface_area(self):
        return 4.0 * np.pi * float(self.radius) ** 2


@dataclass
class Rectangular:
    position: np.ndarray
    length_x: float
    length_y: float

    def surface_area(self):
        return float(self.length_x) * float(self.length_y)


@dataclass
class Circular:
    position: np.ndarray
    radius: float

    def surface_area(self):
        return np.pi * float(self.radius) ** 2


@dataclass
class PhotonTarget:
    shape: Spherical | Rectangular | Circular
    module_id: int


@dataclass
class HomogeneousDetector(PhotonTarget):
    """HomogeneousDetector(shape, active_area, module_id) (README.md:34-38)."""

    def __init__(self, shape, active_area, module_id):
        self.shape = shape
        self.active_area = float(active_area)
        self.module_id = int(module_id)


def spherical_target(position, radius, module_id):
    return PhotonTarget(Spherical(np.asarray(position, dtype=np.float32), float(radius)),
                        int(module_id))


def rectangular_target(position, length_x, length_y, module_id):
    return PhotonTarget(Rectangular(np.asarray(position, dtype=np.float32), float(length_x),
                                    float(length_y)), int(module_id))


def circular_target(position, radius, module_id):
    return PhotonTarget(Circular(np.asarray(position, dtype=np.float32), float(radius)),
                        int(module_id))


@dataclass
class DetectorLine:
    """DetectorLine(x_pos, y_pos, radius, targets) (src/photon_prop_cuda.jl:
    418-423): one string of target shapes; x_pos, y_pos and radius bound the
    string for the reference's own two-level line search."""
    x_pos: float
    y_pos: float
    radius: float
    targets: list


@dataclass
class DetectorLines:
    """DetectorLines(lines) (src/photon_prop_cuda.jl:425-427)."""
    lines: list


def flatten_targets(targets):
    """The target list the library searches. A DetectorLines becomes the
    shapes of its lines in line order, so target k (1-based) is shape t of
    line l with k = (l - 1) * n + t for lines of n shapes. The grid search
    replaces the reference's line kernel (src/photon_prop_cuda.jl:552-694),
    whose quirks are not reproduced: there every line is a separate 2-D test,
    only the first line hit is searched and module_ix = t + n * l counts
    from the second line on."""
    if isinstance(targets, DetectorLines):
        return [shape for line in targets.lines for shape in line.targets]
    return list(targets)


def targets_to_pod(targets):
    """Lower a target list (or a DetectorLines) to the pp_target array
    (index order preserved)."""
    targets = flatten_targets(targets)
    arr = (_lib.pp_target * max(len(targets), 1))()
    for i, t in enumerate(targets):
        shape = t.shape if hasattr(t, "shape") else t
        pos = np.asarray(shape.position, dtype=np.float32).reshape(3)
        arr[i].position[:] = [float(x) for x in pos]
        if isinstance(shape, Spherical):
            arr[i].radius = float(np.float32(shape.radius))
            arr[i].shape_kind = _lib.SHAPE_SPHERE
        elif isinstance(shape, Circular):
            arr[i].radius = float(np.float32(shape.radius))
            arr[i].shape_kind = _lib.SHAPE_DISC
        elif isinstance(shape, Rectangular):
            arr[i].length[:] = [float(np.float32(shape.length_x)),
                                float(np.float32(shape.length_y))]
            arr[i].shape_kind = _lib.SHAPE_RECTANGLE
        else:
            raise TypeError(f"target {i}: unknown shape {type(shape).__name__}")
        arr[i].module_id = int(getattr(t, "module_id", i + 1))
    return arr


def make_detector_grid(n_x, n_y, spacing, n_modules, v_spacing, radius=0.3, z_center=0.0):
    """Strings on an n_x x n_y grid (centred at the origin), each with
    n_modules spheres at v_spacing, centred on z_center. Module ids 1..N in
    string-major order."""
    xs = (np.arange(n_x) - (n_x - 1) / 2.0) * spacing
    ys = (np.arange(n_y) - (n_y - 1) / 2.0) * spacing
    zs = (np.arange(n_modules) - (n_modules - 1) / 2.0) * v_spacing + z_center
    targets = []
    mid = 1
    for x in xs:
        for y in ys:
            for z in zs:
                targets.append(spherical_target((x, y, z), radius, mid))
                mid += 1
    return targets
