"""Block shapes: analytic regions that carve a block's particle lattice into a specimen
(geometry.py:1-548 of the reference, same shape names, parameters and containment
rules; the scene's `shape` entry is read by `region_from_spec`).

- `dogbone` (plane-strain slab, long axis y), `dogbone3d` (the profile revolved about
  y), `dogboneaxi` (the meridional half-section of that bar for an axisymmetric deck:
  long axis x, radius y >= 0 measured from the axis).  The profile's half-width at
  distance d from the centre along the long axis is W_g/2 + R - sqrt(R^2 - d^2) up to
  the junction d_j = sqrt(D (2R - D)), D = (W_t - W_g)/2, and W_t/2 beyond it, out to
  half the derived length d_j + H_t.
- `circle` (a disk in x-y: a cylinder along z in a 3-D block), `sphere`, `cylinder`
  (about one axis, unbounded along it).

Points on a boundary are kept with a tolerance of 1e-6 of the shape's size (1e-6 of
W_t for the dogbones, of the radius otherwise), which keeps a lattice site lying on the
profile on both sides alike.  Centres default to the centroid of the lattice the block
laid (the caller passes it); `dogboneaxi` reads only the axial component.
"""
import math

import numpy as np


class _Profile:
    axial = 1          # column of the long axis
    revolved = False

    def __init__(self, tab_width, tab_height, gauge_width, radius, centre=(0.0, 0.0, 0.0)):
        self.tab_width, self.tab_height = float(tab_width), float(tab_height)
        self.gauge_width, self.radius = float(gauge_width), float(radius)
        self.centre = np.zeros(3)
        c = np.asarray(centre, dtype=np.float64).ravel()[:3]
        self.centre[:len(c)] = c
        if self.gauge_width <= 0.0 or self.tab_width <= 0.0:
            raise ValueError("dogbone: widths must be positive")
        if self.tab_height <= 0.0:
            raise ValueError("dogbone: tab_height must be positive (the grips hold the tabs)")
        if self.gauge_width > self.tab_width:
            raise ValueError("dogbone: gauge_width %g exceeds tab_width %g"
                             % (self.gauge_width, self.tab_width))
        self.reduction = 0.5 * (self.tab_width - self.gauge_width)
        if self.radius < self.reduction:
            raise ValueError("dogbone: radius %g is below (tab_width - gauge_width)/2 = %g, so the "
                             "arc never reaches the tab width" % (self.radius, self.reduction))
        self.shoulder_height = math.sqrt(self.reduction * (2.0 * self.radius - self.reduction))
        self.total_height = 2.0 * (self.shoulder_height + self.tab_height)

    def half_width(self, a):
        d = np.abs(np.asarray(a, dtype=np.float64) - self.centre[self.axial])
        dc = np.minimum(d, self.radius)
        arc = 0.5 * self.gauge_width + self.radius - np.sqrt(np.maximum(self.radius ** 2 - dc * dc, 0.0))
        return np.where(d <= self.shoulder_height, arc, 0.5 * self.tab_width)

    def _tol(self, tol):
        return 1e-6 * self.tab_width if tol is None else float(tol)

    def _within_length(self, p, tol):
        return np.abs(p[:, self.axial] - self.centre[self.axial]) <= 0.5 * self.total_height + tol

    def summary(self):
        return ("%s: tab %g x %g, gauge width %g, radius %g; shoulder %.6g, total length %.6g"
                % (type(self).__name__, self.tab_width, self.tab_height, self.gauge_width, self.radius,
                   self.shoulder_height, self.total_height))


class DogBone2D(_Profile):
    """|x - x_c| <= w(y): a slab taking the block's z layer."""

    def contains(self, points, tol=None):
        t = self._tol(tol)
        p = np.asarray(points, dtype=np.float64)
        return self._within_length(p, t) & (np.abs(p[:, 0] - self.centre[0]) <= self.half_width(p[:, 1]) + t)


class DogBone3D(_Profile):
    """hypot(x - x_c, z - z_c) <= w(y): the profile revolved about y."""
    revolved = True

    def contains(self, points, tol=None):
        t = self._tol(tol)
        p = np.asarray(points, dtype=np.float64)
        r = np.hypot(p[:, 0] - self.centre[0], p[:, 2] - self.centre[2])
        return self._within_length(p, t) & (r <= self.half_width(p[:, 1]) + t)


class DogBoneAxi(_Profile):
    """|x - x_c| <= L/2 and -tol <= y <= w(x): the half-section of DogBone3D in the
    (x, r) plane of an axisymmetric deck."""
    revolved = True
    axial = 0

    def contains(self, points, tol=None):
        t = self._tol(tol)
        p = np.asarray(points, dtype=np.float64)
        return self._within_length(p, t) & (p[:, 1] >= -t) & (p[:, 1] <= self.half_width(p[:, 0]) + t)


class _Round:
    def __init__(self, radius, centre=(0.0, 0.0, 0.0), what="shape"):
        self.radius = float(radius)
        if self.radius <= 0.0:
            raise ValueError("%s: radius must be positive" % what)
        self.centre = np.zeros(3)
        c = np.asarray(centre, dtype=np.float64).ravel()[:3]
        self.centre[:len(c)] = c

    def _tol(self, tol):
        return 1e-6 * self.radius if tol is None else float(tol)

    def summary(self):
        return "%s: radius %g, centre %s" % (type(self).__name__, self.radius, tuple(self.centre))


class Circle2D(_Round):
    def __init__(self, radius, centre=(0.0, 0.0, 0.0)):
        super().__init__(radius, centre, "circle")

    def contains(self, points, tol=None):
        p = np.asarray(points, dtype=np.float64)
        return np.hypot(p[:, 0] - self.centre[0], p[:, 1] - self.centre[1]) <= self.radius + self._tol(tol)


class Sphere3D(_Round):
    def __init__(self, radius, centre=(0.0, 0.0, 0.0)):
        super().__init__(radius, centre, "sphere")

    def contains(self, points, tol=None):
        p = np.asarray(points, dtype=np.float64)
        return np.linalg.norm(p - self.centre[None, :], axis=1) <= self.radius + self._tol(tol)


class Cylinder3D(_Round):
    AXES = {"x": 0, "y": 1, "z": 2}

    def __init__(self, radius, axis="x", centre=(0.0, 0.0, 0.0)):
        super().__init__(radius, centre, "cylinder")
        a = str(axis).lower()
        if a not in self.AXES:
            raise ValueError("cylinder: axis must be one of x, y, z; got %r" % (axis,))
        self.axis = a
        self.others = [k for k in range(3) if k != self.AXES[a]]

    def contains(self, points, tol=None):
        p = np.asarray(points, dtype=np.float64)
        u, v = self.others
        return np.hypot(p[:, u] - self.centre[u], p[:, v] - self.centre[v]) <= self.radius + self._tol(tol)


PROFILES = {"dogbone": DogBone2D, "dogbone3d": DogBone3D, "dogboneaxi": DogBoneAxi}


def region_from_spec(spec, default_centre):
    """The region of a block's `shape` entry, or None (region_from_spec,
    geometry.py:517-548).  `center` defaults to `default_centre`."""
    if not spec:
        return None
    kind = str(spec.get("type", "")).lower()
    centre = spec.get("center", default_centre)
    if kind == "circle":
        return Circle2D(spec["radius"], centre)
    if kind == "sphere":
        return Sphere3D(spec["radius"], centre)
    if kind == "cylinder":
        return Cylinder3D(spec["radius"], spec.get("axis", "x"), centre)
    if kind in PROFILES:
        return PROFILES[kind](spec["tabWidth"], spec["tabHeight"], spec["gaugeWidth"], spec["radius"], centre)
    raise ValueError("unknown block shape %r; known shapes: %s"
                     % (spec.get("type"), ", ".join(sorted(list(PROFILES) + ["circle", "sphere", "cylinder",
                                                                           "gaugeDogbone"]))))
