"""Block shapes (sigmasph_amd/geometry.py) against the reference's stated geometry.

- The dogbone resolution study's particle counts, stated by the reference's manual
  (section 5.3.9 and Table 5.2): 640 particles at dx0 = 0.4 mm, 2508 at 0.2 and 40 114
  at 0.05 for the `dogboneaxi` deck -- the lattice and the carve together reproduce them
  exactly.
- The profile: W_g/2 at the centre, W_t/2 from the junction d_j = sqrt(D (2R - D)) on,
  continuous there, total length 2 (d_j + H_t); the carve is mirror symmetric about the
  centre; the revolved and half-section forms agree with the slab through w.
- Circle, sphere, cylinder: the points within the radius (plus the 1e-6 tolerance).
"""
import math

import numpy as np
import pytest

from sigmasph_amd.config import SimConfig
from sigmasph_amd.geometry import DogBone2D, DogBone3D, DogBoneAxi, region_from_spec
from sigmasph_amd.particle_system import SceneLayout, lattice_points
from sigmasph_amd.scenes import tensile_axi_dogbone_nofail


@pytest.mark.parametrize("dx0,count", [(0.4, 640), (0.2, 2508), (0.05, 40114)])
def test_reference_dogbone_particle_counts(dx0, count):
    lay = SceneLayout(SimConfig(tensile_axi_dogbone_nofail(dx0)))
    assert len(lay.x_init) == count
    x = lay.x_init
    assert x[:, 1].min() == pytest.approx(0.5 * dx0) and np.all(x[:, 2] == 0.0)


def test_dogbone_profile():
    d = DogBone2D(10.0, 3.6, 6.0, 20.36, centre=(0.0, 0.0, 0.0))
    D = 2.0
    yj = math.sqrt(D * (2 * 20.36 - D))
    assert d.shoulder_height == pytest.approx(yj, rel=1e-15)
    assert d.total_height == pytest.approx(2 * (yj + 3.6), rel=1e-15)
    assert d.half_width(0.0) == pytest.approx(3.0, rel=1e-15)
    assert d.half_width(yj * (1 - 1e-12)) == pytest.approx(5.0, rel=1e-9)
    assert d.half_width(yj + 1.0) == 5.0
    with pytest.raises(ValueError):
        DogBone2D(10.0, 3.6, 6.0, 1.0)          # radius below the reduction
    with pytest.raises(ValueError):
        DogBone2D(6.0, 3.6, 10.0, 20.0)          # gauge wider than the tab


def test_carve_symmetry_and_forms():
    pts = lattice_points(np.array([-15.0, -15.0, -6.0]), np.array([30.0, 30.0, 12.0]), 0.25, 3)
    c = pts.mean(axis=0)
    slab = DogBone2D(10.0, 3.6, 6.0, 20.36, c).contains(pts)
    mirrored = DogBone2D(10.0, 3.6, 6.0, 20.36, c).contains(2 * c - pts)
    assert np.array_equal(slab, mirrored)
    bar = DogBone3D(10.0, 3.6, 6.0, 20.36, c).contains(pts)
    r = np.hypot(pts[:, 0] - c[0], pts[:, 2] - c[2])
    w = DogBone2D(10.0, 3.6, 6.0, 20.36, c).half_width(pts[:, 1])
    inside_len = np.abs(pts[:, 1] - c[1]) <= 0.5 * DogBone2D(10, 3.6, 6, 20.36).total_height + 1e-5
    assert np.array_equal(bar, inside_len & (r <= w + 1e-5))
    half = pts[:, [1, 0, 2]].copy()       # (axial, radius, 0) of the same points
    axi = DogBoneAxi(10.0, 3.6, 6.0, 20.36, (c[1], 0.0, 0.0)).contains(half)
    assert np.array_equal(axi, inside_len & (pts[:, 0] >= -1e-5) & (pts[:, 0] <= w + 1e-5))


def test_round_shapes_and_errors():
    pts = lattice_points(np.array([-2.0, -2.0, -2.0]), np.array([4.0, 4.0, 4.0]), 0.2, 3)
    circ = region_from_spec({"type": "circle", "radius": 1.0, "center": [0, 0, 0]}, None)
    assert np.array_equal(circ.contains(pts), np.hypot(pts[:, 0], pts[:, 1]) <= 1.0 + 1e-6)
    sph = region_from_spec({"type": "sphere", "radius": 1.0}, np.zeros(3))
    assert np.array_equal(sph.contains(pts), np.linalg.norm(pts, axis=1) <= 1.0 + 1e-6)
    cyl = region_from_spec({"type": "cylinder", "radius": 1.0, "axis": "y"}, np.zeros(3))
    assert np.array_equal(cyl.contains(pts), np.hypot(pts[:, 0], pts[:, 2]) <= 1.0 + 1e-6)
    with pytest.raises(ValueError, match="unknown block shape"):
        region_from_spec({"type": "hexagon"}, np.zeros(3))
