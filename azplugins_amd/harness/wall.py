"""Wall geometries (HOOMD's ``hoomd.wall``): spheres, cylinders and planes.

``inside`` selects the side of a sphere or cylinder where particles interact
with it (True: the region within the radius); a plane's inside is the
half-space its normal points into.  ``open`` leaves the surface itself out of
the inside (particles exactly on it do not interact unless extrapolated).
"""

import numpy as np


class WallGeometry:
    pass


class Sphere(WallGeometry):
    def __init__(self, radius, origin=(0.0, 0.0, 0.0), inside=True, open=True):
        self.radius = float(radius)
        self.origin = tuple(float(x) for x in origin)
        self.inside = bool(inside)
        self.open = bool(open)

    def _row(self):
        return [*self.origin, self.radius, float(self.inside), float(self.open)]


class Cylinder(WallGeometry):
    def __init__(self, radius, axis, origin=(0.0, 0.0, 0.0), inside=True, open=True):
        self.radius = float(radius)
        self.axis = tuple(float(x) for x in axis)
        self.origin = tuple(float(x) for x in origin)
        self.inside = bool(inside)
        self.open = bool(open)

    def _row(self):
        return [*self.origin, *self.axis, self.radius, float(self.inside), float(self.open)]


class Plane(WallGeometry):
    def __init__(self, origin, normal, open=True):
        self.origin = tuple(float(x) for x in origin)
        self.normal = tuple(float(x) for x in normal)
        if not np.linalg.norm(self.normal) > 0.0:
            raise ValueError("the plane normal must be nonzero")
        self.open = bool(open)

    def _row(self):
        return [*self.origin, *self.normal, float(self.open)]
