"""Projection-geometry base class and the detector part of vector geometries.

``ProjectionGeometry`` mirrors the property protocol of tomosipo
``geometry/base_projection.py:16-264``; ``DetectorVectorGeometry`` follows
``geometry/det_vec.py:59-376``: per-angle detector centre ``det_pos`` and pixel
step vectors ``det_v`` (row, "up") and ``det_u`` (column), all ``(n, 3)`` in
``(z, y, x)`` order. Pixel ``(i, j)`` has its centre at
``det_pos + (i + 1/2 - rows/2) det_v + (j + 1/2 - cols/2) det_u``, which is
consistent with ``lower_left_corner`` (``det_vec.py:317-322``).
"""
from numbers import Integral

import numpy as np

from .. import types as tt
from .. import vector_calc as vc
from ..utils import slice_interval

_EPS = 1e-8


def is_projection(g):
    return isinstance(g, ProjectionGeometry)


def is_cone(g):
    return is_projection(g) and g.is_cone


def is_parallel(g):
    return is_projection(g) and g.is_parallel


class ProjectionGeometry:
    _is_cone = False
    _is_parallel = False
    _is_vector = False

    def __init__(self, shape=1):
        self._shape = tt.to_shape2d(shape)

    def __len__(self):
        return self.num_steps

    @property
    def is_cone(self):
        return self._is_cone

    @property
    def is_parallel(self):
        return self._is_parallel

    @property
    def is_vec(self):
        return self._is_vector

    @property
    def det_shape(self):
        return self._shape

    @property
    def num_steps(self):
        return self.num_angles


def det_vec(shape, det_pos, det_v, det_u):
    return DetectorVectorGeometry(shape, det_pos, det_v, det_u)


class DetectorVectorGeometry(ProjectionGeometry):
    _is_vector = True

    def __init__(self, shape, det_pos, det_v, det_u):
        super().__init__(shape)
        arrs = [tt.to_vec(x) for x in (det_pos, det_v, det_u)]
        try:
            arrs = np.broadcast_arrays(*arrs)
        except ValueError:
            raise ValueError(f"Not all arguments are the same shape. Got: {[a.shape for a in arrs]}")
        self._det_pos, self._det_v, self._det_u = (np.array(a) for a in arrs)

    def __eq__(self, other):
        if not isinstance(other, DetectorVectorGeometry):
            return False
        if self.det_shape != other.det_shape or self.num_angles != other.num_angles:
            return False
        return all(np.all(abs(a - b) < _EPS) for a, b in (
            (self._det_pos, other._det_pos), (self._det_v, other._det_v),
            (self._det_u, other._det_u)))

    def __getitem__(self, key):
        full = slice(None)
        if isinstance(key, (Integral, slice)):
            key = (key, full, full)
        if isinstance(key, tuple):
            key = key + (full,) * (3 - len(key))
            v0, v1, lv, sv = slice_interval(0, self.det_shape[0], self.det_shape[0], key[1])
            u0, u1, lu, su = slice_interval(0, self.det_shape[1], self.det_shape[1], key[2])
            llc = self.lower_left_corner
            new_llc = llc + v0 * self._det_v + u0 * self._det_u
            new_trc = llc + v1 * self._det_v + u1 * self._det_u
            centre = ((new_llc + new_trc) / 2)[key[0]]
            return det_vec((lv, lu), centre, self._det_v[key[0]] * sv, self._det_u[key[0]] * su)
        return det_vec(self.det_shape, self._det_pos[key], self._det_v[key], self._det_u[key])

    def to_vec(self):
        return self

    def to_vol(self):
        from .volume_vec import volume_vec

        rows, cols = self.det_shape
        return volume_vec(shape=(rows, 1, cols), pos=self._det_pos, w=self._det_v,
                          v=self.det_normal, u=self._det_u)

    @property
    def num_angles(self):
        return len(self._det_pos)

    @property
    def det_pos(self):
        return np.copy(self._det_pos)

    @property
    def det_v(self):
        return np.copy(self._det_v)

    @property
    def det_u(self):
        return np.copy(self._det_u)

    @property
    def det_normal(self):
        return vc.cross_product(self._det_u, self._det_v)

    @property
    def det_sizes(self):
        return np.stack([vc.norm(self._det_v * self.det_shape[0]),
                         vc.norm(self._det_u * self.det_shape[1])], axis=1)

    @property
    def det_size(self):
        S = self.det_sizes
        if np.all(abs(S.max(axis=0) - S.min(axis=0)) < _EPS):
            return (float(S[0, 0]), float(S[0, 1]))
        raise ValueError("The size of the detector is not constant. "
                         "To prevent this error, use `pg.det_sizes'. ")

    @property
    def corners(self):
        ds = self._det_pos
        uo = self._det_u * self.det_shape[1] / 2
        vo = self._det_v * self.det_shape[0] / 2
        return np.array([ds - uo - vo, ds - uo + vo, ds + uo - vo, ds + uo + vo]).transpose([1, 0, 2])

    @property
    def lower_left_corner(self):
        return (self._det_pos - self._det_v * self.det_shape[0] / 2
                - self._det_u * self.det_shape[1] / 2)

    def rescale_det(self, scale):
        sv, su = (int(s) for s in tt.to_size2d(scale))
        shape = (self.det_shape[0] // sv, self.det_shape[1] // su)
        return det_vec(shape, self._det_pos, self._det_v * sv, self._det_u * su)

    def reshape(self, new_shape):
        new_shape = tt.to_shape2d(new_shape)
        return det_vec(new_shape, self._det_pos,
                       self._det_v / new_shape[0] * self.det_shape[0],
                       self._det_u / new_shape[1] * self.det_shape[1])

    def _project_along(self, point, direction):
        """Detector coordinates ``(v, u)`` (pixels from the centre) of ``point``
        moved along ``direction`` into the detector plane."""
        if np.isscalar(point):
            point = tt.to_pos(point)
        origin = tt.to_vec(point)
        inter = vc.intersect(origin, direction, self._det_pos, self.det_normal)
        du = vc.dot(inter - self._det_pos, self._det_u) / vc.squared_norm(self._det_u)
        dv = vc.dot(inter - self._det_pos, self._det_v) / vc.squared_norm(self._det_v)
        return np.stack((dv, du), axis=-1)

    def __rmul__(self, other):
        from .transform import Transform

        if isinstance(other, Transform):
            return det_vec(self.det_shape, other.transform_point(self._det_pos),
                           other.transform_vec(self._det_v), other.transform_vec(self._det_u))
        return NotImplemented
