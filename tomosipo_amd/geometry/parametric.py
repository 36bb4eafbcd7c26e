"""Circular parallel-beam and cone-beam geometries and their vector form.

Construction follows tomosipo ``geometry/parallel.py:12-282`` and
``geometry/cone.py:14-399``:

* an integer ``angles`` means ``linspace(0, pi)`` (parallel) or
  ``linspace(0, 2 pi)`` (cone), both without the end point
  (``parallel.py:86-88``, ``cone.py:164-166``);
* ``cone_angle`` sets ``src_det_dist = src_orig_dist = det_height / cone_angle``,
  i.e. the detector passes through the origin (``cone.py:82-85``);
* pixel spacing is ``size / shape`` per axis (``parallel.py:164-168``,
  ``cone.py:255-260``).

``to_vec`` is the closed form of ASTRA's ``geom_2vec`` for these geometries
(reached from ``cone.py:251-252`` and ``parallel.py:188-195``). For angle
``t``, detector spacings ``su`` (columns) and ``sv`` (rows), in ``(z, y, x)``:

* parallel: ``ray = (0, -cos t, sin t)``, ``det = 0``;
* cone: ``src = (0, -cos t, sin t) * SOD``, ``det = (0, cos t, -sin t) * ODD``
  with ``ODD = SDD - SOD``;
* both: ``u = (0, sin t, cos t) * su``, ``v = (sv, 0, 0)``.

At ``t = 0`` this puts ``v`` along ``+z``, ``u`` along ``+x`` and the cone source
on ``-y``, the axis convention pinned by
``tests/geometry/test_cone_vec.py:143-173``.
"""
import warnings

import numpy as np

from .. import types as tt
from .det_vec import ProjectionGeometry
from .vec import ParallelVectorGeometry, ConeVectorGeometry

_EPS = 1e-8


def _angles(angles, full_turn, cls_name):
    original = angles
    if np.isscalar(angles) and isinstance(angles, (int, np.integer)) and not isinstance(angles, bool):
        angles = np.linspace(0, full_turn, int(angles), endpoint=False)
    else:
        angles = tt.to_scalars(angles, var_name="angles", accept_empty=True)
    if len(angles) == 0:
        raise (TypeError if cls_name == "ParallelGeometry" else ValueError)(
            f"{cls_name} expects non-empty array of angles; got {original}")
    return original, angles


def _det_axes(angles, size, shape):
    sv, su = np.array(size, dtype=np.float64) / np.array(shape, dtype=np.float64)
    c, s = np.cos(angles), np.sin(angles)
    z = np.zeros_like(angles)
    u = np.stack([z, s * su, c * su], axis=1)
    v = np.stack([z + sv, z, z], axis=1)
    return c, s, z, u, v


def parallel(*, angles=1, shape=1, size=None):
    return ParallelGeometry(angles, shape, size)


def random_parallel(rng=None):
    rng = np.random.default_rng() if rng is None else rng
    return parallel(angles=rng.normal(size=20), shape=rng.uniform(10, 20, size=2).astype(int),
                    size=rng.uniform(10, 20, size=2))


class ParallelGeometry(ProjectionGeometry):
    _is_parallel = True

    def __init__(self, angles=1, shape=1, size=None):
        super().__init__(shape)
        self._angles_original, self._angles = _angles(angles, np.pi, "ParallelGeometry")
        self._size = tt.to_size2d(self.det_shape if size is None else size)

    def __repr__(self):
        return (f"ts.parallel(\n    angles={self._angles_original!r},\n    shape={self.det_shape!r},\n"
                f"    size={self._size!r},\n)")

    def __eq__(self, other):
        return (isinstance(other, ParallelGeometry) and self.det_shape == other.det_shape
                and len(self._angles) == len(other._angles)
                and bool(np.all(abs(self._angles - other._angles) < _EPS))
                and bool(np.all(abs(np.subtract(self._size, other._size)) < _EPS)))

    def __getitem__(self, key):
        if isinstance(key, tuple):
            raise ValueError(
                f"Expected 1 index to ParallelGeometry, got {len(key)}. Indexing on the detector "
                "plane is not supported, since it might move the detector center. "
                "To prevent this error, use `pg.to_vec()[a:b, c:d, e:f]'. ")
        return parallel(angles=np.atleast_1d(self._angles[key]), shape=self.det_shape, size=self._size)

    def to_astra(self):
        rows, cols = self.det_shape
        sp = np.array(self._size) / np.array(self.det_shape)
        return {"type": "parallel3d", "DetectorSpacingX": sp[1], "DetectorSpacingY": sp[0],
                "DetectorRowCount": rows, "DetectorColCount": cols,
                "ProjectionAngles": np.copy(self._angles)}

    @staticmethod
    def from_astra(astra_pg):
        if astra_pg["type"] != "parallel3d":
            raise ValueError("ParallelGeometry.from_astra only supports 'parallel3d' type astra geometries.")
        shape = (astra_pg["DetectorRowCount"], astra_pg["DetectorColCount"])
        sp = (astra_pg["DetectorSpacingY"], astra_pg["DetectorSpacingX"])
        return parallel(angles=np.copy(astra_pg["ProjectionAngles"]), shape=shape,
                        size=np.array(sp) * np.array(shape))

    def to_vec(self):
        c, s, z, u, v = _det_axes(self._angles, self._size, self.det_shape)
        ray = np.stack([z, -c, s], axis=1)
        return ParallelVectorGeometry(self.det_shape, ray, np.zeros_like(ray), v, u)

    def to_vol(self):
        return self.to_vec().to_vol()

    @property
    def num_angles(self):
        return len(self._angles)

    @property
    def angles(self):
        return np.copy(self._angles)

    @property
    def src_pos(self):
        raise NotImplementedError()

    @property
    def ray_dir(self):
        return self.to_vec().ray_dir

    det_pos = property(lambda self: self.to_vec().det_pos)
    det_v = property(lambda self: self.to_vec().det_v)
    det_u = property(lambda self: self.to_vec().det_u)
    det_normal = property(lambda self: self.to_vec().det_normal)
    corners = property(lambda self: self.to_vec().corners)
    lower_left_corner = property(lambda self: self.to_vec().lower_left_corner)

    @property
    def det_size(self):
        return self._size

    @property
    def det_sizes(self):
        return np.repeat([self._size], self.num_angles, axis=0)

    def rescale_det(self, scale):
        sv, su = (int(x) for x in tt.to_size2d(scale))
        return parallel(angles=np.copy(self._angles),
                        shape=(self.det_shape[0] // sv, self.det_shape[1] // su), size=self._size)

    def reshape(self, new_shape):
        return parallel(angles=np.copy(self._angles), shape=new_shape, size=self._size)

    def project_point(self, point):
        return self.to_vec().project_point(point)

    def __rmul__(self, other):
        from .transform import Transform

        if isinstance(other, Transform):
            warnings.warn("Converting parallel geometry to vector geometry. "
                          "Use `T * pg.to_vec()` to inhibit this warning. ", stacklevel=2)
            return other * self.to_vec()
        return NotImplemented


def cone(*, angles=1, shape=(1, 1), size=None, cone_angle=None, src_orig_dist=None,
         src_det_dist=None):
    shape = tt.to_shape2d(shape)
    size = tt.to_size2d(shape if size is None else size)
    if cone_angle is None and src_orig_dist is None and src_det_dist is None:
        raise ValueError("ts.cone requires at least one of `cone_angle`, `src_orig_dist`, "
                         "or `src_det_dist` parameters. ")
    if cone_angle is not None and src_orig_dist is not None:
        raise ValueError("ts.cone does not accept both `cone_angle` and src_orig_dist` arguments at the same time. ")
    if cone_angle is not None and src_det_dist is not None:
        raise ValueError("ts.cone does not accept both `cone_angle` and src_det_dist` arguments at the same time. ")
    if cone_angle is not None:
        src_det_dist = size[0] / cone_angle
        src_orig_dist = src_det_dist
    elif src_orig_dist is None:
        src_orig_dist = src_det_dist
    elif src_det_dist is None:
        src_det_dist = src_orig_dist
    return ConeGeometry(angles=angles, shape=shape, size=size, src_orig_dist=src_orig_dist,
                        src_det_dist=src_det_dist)


def random_cone(rng=None):
    rng = np.random.default_rng() if rng is None else rng
    return cone(angles=rng.normal(size=20), shape=rng.uniform(10, 20, size=2).astype(int),
                size=rng.uniform(10, 20, size=2), src_orig_dist=rng.uniform(0, 10),
                src_det_dist=rng.uniform(0, 20))


class ConeGeometry(ProjectionGeometry):
    _is_cone = True

    def __init__(self, angles=1, shape=1, size=None, src_orig_dist=None, src_det_dist=None):
        super().__init__(shape)
        self.angles_original, self._angles = _angles(angles, 2 * np.pi, "ConeGeometry")
        self._size = tt.to_size2d(self.det_shape if size is None else size)
        if src_orig_dist is None:
            raise ValueError("Expected `src_orig_dist` parameter. Got `None`. ")
        if src_det_dist is None:
            raise ValueError("Expected `src_det_dist` parameter. Got `None`. ")
        self._src_orig_dist = float(src_orig_dist)
        self._src_det_dist = float(src_det_dist)

    def __repr__(self):
        return (f"ts.cone(\n    angles={self.angles_original!r},\n    shape={self.det_shape},\n"
                f"    size={self._size},\n    src_orig_dist={self._src_orig_dist},\n"
                f"    src_det_dist={self._src_det_dist},\n)")

    def __eq__(self, other):
        return (isinstance(other, ConeGeometry) and len(self._angles) == len(other._angles)
                and self.det_shape == other.det_shape
                and bool(np.all(abs(self._angles - other._angles) < _EPS))
                and bool(np.all(abs(np.subtract(self._size, other._size)) < _EPS))
                and abs(self._src_orig_dist - other._src_orig_dist) < _EPS
                and abs(self._src_det_dist - other._src_det_dist) < _EPS)

    def __getitem__(self, key):
        if isinstance(key, tuple):
            raise ValueError(
                f"Expected 1 index to ConeGeometry, got {len(key)}. Indexing on the detector "
                "plane is not supported, since it might move the detector center. ")
        return ConeGeometry(angles=np.atleast_1d(self._angles[key]), shape=self.det_shape,
                            size=self._size, src_orig_dist=self._src_orig_dist,
                            src_det_dist=self._src_det_dist)

    def to_astra(self):
        rows, cols = self.det_shape
        sp = np.array(self._size) / np.array(self.det_shape)
        return {"type": "cone", "DetectorSpacingX": sp[1], "DetectorSpacingY": sp[0],
                "DetectorRowCount": rows, "DetectorColCount": cols,
                "ProjectionAngles": np.copy(self._angles),
                "DistanceOriginSource": self._src_orig_dist,
                "DistanceOriginDetector": self._src_det_dist - self._src_orig_dist}

    @staticmethod
    def from_astra(astra_pg):
        if astra_pg["type"] != "cone":
            raise ValueError("ConeGeometry.from_astra only supports 'cone' type astra geometries.")
        shape = (astra_pg["DetectorRowCount"], astra_pg["DetectorColCount"])
        sp = (astra_pg["DetectorSpacingY"], astra_pg["DetectorSpacingX"])
        sod = astra_pg["DistanceOriginSource"]
        return ConeGeometry(angles=astra_pg["ProjectionAngles"], shape=shape,
                            size=np.array(sp) * np.array(shape), src_orig_dist=sod,
                            src_det_dist=sod + astra_pg["DistanceOriginDetector"])

    def to_vec(self):
        c, s, z, u, v = _det_axes(self._angles, self._size, self.det_shape)
        sod = self._src_orig_dist
        odd = self._src_det_dist - self._src_orig_dist
        src = np.stack([z, -c * sod, s * sod], axis=1)
        det = np.stack([z, c * odd, -s * odd], axis=1)
        return ConeVectorGeometry(shape=self.det_shape, src_pos=src, det_pos=det, det_v=v, det_u=u)

    def to_vol(self):
        return self.to_vec().to_vol()

    @property
    def src_orig_dist(self):
        return self._src_orig_dist

    @property
    def src_det_dist(self):
        return self._src_det_dist

    @property
    def num_angles(self):
        return len(self._angles)

    @property
    def angles(self):
        return np.copy(self._angles)

    @property
    def ray_dir(self):
        raise NotImplementedError()

    src_pos = property(lambda self: self.to_vec().src_pos)
    det_pos = property(lambda self: self.to_vec().det_pos)
    det_v = property(lambda self: self.to_vec().det_v)
    det_u = property(lambda self: self.to_vec().det_u)
    det_normal = property(lambda self: self.to_vec().det_normal)
    corners = property(lambda self: self.to_vec().corners)
    lower_left_corner = property(lambda self: self.to_vec().lower_left_corner)

    @property
    def det_size(self):
        return self._size

    @property
    def det_sizes(self):
        return np.repeat([self._size], self.num_angles, axis=0)

    def rescale_det(self, scale):
        sv, su = (int(x) for x in tt.to_size2d(scale))
        return ConeGeometry(self.angles_original, (self.det_shape[0] // sv, self.det_shape[1] // su),
                            self._size, self._src_orig_dist, self._src_det_dist)

    def reshape(self, new_shape):
        return ConeGeometry(self.angles_original, tt.to_shape2d(new_shape), self._size,
                            self._src_orig_dist, self._src_det_dist)

    def project_point(self, point):
        return self.to_vec().project_point(point)

    def __rmul__(self, other):
        from .transform import Transform

        if isinstance(other, Transform):
            warnings.warn("Converting cone geometry to vector geometry. "
                          "Use `T * pg.to_vec()` to inhibit this warning. ", stacklevel=2)
            return other * self.to_vec()
        return NotImplemented
