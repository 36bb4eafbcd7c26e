"""Vector projection geometries: ``parallel_vec`` and ``cone_vec``.

Follows tomosipo ``geometry/parallel_vec.py:15-343`` and
``geometry/cone_vec.py:10-339``: one ray direction (parallel) or source
position (cone) per angle plus the detector vectors, all ``(n, 3)`` in
``(z, y, x)`` order.

``to_astra`` produces ASTRA's ``parallel3d_vec`` / ``cone_vec`` dictionary
(``parallel_vec.py:175-192``, ``cone_vec.py:174-191``): ``Vectors`` is
``(n, 12)`` with rows ``[ray|src, det, u, v]``, each 3-vector in ``(x, y, z)``
order. This dictionary is what the projector boundary receives.
"""
import numpy as np

from .. import types as tt
from .. import vector_calc as vc
from .det_vec import ProjectionGeometry, DetectorVectorGeometry, det_vec

_EPS = 1e-8


def _vectors(first, dv):
    return np.concatenate([first[:, ::-1], dv._det_pos[:, ::-1], dv._det_u[:, ::-1],
                           dv._det_v[:, ::-1]], axis=1)


def _unpack(astra_pg, kind):
    if astra_pg["type"] != kind:
        raise ValueError(f"Expected an ASTRA geometry of type '{kind}'. Got '{astra_pg['type']}'.")
    V = np.asarray(astra_pg["Vectors"], dtype=np.float64)
    shape = (astra_pg["DetectorRowCount"], astra_pg["DetectorColCount"])
    return shape, V[:, 0:3][:, ::-1], V[:, 3:6][:, ::-1], V[:, 9:12][:, ::-1], V[:, 6:9][:, ::-1]


class _VecGeometry(ProjectionGeometry):
    _is_vector = True

    def _init_det(self, shape, first, det_pos, det_v, det_u, first_name):
        arrs = [tt.to_vec(first, first_name), tt.to_vec(det_pos, "detector position"),
                tt.to_vec(det_v, "v axis"), tt.to_vec(det_u, "u axis")]
        try:
            arrs = np.broadcast_arrays(*arrs)
        except ValueError:
            raise ValueError(
                f"Not all arguments {first_name}, det_pos, det_v, det_u are the same shape. "
                f"Got: {[a.shape for a in arrs]}")
        self._first = np.array(arrs[0])
        self._det_vec = det_vec(shape, *arrs[1:])

    @property
    def num_angles(self):
        return self._det_vec.num_angles

    @property
    def angles(self):
        raise NotImplementedError()

    @property
    def det_pos(self):
        return self._det_vec.det_pos

    @property
    def det_v(self):
        return self._det_vec.det_v

    @property
    def det_u(self):
        return self._det_vec.det_u

    @property
    def det_normal(self):
        return self._det_vec.det_normal

    @property
    def det_size(self):
        return self._det_vec.det_size

    @property
    def det_sizes(self):
        return self._det_vec.det_sizes

    @property
    def corners(self):
        return self._det_vec.corners

    @property
    def lower_left_corner(self):
        return self._det_vec.lower_left_corner

    def to_vec(self):
        return self

    def to_vol(self):
        return self._det_vec.to_vol()

    def _rebuild(self, first, dv):
        return type(self)._make(dv.det_shape, first, dv.det_pos, dv.det_v, dv.det_u)

    def __getitem__(self, key):
        dv = self._det_vec[key]
        k0 = key[0] if isinstance(key, tuple) else key
        return self._rebuild(self._first[k0], dv)

    def rescale_det(self, scale):
        return self._rebuild(self._first, self._det_vec.rescale_det(scale))

    def reshape(self, new_shape):
        return self._rebuild(self._first, self._det_vec.reshape(new_shape))


def parallel_vec(*, shape, ray_dir, det_pos, det_v, det_u):
    return ParallelVectorGeometry(shape, ray_dir, det_pos, det_v, det_u)


def random_parallel_vec(rng=None):
    rng = np.random.default_rng() if rng is None else rng
    n = int(rng.uniform(1, 100))
    return parallel_vec(shape=rng.uniform(10, 20, size=2).astype(int),
                        ray_dir=rng.normal(size=(n, 3)), det_pos=rng.normal(size=(n, 3)),
                        det_v=rng.normal(size=(n, 3)), det_u=rng.normal(size=(n, 3)))


class ParallelVectorGeometry(_VecGeometry):
    _is_parallel = True

    def __init__(self, shape, ray_dir, det_pos, det_v, det_u):
        super().__init__(shape)
        self._init_det(shape, ray_dir, det_pos, det_v, det_u, "ray direction")

    @classmethod
    def _make(cls, shape, first, det_pos, det_v, det_u):
        return cls(shape, first, det_pos, det_v, det_u)

    def __repr__(self):
        return (f"ts.parallel_vec(\n    shape={self.det_shape},\n    ray_dir={self._first!r},\n"
                f"    det_pos={self.det_pos!r},\n    det_v={self.det_v!r},\n    det_u={self.det_u!r},\n)")

    def __eq__(self, other):
        return (isinstance(other, ParallelVectorGeometry) and self._det_vec == other._det_vec
                and bool(np.all(abs(self._first - other._first) < _EPS)))

    @property
    def ray_dir(self):
        return np.copy(self._first)

    @property
    def src_pos(self):
        raise NotImplementedError()

    def to_astra(self):
        rows, cols = self.det_shape
        return {"type": "parallel3d_vec", "DetectorRowCount": rows, "DetectorColCount": cols,
                "Vectors": _vectors(self._first, self._det_vec)}

    @staticmethod
    def from_astra(astra_pg):
        shape, ray, dpos, dv, du = _unpack(astra_pg, "parallel3d_vec")
        return ParallelVectorGeometry(shape, ray, dpos, dv, du)

    def project_point(self, point):
        return self._det_vec._project_along(point, self._first)

    def __rmul__(self, other):
        from .transform import Transform

        if isinstance(other, Transform):
            dv = other * self._det_vec
            return self._rebuild(other.transform_vec(self._first), dv)
        return NotImplemented


def cone_vec(*, shape, src_pos, det_pos, det_v, det_u):
    return ConeVectorGeometry(shape=shape, src_pos=src_pos, det_pos=det_pos, det_v=det_v,
                              det_u=det_u)


def random_cone_vec(rng=None):
    from .transform import random_transform
    from .parametric import random_cone

    rng = np.random.default_rng() if rng is None else rng
    return random_transform(rng) * random_cone(rng).to_vec()


class ConeVectorGeometry(_VecGeometry):
    _is_cone = True

    def __init__(self, *, shape, src_pos, det_pos, det_v, det_u):
        super().__init__(shape)
        self._init_det(shape, src_pos, det_pos, det_v, det_u, "source position")

    @classmethod
    def _make(cls, shape, first, det_pos, det_v, det_u):
        return cls(shape=shape, src_pos=first, det_pos=det_pos, det_v=det_v, det_u=det_u)

    def __repr__(self):
        return (f"ts.cone_vec(\n    shape={self.det_shape},\n    src_pos={self._first!r},\n"
                f"    det_pos={self.det_pos!r},\n    det_v={self.det_v!r},\n    det_u={self.det_u!r},\n)")

    def __eq__(self, other):
        return (isinstance(other, ConeVectorGeometry) and self._det_vec == other._det_vec
                and bool(np.all(abs(self._first - other._first) < _EPS)))

    @property
    def src_pos(self):
        return np.copy(self._first)

    @property
    def ray_dir(self):
        raise NotImplementedError()

    def to_astra(self):
        rows, cols = self.det_shape
        return {"type": "cone_vec", "DetectorRowCount": rows, "DetectorColCount": cols,
                "Vectors": _vectors(self._first, self._det_vec)}

    @staticmethod
    def from_astra(astra_pg):
        shape, src, dpos, dv, du = _unpack(astra_pg, "cone_vec")
        return ConeVectorGeometry(shape=shape, src_pos=src, det_pos=dpos, det_v=dv, det_u=du)

    def project_point(self, point):
        if np.isscalar(point):
            point = tt.to_pos(point)
        return self._det_vec._project_along(point, tt.to_vec(point) - self._first)

    def __rmul__(self, other):
        from .transform import Transform

        if isinstance(other, Transform):
            dv = other * self._det_vec
            return self._rebuild(other.transform_point(self._first), dv)
        return NotImplemented
