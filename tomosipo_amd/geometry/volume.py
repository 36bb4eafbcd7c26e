"""Axis-aligned volume geometry.

Follows tomosipo ``geometry/volume.py``: construction from shape / pos / size
/ extent (``volume.py:34-122``), properties and slicing (``:161-453``),
translate / scale / multiply / reshape (``:455-537``), transforms that stay
axis-aligned (``:539-566``), and the ASTRA volume dictionary
(``to_astra`` ``:286-301``, ``from_astra`` ``:569-580``). The dictionary uses
ASTRA's keys (``GridColCount`` = x, ``GridRowCount`` = y,
``GridSliceCount`` = z, ``option.WindowMin/Max{X,Y,Z}``) because that is the
data format that crosses the projector boundary.
"""
import warnings

import numpy as np

from .. import types as tt
from .volume_vec import VolumeVectorGeometry, volume_vec


def is_volume(g):
    return isinstance(g, (VolumeGeometry, VolumeVectorGeometry))


def volume(*, shape=(1, 1, 1), pos=None, size=None, extent=None):
    shape = tt.to_shape3d(shape)
    if extent is not None and pos is not None:
        raise ValueError("ts.volume does not accept both `extent` and `pos` arguments. ")
    if extent is not None and size is not None:
        raise ValueError("ts.volume does not accept both `extent` and `size` arguments. ")
    if extent is not None:
        pos, size = _extent_to_pos_size(extent)
        return VolumeGeometry(shape, pos, size)
    if size is None:
        size = shape
    if pos is None:
        pos = 0
    return VolumeGeometry(shape, pos=pos, size=size)


def random_volume(rng=None):
    rng = np.random.default_rng() if rng is None else rng
    return volume(shape=rng.uniform(2, 100, 3).astype(int), pos=rng.normal(size=3),
                  size=rng.uniform(1, 10, size=3))


def _pos_size_to_extent(pos, size):
    pos = np.array(tt.to_pos(pos))
    size = np.array(tt.to_size3d(size))
    return tuple((float(l), float(r)) for l, r in zip(pos - 0.5 * size, pos + 0.5 * size))


def _extent_to_pos_size(extent):
    extent = tuple(extent)
    if len(extent) == 2 and np.isscalar(extent[0]):
        extent = (extent,) * 3
    size = tt.to_size3d(tuple(r - l for l, r in extent))
    pos = tt.to_pos(tuple((r + l) / 2 for l, r in extent))
    return pos, size


class VolumeGeometry:
    def __init__(self, shape=(1, 1, 1), pos=0, size=None):
        shape = tt.to_shape3d(shape)
        pos = tt.to_pos(pos)
        if size is None:
            self._inner = volume_vec(shape=shape, pos=pos)
        else:
            size = tt.to_size3d(size)
            vs = tuple(s / n for s, n in zip(size, shape))
            self._inner = volume_vec(shape=shape, pos=pos, w=(vs[0], 0, 0), v=(0, vs[1], 0),
                                     u=(0, 0, vs[2]))

    def __repr__(self):
        return (f"ts.volume(\n    shape={self.shape},\n    pos={tuple(float(p) for p in self.pos[0])},\n"
                f"    size={self.size},\n)")

    def __eq__(self, other):
        return isinstance(other, VolumeGeometry) and self._inner == other._inner

    def __getitem__(self, key):
        inner = self._inner[(0, *key)] if isinstance(key, tuple) else self._inner[0, key]
        return VolumeGeometry(inner.shape, inner.pos[0], inner.size)

    def __contains__(self, other):
        return all(s[0] <= o[0] and o[1] <= s[1] for s, o in zip(self.extent, other.extent))

    def __len__(self):
        return 1

    def to_astra(self):
        nz, ny, nx = self.shape
        (z0, z1), (y0, y1), (x0, x1) = self.extent
        return {
            "GridRowCount": ny,
            "GridColCount": nx,
            "GridSliceCount": nz,
            "option": {
                "WindowMinX": x0, "WindowMaxX": x1,
                "WindowMinY": y0, "WindowMaxY": y1,
                "WindowMinZ": z0, "WindowMaxZ": z1,
            },
        }

    def to_vec(self):
        return self._inner

    @property
    def num_steps(self):
        return 1

    @property
    def pos(self):
        return self._inner.pos

    @property
    def w(self):
        return self._inner.w

    @property
    def v(self):
        return self._inner.v

    @property
    def u(self):
        return self._inner.u

    @property
    def shape(self):
        return self._inner.shape

    @property
    def sizes(self):
        return self._inner.sizes

    @property
    def size(self):
        return self._inner.size

    @property
    def voxel_sizes(self):
        return self._inner.voxel_sizes

    @property
    def voxel_size(self):
        return self._inner.voxel_size

    @property
    def extent(self):
        return _pos_size_to_extent(self._inner.pos[0], self._inner.size)

    @property
    def corners(self):
        return self._inner.corners

    @property
    def lower_left_corner(self):
        return self._inner.lower_left_corner

    def with_voxel_size(self, voxel_size):
        voxel_size = tt.to_size3d(voxel_size)
        new_shape = (np.array(self.size) / voxel_size).astype(int)
        return VolumeGeometry(tuple(int(s) for s in new_shape), pos=self.pos[0],
                              size=new_shape * np.array(voxel_size))

    def reshape(self, new_shape):
        return VolumeGeometry(new_shape, pos=self.pos[0], size=self.size)

    def translate(self, t):
        t = tt.to_pos(t)
        return VolumeGeometry(self.shape, tuple(p + d for p, d in zip(self.pos[0], t)), self.size)

    def untranslate(self, t):
        return self.translate(-np.array(t))

    def scale(self, s):
        s = tt.to_size3d(s)
        return VolumeGeometry(self.shape, self.pos[0], tuple(a * b for a, b in zip(s, self.size)))

    def multiply(self, s):
        s = tt.to_size3d(s)
        return VolumeGeometry(self.shape, tuple(a * b for a, b in zip(s, self.pos[0])),
                              tuple(a * b for a, b in zip(s, self.size)))

    def __rmul__(self, other):
        from .transform import Transform, translate, scale

        if isinstance(other, Transform):
            if other.num_steps == 1:
                t = other.matrix[0, :3, 3]
                s = abs(other.matrix[0].diagonal()[:3])
                if np.all(s > 0) and translate(t) * scale(s) == other:
                    return self.multiply(s).translate(t)
            warnings.warn("Converting VolumeGeometry to VolumeVectorGeometry. "
                          "Use `T * vg.to_vec()` to inhibit this warning. ", stacklevel=2)
            return other * self.to_vec()
        return NotImplemented


def from_astra(avg):
    o = avg["option"]
    shape = (avg["GridSliceCount"], avg["GridRowCount"], avg["GridColCount"])
    extent = ((o["WindowMinZ"], o["WindowMaxZ"]), (o["WindowMinY"], o["WindowMaxY"]),
              (o["WindowMinX"], o["WindowMaxX"]))
    pos, size = _extent_to_pos_size(extent)
    return VolumeGeometry(shape=shape, pos=pos, size=size)
