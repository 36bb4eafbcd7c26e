"""Arbitrarily oriented (possibly moving) volume geometry.

Follows tomosipo ``geometry/volume_vec.py``: ``pos`` is the volume centre and
``w, v, u`` are the per-voxel axis vectors, all ``(num_steps, 3)`` in
``(z, y, x)`` order (``volume_vec.py:77-160``); slicing (``:175-231``),
``sizes``/``size``/``voxel_size`` (``:279-345``), corners and lower-left
corner (``:347-399``), ``reshape`` and transforms (``:404-441``).
"""
from numbers import Integral

import numpy as np

from .. import types as tt
from .. import vector_calc as vc
from ..utils import slice_interval

_EPS = 1e-8


def volume_vec(*, shape, pos=0, w=(1, 0, 0), v=(0, 1, 0), u=(0, 0, 1)):
    return VolumeVectorGeometry(shape, pos, w, v, u)


def random_volume_vec(rng=None):
    from .transform import random_transform

    rng = np.random.default_rng() if rng is None else rng
    vg = volume_vec(shape=rng.uniform(4, 10, size=3).astype(int), pos=rng.normal(size=3))
    return random_transform(rng) * vg


class VolumeVectorGeometry:
    def __init__(self, shape, pos, w=(1, 0, 0), v=(0, 1, 0), u=(0, 0, 1)):
        self._shape = tt.to_shape3d(shape)
        if np.isscalar(pos) and pos == 0.0:
            pos = (0.0, 0.0, 0.0)
        arrs = [tt.to_vec(x, name) for x, name in
                ((pos, "position"), (w, "w axis"), (v, "v axis"), (u, "u axis"))]
        try:
            arrs = np.broadcast_arrays(*arrs)
        except ValueError:
            raise ValueError(
                f"Not all arguments pos, w, v, u are the same shape. Got: {[a.shape for a in arrs]}"
            )
        self._pos, self._w, self._v, self._u = (np.array(a) for a in arrs)

    def __repr__(self):
        return (
            f"ts.volume_vec(\n    shape={self._shape},\n    pos={self.pos!r},\n"
            f"    w={self.w!r},\n    v={self.v!r},\n    u={self.u!r},\n)"
        )

    def __eq__(self, other):
        if not isinstance(other, VolumeVectorGeometry):
            return False
        if self.shape != other.shape or self.num_steps != other.num_steps:
            return False
        return all(
            np.all(abs(a - b) < _EPS)
            for a, b in ((self._pos, other._pos), (self._w, other._w),
                         (self._v, other._v), (self._u, other._u))
        )

    def __getitem__(self, key):
        full = slice(None)
        if isinstance(key, tuple) and len(key) > 4:
            raise ValueError(f"VolumeVectorGeometry supports indexing in 4 dimensions. Got {key}.")
        if isinstance(key, (Integral, slice)):
            key = (key, full, full, full)
        if isinstance(key, tuple):
            key = key + (full,) * (4 - len(key))
            w0, w1, lw, sw = slice_interval(0, self.shape[0], self.shape[0], key[1])
            v0, v1, lv, sv = slice_interval(0, self.shape[1], self.shape[1], key[2])
            u0, u1, lu, su = slice_interval(0, self.shape[2], self.shape[2], key[3])
            llc = self.lower_left_corner
            new_llc = llc + w0 * self._w + v0 * self._v + u0 * self._u
            new_trc = llc + w1 * self._w + v1 * self._v + u1 * self._u
            new_pos = ((new_llc + new_trc) / 2)[key[0]]
            return VolumeVectorGeometry(
                (lw, lv, lu), new_pos, sw * self._w[key[0]], sv * self._v[key[0]],
                su * self._u[key[0]],
            )
        return VolumeVectorGeometry(self.shape, self._pos[key], self._w[key], self._v[key], self._u[key])

    def __len__(self):
        return self.num_steps

    def to_vec(self):
        return self

    @property
    def num_steps(self):
        return len(self._pos)

    @property
    def pos(self):
        return np.copy(self._pos)

    @property
    def w(self):
        return np.copy(self._w)

    @property
    def v(self):
        return np.copy(self._v)

    @property
    def u(self):
        return np.copy(self._u)

    @property
    def shape(self):
        return self._shape

    @property
    def sizes(self):
        return np.stack(
            [self._shape[0] * vc.norm(self._w), self._shape[1] * vc.norm(self._v),
             self._shape[2] * vc.norm(self._u)], axis=1)

    @property
    def size(self):
        S = self.sizes
        if np.all(abs(S.max(axis=0) - S.min(axis=0)) < _EPS):
            return tuple(float(s) for s in S[0])
        raise ValueError("The size of the volume is not constant. To prevent this error, use `vg.sizes'.")

    @property
    def voxel_sizes(self):
        return self.sizes / np.array([self.shape])

    @property
    def voxel_size(self):
        return tuple(s / n for s, n in zip(self.size, self.shape))

    @property
    def corners(self):
        c = np.array([(a, b, d) for a in (0, 1) for b in (0, 1) for d in (0, 1)], dtype=float) - 0.5
        sh = self._shape
        out = (self._pos[:, None, :]
               + c[None, :, 0:1] * self._w[:, None, :] * sh[0]
               + c[None, :, 1:2] * self._v[:, None, :] * sh[1]
               + c[None, :, 2:3] * self._u[:, None, :] * sh[2])
        return out

    @property
    def lower_left_corner(self):
        return (self._pos - self._w * self._shape[0] / 2 - self._v * self._shape[1] / 2
                - self._u * self._shape[2] / 2)

    def reshape(self, new_shape):
        new_shape = tt.to_shape3d(new_shape)
        S = self.sizes
        return VolumeVectorGeometry(
            new_shape, self._pos,
            self._w / vc.norm(self._w)[:, None] * (S[:, 0] / new_shape[0])[:, None],
            self._v / vc.norm(self._v)[:, None] * (S[:, 1] / new_shape[1])[:, None],
            self._u / vc.norm(self._u)[:, None] * (S[:, 2] / new_shape[2])[:, None],
        )

    def __rmul__(self, other):
        from .transform import Transform

        if isinstance(other, Transform):
            return VolumeVectorGeometry(
                self.shape, other.transform_point(self._pos), other.transform_vec(self._w),
                other.transform_vec(self._v), other.transform_vec(self._u))
        return NotImplemented
