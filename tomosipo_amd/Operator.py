"""Tomographic projection operator (tomosipo ``Operator.py``).

``ts.operator(vg, pg)`` returns an ``Operator`` whose call is the forward
projection and whose ``.T`` is the backprojection, with the argument meaning,
output allocation and error behaviour of tomosipo:

* a ``VolumeVectorGeometry`` is moved to an axis-aligned frame on the origin
  and the same per-step transform is applied to the projection geometry
  (``to_astra_compatible_operator_geometry``, ``Operator.py:11-60``);
* ``out=None`` allocates ``empty`` (or ``zeros`` when ``additive``) on the
  input's array type and device (``Operator.py:176-233``);
* ``Data`` in gives ``Data`` out (``Operator.py:189-192, 230-233``);
* ``A.T is A.T.T.T`` (``Operator.py:288-372``).

The backprojection is the exact transpose of the forward projection.
"""
import numpy as np

from . import vector_calc as vc
from .Data import Data, data as make_data
from .geometry import (VolumeGeometry, VolumeVectorGeometry, volume, from_perspective)
from .links.base import link, geometry_shape
from .projector import create_projector, direct_fp, direct_bp


def to_astra_compatible_operator_geometry(vg, pg):
    if isinstance(vg, VolumeGeometry):
        return (vg, pg)
    if not isinstance(vg, VolumeVectorGeometry):
        raise TypeError(f"Expected volume geometry. Got {type(vg)}. ")
    vg = vg.to_vec()
    P = from_perspective(pos=vg.pos, w=vg.w / vc.norm(vg.w)[:, None],
                         v=vg.v / vc.norm(vg.v)[:, None], u=vg.u / vc.norm(vg.u)[:, None])
    vg = P * vg
    pg = P * pg.to_vec()
    voxel_size = vg.voxel_size
    assert np.allclose(vg.pos, np.array([0, 0, 0]))
    assert np.allclose(vg.w, voxel_size[0] * np.array([1, 0, 0]))
    assert np.allclose(vg.v, voxel_size[1] * np.array([0, 1, 0]))
    assert np.allclose(vg.u, voxel_size[2] * np.array([0, 0, 1]))
    return volume(shape=vg.shape, pos=0, size=vg.size), pg


def operator(volume_geometry, projection_geometry, voxel_supersampling=1,
             detector_supersampling=1, additive=False):
    return Operator(volume_geometry, projection_geometry, voxel_supersampling=voxel_supersampling,
                    detector_supersampling=detector_supersampling, additive=additive)


def _to_link(geometry, x):
    return x.link if isinstance(x, Data) else link(geometry, x)


class Operator:
    def __init__(self, volume_geometry, projection_geometry, voxel_supersampling=1,
                 detector_supersampling=1, additive=False):
        self.volume_geometry = volume_geometry
        self.projection_geometry = projection_geometry
        vg, pg = to_astra_compatible_operator_geometry(volume_geometry, projection_geometry)
        self.astra_compat_vg = vg
        self.astra_compat_pg = pg
        self.projector = create_projector(vg, pg, voxel_supersampling=voxel_supersampling,
                                          detector_supersampling=detector_supersampling)
        self.additive = additive
        self._transpose = BackprojectionOperator(self)

    # tomosipo keeps the ASTRA projector id under this name (Operator.py:165-171)
    @property
    def astra_projector(self):
        return self.projector

    def _fp(self, volume, out=None):
        vlink = _to_link(self.astra_compat_vg, volume)
        if out is not None:
            plink = _to_link(self.astra_compat_pg, out)
        else:
            plink = vlink.new_zeros(self.range_shape) if self.additive else vlink.new_empty(self.range_shape)
        direct_fp(self.projector, vlink, plink, additive=self.additive)
        if isinstance(volume, Data):
            return make_data(self.projection_geometry, plink.data)
        return plink.data

    def _bp(self, projection, out=None):
        plink = _to_link(self.astra_compat_pg, projection)
        if out is not None:
            vlink = _to_link(self.astra_compat_vg, out)
        else:
            vlink = plink.new_zeros(self.domain_shape) if self.additive else plink.new_empty(self.domain_shape)
        direct_bp(self.projector, vlink, plink, additive=self.additive)
        if isinstance(projection, Data):
            return make_data(self.volume_geometry, vlink.data)
        return vlink.data

    def __call__(self, volume, out=None):
        return self._fp(volume, out)

    def transpose(self):
        return self._transpose

    @property
    def T(self):
        return self.transpose()

    @property
    def domain(self):
        return self.volume_geometry

    @property
    def range(self):
        return self.projection_geometry

    @property
    def domain_shape(self):
        return geometry_shape(self.astra_compat_vg)

    @property
    def range_shape(self):
        return geometry_shape(self.astra_compat_pg)


class BackprojectionOperator:
    """Transpose of an operator; ``A.T.T is A``."""

    def __init__(self, parent):
        self.parent = parent

    def __call__(self, projection, out=None):
        return self.parent._bp(projection, out)

    def transpose(self):
        return self.parent

    @property
    def T(self):
        return self.transpose()

    @property
    def additive(self):
        return self.parent.additive

    @property
    def domain(self):
        return self.parent.range

    @property
    def range(self):
        return self.parent.domain

    @property
    def domain_shape(self):
        return self.parent.range_shape

    @property
    def range_shape(self):
        return self.parent.domain_shape
