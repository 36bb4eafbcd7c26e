"""Geometry layer: volumes, projection geometries and transforms (FP64, host)."""
from .transform import (Transform, identity, translate, scale, rotate, reflect, to_perspective,
                        from_perspective, random_transform)
from .volume_vec import VolumeVectorGeometry, volume_vec, random_volume_vec
from .volume import VolumeGeometry, volume, random_volume, is_volume
from .det_vec import (ProjectionGeometry, DetectorVectorGeometry, det_vec, is_projection, is_cone,
                      is_parallel)
from .vec import (ParallelVectorGeometry, ConeVectorGeometry, parallel_vec, cone_vec,
                  random_parallel_vec, random_cone_vec)
from .parametric import (ParallelGeometry, ConeGeometry, parallel, cone, random_parallel,
                         random_cone)
from .concatenate import concatenate
from .volume import from_astra as _volume_from_astra


def from_astra_projection_geometry(astra_pg):
    kinds = {"parallel3d": ParallelGeometry.from_astra, "cone": ConeGeometry.from_astra,
             "parallel3d_vec": ParallelVectorGeometry.from_astra,
             "cone_vec": ConeVectorGeometry.from_astra}
    try:
        return kinds[astra_pg["type"]](astra_pg)
    except KeyError:
        raise ValueError(f"ASTRA projection geometry type '{astra_pg.get('type')}' is not supported.")


def from_astra(astra_geom):
    """Volume or projection geometry from its ASTRA dictionary (tomosipo ``astra.py:35-60``)."""
    if not isinstance(astra_geom, dict):
        raise TypeError(f"Expected an ASTRA geometry dictionary. Got {type(astra_geom)}.")
    if "GridSliceCount" in astra_geom:
        return _volume_from_astra(astra_geom)
    return from_astra_projection_geometry(astra_geom)


def to_astra(x):
    try:
        return x.to_astra()
    except AttributeError:
        raise TypeError(f"The object of type {type(x)} does not support conversion to ASTRA.")
