"""Concatenation of geometries and transforms along the angle / step axis.

Follows tomosipo ``geometry/concatenate.py:16-88``: parallel and cone
geometries are converted to vector form; volume geometries are concatenated
as volume vector geometries of equal shape.
"""
import numpy as np

from .transform import Transform
from .det_vec import is_parallel, is_cone
from .volume import is_volume
from .vec import parallel_vec, cone_vec
from .volume_vec import volume_vec


def concatenate(items):
    items = list(items)
    if len(items) == 0:
        raise ValueError("ts.concatenate expected at least one argument. ")
    if all(isinstance(i, Transform) for i in items):
        return Transform(np.concatenate([i.matrix for i in items]))
    if all(is_parallel(i) for i in items) or all(is_cone(i) for i in items):
        if not all(i.det_shape == items[0].det_shape for i in items):
            raise ValueError("Cannot concatenate geometries. Not all detector shapes are equal.")
        vecs = [i.to_vec() for i in items]
        cat = lambda name: np.concatenate([getattr(v, name) for v in vecs])  # noqa: E731
        if is_parallel(items[0]):
            return parallel_vec(shape=items[0].det_shape, ray_dir=cat("ray_dir"),
                                det_pos=cat("det_pos"), det_v=cat("det_v"), det_u=cat("det_u"))
        return cone_vec(shape=items[0].det_shape, src_pos=cat("src_pos"), det_pos=cat("det_pos"),
                        det_v=cat("det_v"), det_u=cat("det_u"))
    if all(is_volume(i) for i in items):
        if not all(i.shape == items[0].shape for i in items):
            raise ValueError("Cannot concatenate volumes. Not all shapes are equal.")
        vecs = [i.to_vec() for i in items]
        cat = lambda name: np.concatenate([getattr(v, name) for v in vecs])  # noqa: E731
        return volume_vec(shape=items[0].shape, pos=cat("pos"), w=cat("w"), v=cat("v"), u=cat("u"))
    raise TypeError("Concatenating these types of objects is not supported: "
                    f"{[type(i) for i in items]}")
