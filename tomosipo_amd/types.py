"""Argument normalisation for geometry constructors.

Behaviour follows tomosipo's ``types.py`` (``to_shape2d``/``to_shape3d``
``types.py:88-140``, ``to_size2d``/``to_size3d`` ``types.py:143-190``,
``to_pos`` ``types.py:193-214``, ``to_scalars`` ``types.py:217-270``,
``to_vec``): shapes are positive integer tuples, sizes are non-negative float
tuples, positions are 3-tuples and vectors are ``(n, 3)`` float64 arrays.
"""
from collections import abc
from numbers import Integral

import numpy as np

_EPS = 1e-8


def to_tuple(val, n):
    n = int(n)
    if isinstance(val, abc.Iterable) and not isinstance(val, (str, bytes)):
        val = tuple(val)
        if len(val) != n:
            raise TypeError(f"Expected tuple with {n} elements. Got {val!r}.")
        return val
    return (val,) * n


def to_float_tuple(val, n, var_name="value"):
    val = to_tuple(val, n)
    try:
        return tuple(float(v) for v in val)
    except (TypeError, ValueError):
        raise TypeError(f"{var_name} must contain only floats. Got {val!r}.")


def _to_shape(shape, n):
    shape = to_tuple(shape, n)
    for s in shape:
        if not isinstance(s, Integral):
            raise TypeError(
                f"Shape must contain only integers. Got {shape} with type {type(s)}."
            )
    shape = tuple(int(s) for s in shape)
    if not all(s >= 1 for s in shape):
        raise TypeError(f"Shape must be positive. Got {shape}.")
    return shape


def to_shape2d(shape):
    return _to_shape(shape, 2)


def to_shape3d(shape):
    return _to_shape(shape, 3)


def _to_size(size, n):
    size = to_float_tuple(size, n, var_name="Size")
    if not all(s >= -_EPS for s in size):
        raise TypeError(f"Size must be non-negative. Got {size}.")
    return size


def to_size2d(size):
    return _to_size(size, 2)


def to_size3d(size):
    return _to_size(size, 3)


def to_pos(pos):
    if np.isscalar(pos) and pos == 0.0:
        return (0.0, 0.0, 0.0)
    if isinstance(pos, abc.Iterable):
        return to_float_tuple(pos, 3, "Position")
    raise TypeError(
        f"Cannot convert value to position. Expected (float, float, float). Got {pos}."
    )


def to_scalars(s, var_name="scalars", accept_empty=False):
    """A 1-d float64 array; a scalar becomes shape ``(1,)``."""
    try:
        arr = np.array(s, dtype=np.float64, copy=True, ndmin=1)
    except (TypeError, ValueError):
        raise TypeError(f"Could not convert {var_name} to np.array. Got: {s!r}.")
    if arr.ndim != 1 or (arr.size == 0 and not accept_empty):
        raise TypeError(
            f"Value cannot be converted to {var_name}. Expected shape: (N,). "
            f"Got shape: {arr.shape}."
        )
    if np.any(np.isnan(arr)):
        raise TypeError(f"Could not convert to array of {var_name}: array contains NaN.")
    return arr


def to_vec(x, var_name="vector"):
    """An ``(n, 3)`` float64 array; a single 3-vector becomes ``(1, 3)``."""
    arr = np.array(x, dtype=np.float64, copy=True)
    if arr.ndim == 1:
        arr = arr[None, :]
    if arr.ndim != 2 or arr.shape[1] != 3:
        raise ValueError(
            f"Expected {var_name} to have shape (n, 3) or (3,). Got {arr.shape}."
        )
    return arr


def to_homogeneous(x, s):
    x = np.array(x, dtype=np.float64, copy=True)
    if x.ndim == 1:
        x = x[None, :]
    if x.ndim != 2 or x.shape[1] not in (3, 4):
        raise ValueError(f"Expected shape (n, 3) or (n, 4). Got {x.shape}.")
    if x.shape[1] == 3:
        x = np.concatenate([x, np.full((x.shape[0], 1), float(s))], axis=1)
    return x


def to_homogeneous_vec(x):
    return to_homogeneous(x, 0.0)


def to_homogeneous_pos(x):
    return to_homogeneous(x, 1.0)
