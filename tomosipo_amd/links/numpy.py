"""numpy link (tomosipo ``links/numpy.py:7-144``).

Shape mismatch raises ``ValueError``; non-float32 or non-contiguous arrays are
converted with a ``UserWarning``; a scalar fills a new float32 array.
"""
import warnings

import numpy as np

from .base import Link, backends


class NumpyLink(Link):
    def __init__(self, shape, initial_value):
        super().__init__(shape, initial_value)
        if initial_value is None:
            self._data = np.zeros(shape, dtype=np.float32)
        elif np.isscalar(initial_value):
            self._data = np.zeros(shape, dtype=np.float32)
            self._data[...] = initial_value
        else:
            initial_value = np.asarray(initial_value)
            if initial_value.shape != tuple(shape):
                raise ValueError(f"Cannot link array. Expected array of shape {tuple(shape)}. "
                                 f"Got {initial_value.shape}")
            if initial_value.dtype != np.float32:
                warnings.warn(
                    f"The parameter initial_value is of type {initial_value.dtype}; expected "
                    "`np.float32`. The type has been Automatically converted. "
                    "Use `ts.link(x.astype(np.float32))' to inhibit this warning. ")
                initial_value = initial_value.astype(np.float32)
            if not (initial_value.flags["C_CONTIGUOUS"] and initial_value.flags["ALIGNED"]):
                warnings.warn(
                    "The parameter initial_value should be C_CONTIGUOUS and ALIGNED. It has been "
                    "automatically made contiguous and aligned. "
                    "Use `ts.link(np.ascontiguousarray(x))' to inhibit this warning. ")
                initial_value = np.ascontiguousarray(initial_value)
            self._data = initial_value

    @staticmethod
    def __accepts__(initial_value):
        return initial_value is None or isinstance(initial_value, np.ndarray) or np.isscalar(initial_value)

    def __compatible_with__(self, other):
        return True if isinstance(other, NumpyLink) else NotImplemented

    @property
    def data(self):
        return self._data

    @data.setter
    def data(self, val):
        Link.data.fset(self, val)

    def new_zeros(self, shape):
        return NumpyLink(shape, np.zeros(shape, dtype=self._data.dtype))

    def new_full(self, shape, value):
        return NumpyLink(shape, np.full(shape, value, dtype=self._data.dtype))

    def new_empty(self, shape):
        return NumpyLink(shape, np.empty(shape, dtype=self._data.dtype))

    def clone(self):
        return NumpyLink(self._data.shape, np.copy(self._data))


backends.append(NumpyLink)
