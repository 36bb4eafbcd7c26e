"""``ts.data``: a geometry paired with a linked array (tomosipo ``Data.py:5-135``).

The ASTRA data id of the reference has no counterpart here; the container
keeps the geometry and the link so operators can take and return ``Data``
(``Operator.py:109-113, 189-192, 230-233``).
"""
from .geometry import is_volume, is_projection
from .links.base import link


def data(geometry, initial_value=None):
    if isinstance(initial_value, Data):
        if geometry == initial_value.geometry:
            return initial_value
        raise ValueError(f"Got initial_value={initial_value}, but its geometry does not match {geometry}.")
    return Data(geometry, initial_value)


class Data:
    def __init__(self, geometry, initial_value=None):
        if not (is_volume(geometry) or is_projection(geometry)):
            raise ValueError(f"Geometry '{type(geometry)}' is not supported. "
                             "Cannot determine if volume or projection geometry.")
        self.geometry = geometry
        self._link = link(geometry, initial_value)

    def clone(self):
        return Data(self.geometry, self._link.clone().data)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False

    @property
    def data(self):
        return self._link.data

    @data.setter
    def data(self, val):
        self._link.data = val

    @property
    def link(self):
        return self._link

    def is_volume(self):
        return is_volume(self.geometry)

    def is_projection(self):
        return is_projection(self.geometry)
