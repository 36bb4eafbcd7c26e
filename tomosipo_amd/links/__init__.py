"""Array links: numpy always, torch when importable (tomosipo ``links/__init__.py``)."""
from . import base
from . import numpy
from .base import are_compatible, geometry_shape, link

try:
    from . import torch  # noqa: F401
except ImportError:  # torch is optional for geometry-only use
    pass
