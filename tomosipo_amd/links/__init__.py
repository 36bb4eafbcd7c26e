"""Array links: numpy always; torch and DLPack arrays when torch is importable
(tomosipo ``links/__init__.py``)."""
from . import base
from . import numpy
from .base import are_compatible, geometry_shape, link

try:
    from . import torch  # noqa: F401
    from . import dlpack  # noqa: F401  (registered after numpy and torch)
except ImportError:  # torch is optional for geometry-only use
    pass
