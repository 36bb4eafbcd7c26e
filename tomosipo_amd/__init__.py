"""tomosipo_amd: tomosipo's projection operator API on hand-written HIP kernels for MI355X.

``import tomosipo_amd as ts`` exposes the same names as ``tomosipo`` for the
projection path (tomosipo ``__init__.py:10-41``): geometry constructors,
transforms, ``ts.operator``, ``ts.link``, ``ts.data``, ``ts.phantom`` and
``ts.epsilon``. Projections run in ``libtsx.so`` (HIP, gfx950); there is no
CPU fallback.
"""
__version__ = "0.1.0"

# Geometry equality tolerance and the SIRT clamp constant (tomosipo __init__.py:41).
epsilon = 1e-8

from . import types  # noqa: E402
from . import vector_calc  # noqa: E402
from . import geometry  # noqa: E402
from .geometry import (  # noqa: E402
    volume, volume_vec, parallel, parallel_vec, cone, cone_vec, translate, scale, rotate, reflect,
    to_perspective, from_perspective, concatenate, from_astra, to_astra,
)
from .geometry import random_transform  # noqa: E402,F401
from .links.base import link  # noqa: E402
from .Data import data  # noqa: E402
from .Operator import operator  # noqa: E402
from . import phantom  # noqa: E402
from . import astra  # noqa: E402
from .fdk import fdk  # noqa: E402
