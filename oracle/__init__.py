"""FP64 CPU oracle for the projection path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker or the reported
CPU baseline. Nothing under ``tomosipo_amd/`` imports it; the product path
runs on the HIP library or raises.

Two independent restatements of the same projector:

* ``liboracle.so`` (``joseph_oracle.c``): Joseph forward projection in grid
  coordinates and its exact transpose, FP64, OpenMP. Normative.
* ``numpy_twin``: a dense-matrix restatement in world coordinates for tiny
  problems (a few hundred voxels), written separately so that a convention
  mistake in one of them shows up as a disagreement.
* ``fdk``: the FDK reconstruction (cosine weighting, Ram-Lak filter,
  voxel-driven backprojection) in FP64 numpy, for small problems.

The geometry handed to both is the ASTRA form that tomosipo passes to its
projector (``tomosipo/astra.py:80-98``): the volume dictionary
(``geometry/volume.py:286-301``) and ``(n, 12)`` vectors
``[ray|src, det, u, v]`` in ``(x, y, z)`` order (``geometry/cone_vec.py:174-191``,
``geometry/parallel_vec.py:175-192``).
"""
from .oracle import (OracleGeometry, oracle_geometry, fp, bp, axes, num_threads, lib_path,
                     build)
from . import numpy_twin
from . import fdk

__all__ = ["OracleGeometry", "oracle_geometry", "fp", "bp", "axes", "num_threads", "lib_path",
           "build", "numpy_twin", "fdk"]
