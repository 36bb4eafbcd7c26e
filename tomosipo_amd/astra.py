"""The ASTRA-shaped entry points of tomosipo (``tomosipo/astra.py``), on the HIP library.

``create_astra_projector`` / ``direct_project`` / ``direct_fp`` / ``direct_bp``
are the projector bridge (``astra.py:80-215``; implemented in
``tomosipo_amd.projector``). ``fdk`` follows ``astra.py:374-406``: the FDK
reconstruction of a projection dataset is added to a volume dataset.
"""
from .projector import (create_projector as create_astra_projector, direct_project, direct_fp,  # noqa: F401
                        direct_bp, direct_fdk)


def fdk(vol_data, proj_data, *, voxel_supersampling=1, detector_supersampling=1):
    """Do an FDK reconstruction of the given geometry (``astra.py:374-406``).

    Expects a single volume dataset and a single projection dataset (``ts.data``)
    of a cone-beam geometry; the reconstruction is added to ``vol_data``, as
    ASTRA's ``accumulate_FDK`` does. Supersampling other than 1 raises
    ``ValueError``.
    """
    if voxel_supersampling != 1 or detector_supersampling != 1:
        raise ValueError("FDK supports voxel_supersampling = detector_supersampling = 1 only")
    projector = create_astra_projector(vol_data.geometry, proj_data.geometry)
    direct_fdk(projector, vol_data.link, proj_data.link, additive=True)
