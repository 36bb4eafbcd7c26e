"""The ASTRA-shaped entry points of tomosipo (``tomosipo/astra.py``), on the HIP library.

* ``create_astra_projector`` / ``direct_project`` / ``direct_fp`` /
  ``direct_bp``: the projector bridge (``astra.py:80-215``; implemented in
  ``tomosipo_amd.projector``).
* ``project`` / ``forward`` / ``backward``: the legacy composite projection
  (``astra.py:221-371``), which projects every volume dataset onto every
  projection dataset. Forward: each projection dataset becomes (SET) or gains
  (ADD) the sum over all volumes of their forward projections; backward:
  each volume becomes or gains the sum over all projection datasets of their
  backprojections (the semantics of ASTRA's composite call that tomosipo
  makes, ``astra.py:279-287``). Each (volume, projection) pair is projected
  with its own geometries; the options come from the keyword arguments or
  from a given ``projector``.
* ``fdk``: ``astra.py:374-406``; the FDK reconstruction of a projection
  dataset is added to a volume dataset.

Volume supersampling (``voxel_supersampling``) applies to the backprojection
only, as in ASTRA's projector options (``astra.py:94-97``): ``backward`` then
is the exact transpose of the forward projection of the k-times refined
volume (``tomosipo_amd.projector``), and ``forward`` ignores it.
"""
from .projector import (create_projector as create_astra_projector, direct_project, direct_fp,  # noqa: F401
                        direct_bp, direct_fdk, Projector)
from .geometry import from_astra, to_astra  # noqa: F401  (astra.py:37-77)


def _pair_projector(vg, pg, voxel_supersampling, detector_supersampling, forward, given, cache):
    key = (id(vg), id(pg))
    p = cache.get(key)
    if p is None:
        vs = 1 if forward else voxel_supersampling
        if (given is not None and given.voxel_supersampling == vs
                and given.vol_shape == tuple(vg.shape) and _same(given, vg, pg)):
            p = given
        else:
            p = create_astra_projector(vg, pg, voxel_supersampling=vs,
                                       detector_supersampling=detector_supersampling)
        cache[key] = p
    return p


def _same(projector, vg, pg):
    import numpy as np

    a = vg.to_astra()
    b = pg.to_vec().to_astra()
    if a != projector.vol_geom:
        return False
    return (b["type"] == projector.proj_geom["type"]
            and b["DetectorRowCount"] == projector.proj_geom["DetectorRowCount"]
            and b["DetectorColCount"] == projector.proj_geom["DetectorColCount"]
            and np.array_equal(b["Vectors"], projector.proj_geom["Vectors"]))


def _fused_ok(vol_data, proj_data):
    try:
        import torch
    except ImportError:
        return False
    arrays = [d.data for d in vol_data + proj_data]
    if not all(isinstance(a, torch.Tensor) and a.is_cuda for a in arrays):
        return False
    return len({a.device for a in arrays}) == 1


def project(*data, voxel_supersampling=1, detector_supersampling=1, forward=None, additive=False,
            projector=None):
    """Project all volumes on all projection datasets (``forward=True``) or
    backproject all projection datasets on all volumes (``forward=False``);
    tomosipo ``astra.py:221-287``. ``data``: ``ts.data`` objects, at least one
    volume and one projection dataset. ``additive`` adds to the existing data
    instead of overwriting it. ``projector`` (as ``Operator.astra_projector``)
    supplies the supersampling options."""
    vol_data = [d for d in data if d.is_volume()]
    proj_data = [d for d in data if d.is_projection()]
    if forward is None:
        raise ValueError("project must be given a forward argument (True/False).")
    if len(vol_data) < 1 or len(proj_data) < 1:
        raise ValueError("Expected at least one projection dataset and one volume dataset")
    if projector is not None:
        if not isinstance(projector, Projector):
            raise TypeError(f"Expected a projector from create_astra_projector; got {type(projector)}")
        voxel_supersampling = projector.voxel_supersampling
        detector_supersampling = projector.detector_supersampling
    cache = {}

    def proj_of(vd, pd):
        return _pair_projector(vd.geometry, pd.geometry, voxel_supersampling, detector_supersampling,
                               forward, projector, cache)

    # outputs: the projection datasets (forward) or the volumes (backward); the
    # first contribution to an output overwrites it unless additive
    outs, ins = (proj_data, vol_data) if forward else (vol_data, proj_data)
    if _fused_ok(vol_data, proj_data) and (forward or voxel_supersampling == 1):
        # one fused launch per input dataset over every output dataset (no two
        # instances of a launch write the same output)
        from .fused import fused_forward, fused_backward

        for n, src in enumerate(ins):
            pairs = [(src, dst) if forward else (dst, src) for dst in outs]
            ops = [_PairOperator(proj_of(vd, pd), vd, pd) for vd, pd in pairs]
            add = additive or n > 0
            fn = fused_forward if forward else fused_backward
            fn(ops, [src.data] * len(outs), out=[d.data for d in outs], additive=add)
        return
    for n, src in enumerate(ins):
        for dst in outs:
            vd, pd = (src, dst) if forward else (dst, src)
            direct_project(proj_of(vd, pd), vd.link, pd.link, forward=forward,
                           additive=additive or n > 0)


class _PairOperator:
    """The slice of the Operator interface the fused launches read."""

    def __init__(self, projector, vd, pd):
        from .links.base import geometry_shape

        self.projector = projector
        self.domain_shape = geometry_shape(vd.geometry)
        self.range_shape = geometry_shape(pd.geometry)


def forward(*data, voxel_supersampling=1, detector_supersampling=1, additive=False, projector=None):
    """Forward project every volume onto every projection dataset (``astra.py:290-328``)."""
    project(*data, voxel_supersampling=voxel_supersampling, detector_supersampling=detector_supersampling,
            additive=additive, forward=True, projector=projector)


def backward(*data, voxel_supersampling=1, detector_supersampling=1, additive=False, projector=None):
    """Backproject every projection dataset onto every volume (``astra.py:331-371``)."""
    project(*data, voxel_supersampling=voxel_supersampling, detector_supersampling=detector_supersampling,
            additive=additive, forward=False, projector=projector)


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
