"""FDK reconstruction for a cone-beam operator.

``fdk(A, y)`` reconstructs projections ``y`` (shape ``A.range_shape``) of the
operator ``A = ts.operator(vg, pg)`` into a new volume (``A.domain_shape``),
in the operator's normalised geometry (``Operator.py:11-60``), on the HIP
library (``tsx_fdk``). The reference reaches FDK only through its legacy
``ts.astra.fdk`` (``astra.py:374-406``), which ``tomosipo_amd.astra.fdk``
mirrors.
"""
from .links.base import link
from .projector import direct_fdk


def fdk(A, y, out=None):
    """FDK reconstruction of ``y`` for ``A``; ``out`` (optional) is overwritten."""
    from .Operator import Operator

    if not isinstance(A, Operator):
        raise TypeError("fdk needs a tomosipo_amd Operator")
    plink = link(A.astra_compat_pg, y)
    vlink = plink.new_empty(A.domain_shape) if out is None else link(A.astra_compat_vg, out)
    direct_fdk(A.projector, vlink, plink, additive=False)
    return vlink.data
