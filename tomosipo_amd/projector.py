"""Projector creation and direct projection on the HIP library.

Replaces tomosipo's ASTRA bridge (``tomosipo/astra.py:80-215``):

* ``create_projector(vg, pg, ...)`` takes the place of
  ``create_astra_projector`` (``astra.py:80-98``): the geometry is serialised
  to the ASTRA dictionaries (``vg.to_astra()``, ``pg.to_vec().to_astra()``) and
  validated by ``tsx_plan_create`` on the host at once; device tables are
  built lazily, once per device.
* ``direct_fp`` / ``direct_bp`` / ``direct_project`` follow ``astra.py:104-215``:
  links must be compatible (same device) or ``ValueError``; SET overwrites,
  ADD accumulates.

Voxel supersampling k (``astra.py:94-97``) is defined here so that the
backprojection stays the exact transpose of the forward projection: the
volume is refined k times along each axis (each voxel becomes k^3 voxels of
its value, operator U) and the refined grid is projected, A_k = A_fine U;
the backprojection is A_k^T = U^T A_fine^T (the refined backprojection summed
over each voxel's k^3 sub-voxels). ASTRA applies VoxelSuperSampling to its
voxel-driven backprojector only; A_k x differs from A x only by the finer
sampling of the same piecewise-constant volume.

Torch tensors on a HIP device go to the kernels by pointer on the current
stream of their device (asynchronous, like any torch op). numpy arrays and
CPU tensors are copied to device 0, projected and copied back, which is what
ASTRA does for host arrays (``links/numpy.py:70-77``). Without a HIP device
the call raises ``RuntimeError``; there is no CPU path.
"""
import threading

import numpy as np

from . import _tsx
from .links.base import are_compatible


# Plans with at least this many voxel-angles are autotuned on their device
# when first used (a one-off cost of a few trial launches on bands of detector
# rows and on at most 64 angles, cached per shape by the library).
AUTOTUNE_MIN_VOXEL_ANGLES = 1 << 30


class Projector:
    def __init__(self, volume_geometry, projection_geometry, *, voxel_supersampling=1,
                 detector_supersampling=1, autotune=None):
        from .geometry import VolumeGeometry

        if not isinstance(volume_geometry, VolumeGeometry):
            raise TypeError(f"Expected an axis-aligned VolumeGeometry. Got {type(volume_geometry)}.")
        self.vol_geom = volume_geometry.to_astra()
        self.proj_geom = projection_geometry.to_vec().to_astra()
        self.voxel_supersampling = int(voxel_supersampling)
        self.detector_supersampling = int(detector_supersampling)
        if self.voxel_supersampling < 1:
            raise ValueError(f"voxel_supersampling must be >= 1; got {voxel_supersampling}")
        k = self.voxel_supersampling
        # the kernels project the k-times refined grid of the same window
        self.plan_vol_geom = dict(self.vol_geom)
        for key in ("GridColCount", "GridRowCount", "GridSliceCount"):
            self.plan_vol_geom[key] = int(self.vol_geom[key]) * k
        self.vol_shape = tuple(volume_geometry.shape)
        rows, cols = projection_geometry.det_shape
        self.proj_shape = (rows, projection_geometry.num_angles, cols)
        if autotune is None:
            n = int(np.prod(self.vol_shape)) * projection_geometry.num_angles
            autotune = n >= AUTOTUNE_MIN_VOXEL_ANGLES
        self.autotune = bool(autotune)
        # Host-only plan: validates now, as astra.create_projector would.
        self._host_plan = self._make_plan(-1)
        self._plans = {}
        self._lock = threading.Lock()

    def _make_plan(self, device):
        return _tsx.Plan(self.plan_vol_geom, self.proj_geom, device=device,
                         det_supersampling=self.detector_supersampling, vox_supersampling=1)

    def plan(self, device):
        """The device plan for ``device`` (created on first use)."""
        p = self._plans.get(device)
        if p is None:
            with self._lock:
                p = self._plans.get(device)
                if p is None:
                    p = self._make_plan(device)
                    if self.autotune:
                        torch = _torch()
                        p.autotune(torch.cuda.current_stream(device).cuda_stream)
                    self._plans[device] = p
        return p

    def axes(self):
        return self._host_plan.axes()


def create_projector(volume_geometry, projection_geometry, *, voxel_supersampling=1,
                     detector_supersampling=1, autotune=None):
    return Projector(volume_geometry, projection_geometry, voxel_supersampling=voxel_supersampling,
                     detector_supersampling=detector_supersampling, autotune=autotune)


def _torch():
    try:
        import torch
    except ImportError:
        raise RuntimeError("tomosipo_amd projects on a HIP device through torch; torch is not importable")
    return torch


def _require_device(torch):
    if not torch.cuda.is_available():
        raise RuntimeError("tomosipo_amd needs a HIP (ROCm) device for projections; none is "
                           "available and there is no CPU fallback")


def _refine(x, k):
    """U: each voxel of the trailing three dims becomes k^3 voxels of its value."""
    for d in (-3, -2, -1):
        x = x.repeat_interleave(k, dim=d)
    return x.contiguous()


def _coarsen(x, k, shape):
    """U^T: the sum over each voxel's k^3 sub-voxels."""
    nz, ny, nx = shape
    lead = tuple(x.shape[:-3])
    return x.reshape(lead + (nz, k, ny, k, nx, k)).sum(dim=(-5, -3, -1))


def _run(projector, vol_t, proj_t, forward, additive, batch=1, vol_stride=0, proj_stride=0):
    """Launch on device tensors (contiguous, same device; float32, or float64 for the
    FP64 kernels)."""
    k = projector.voxel_supersampling
    if k > 1:
        # through the refined grid: A_k = A_fine U, A_k^T = U^T A_fine^T
        torch = _torch()
        shape = projector.vol_shape
        lead = (batch,) if batch > 1 else ()
        fine_shape = tuple(n * k for n in shape)
        fine_numel = int(np.prod(fine_shape))
        vol_view = vol_t.reshape(-1)[:batch * int(np.prod(shape))].reshape(lead + shape) \
            if vol_stride in (0, int(np.prod(shape))) else None
        if vol_view is None:
            raise ValueError("voxel supersampling needs densely packed batch entries")
        if forward:
            fine = _refine(vol_view, k)
            _run_plan(projector, fine, proj_t, True, additive, batch, fine_numel, proj_stride)
        else:
            fine = torch.empty(lead + fine_shape, dtype=vol_t.dtype, device=vol_t.device)
            _run_plan(projector, fine, proj_t, False, False, batch, fine_numel, proj_stride)
            coarse = _coarsen(fine, k, shape)
            with torch.no_grad():
                if additive:
                    vol_view += coarse
                else:
                    vol_view.copy_(coarse)
        return
    _run_plan(projector, vol_t, proj_t, forward, additive, batch, vol_stride, proj_stride)


def _run_plan(projector, vol_t, proj_t, forward, additive, batch, vol_stride, proj_stride):
    """Launch on device tensors (contiguous, same device; float32, or float64 for the
    FP64 kernels)."""
    torch = _torch()
    dev = vol_t.device.index if vol_t.device.index is not None else torch.cuda.current_device()
    plan = projector.plan(dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    if vol_t.dtype == torch.float64 and proj_t.dtype == torch.float64:
        fn = plan.forward_f64 if forward else plan.backward_f64
        fn(vol_t.data_ptr(), proj_t.data_ptr(), additive, stream, batch, vol_stride, proj_stride)
        return
    if vol_t.dtype != torch.float32 or proj_t.dtype != torch.float32:
        raise TypeError("projection kernels take float32 (or float64 on both sides)")
    # Scratch memory comes from torch's caching allocator on the launch stream, so
    # that torch owns every allocation (and HIP graph capture sees none).
    nbytes = plan.workspace_size(_tsx.TSX_OP_FP if forward else _tsx.TSX_OP_BP, batch)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=vol_t.device)
    if forward:
        plan.forward(vol_t.data_ptr(), proj_t.data_ptr(), additive, stream, batch, vol_stride,
                     proj_stride, ws.data_ptr(), nbytes)
    else:
        plan.backward(vol_t.data_ptr(), proj_t.data_ptr(), additive, stream, batch, vol_stride,
                      proj_stride, ws.data_ptr(), nbytes)


def direct_project(projector, vol_link, proj_link, forward=None, additive=False):
    if forward is None:
        raise ValueError("project must be given a forward argument (True/False).")
    if not are_compatible(vol_link, proj_link):
        raise ValueError(
            "Cannot perform projection on volume and projection data, because they are not "
            "compatible. Usually, this indicates that the data are located on different computing "
            "devices. ")
    torch = _torch()
    # a link over another library's array (DLPack) is read through its torch view
    vol = getattr(vol_link, "tensor", vol_link.data)
    proj = getattr(proj_link, "tensor", proj_link.data)
    on_device = isinstance(vol, torch.Tensor) and vol.is_cuda
    with vol_link.context():
        if on_device:
            _run(projector, vol, proj, forward, additive)
            return
        _require_device(torch)
        dev = torch.device("cuda", torch.cuda.current_device())
        to_dev = (lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)) \
            if isinstance(vol, np.ndarray) else (lambda a: a.detach().to(dev))
        vol_d = to_dev(vol)
        proj_d = to_dev(proj)
        _run(projector, vol_d, proj_d, forward, additive)
        dst, src = (proj, proj_d) if forward else (vol, vol_d)
        if isinstance(dst, np.ndarray):
            dst[...] = src.cpu().numpy()
        else:
            with torch.no_grad():
                dst.copy_(src.cpu())


def direct_fp(projector, vol_link, proj_link, additive=False):
    return direct_project(projector, vol_link, proj_link, forward=True, additive=additive)


def direct_bp(projector, vol_link, proj_link, additive=False):
    return direct_project(projector, vol_link, proj_link, forward=False, additive=additive)


def _run_fdk(projector, vol_t, proj_t, additive):
    torch = _torch()
    if projector.voxel_supersampling != 1:
        raise ValueError("FDK supports voxel_supersampling = 1 only")
    if vol_t.dtype != torch.float32 or proj_t.dtype != torch.float32:
        raise TypeError("FDK takes float32 volume and projections")
    dev = vol_t.device.index if vol_t.device.index is not None else torch.cuda.current_device()
    plan = projector.plan(dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    nbytes = plan.fdk_workspace_size()
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=vol_t.device)
    plan.fdk(vol_t.data_ptr(), proj_t.data_ptr(), additive, stream, ws.data_ptr(), nbytes)


def direct_fdk(projector, vol_link, proj_link, additive=True):
    """FDK reconstruction of ``proj_link`` into ``vol_link`` (``tsx_fdk``), with the
    staging rules of ``direct_project``: device tensors in place on their stream,
    numpy arrays and CPU tensors through device memory."""
    if not are_compatible(vol_link, proj_link):
        raise ValueError(
            "Cannot perform FDK on volume and projection data, because they are not "
            "compatible. Usually, this indicates that the data are located on different computing "
            "devices. ")
    torch = _torch()
    vol = getattr(vol_link, "tensor", vol_link.data)
    proj = getattr(proj_link, "tensor", proj_link.data)
    with vol_link.context():
        if isinstance(vol, torch.Tensor) and vol.is_cuda:
            _run_fdk(projector, vol, proj, additive)
            return
        _require_device(torch)
        dev = torch.device("cuda", torch.cuda.current_device())
        to_dev = (lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)) \
            if isinstance(vol, np.ndarray) else (lambda a: a.detach().to(dev))
        vol_d = to_dev(vol)
        _run_fdk(projector, vol_d, to_dev(proj), additive)
        if isinstance(vol, np.ndarray):
            vol[...] = vol_d.cpu().numpy()
        else:
            with torch.no_grad():
                vol.copy_(vol_d.cpu())
