"""Fused projection of many (volume, geometry) instances in one launch.

``fused_forward([A_1, ..., A_n], [x_1, ..., x_n])`` returns ``[A_k(x_k)]`` and
``fused_backward`` the backprojections, as one fused kernel launch per up to
16 instances (shapes and geometries may differ) instead of n launches; the
results are bit-identical to the per-operator calls. All tensors must be
float32, contiguous and on the same HIP device. There is no counterpart in
tomosipo, which applies one operator per call (tomosipo/Operator.py:176-233).
"""
import torch

from . import _tsx


def _device_plans(ops, tensors):
    dev = tensors[0].device
    for t in tensors:
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.device == dev):
            raise ValueError("fused projection needs torch tensors on one HIP device")
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError("fused projection needs contiguous float32 tensors")
    index = dev.index if dev.index is not None else torch.cuda.current_device()
    return dev, [A.projector.plan(index) for A in ops]


def _run(ops, inputs, out, forward, additive):
    if len(ops) != len(inputs):
        raise ValueError("one input per operator is required")
    if len(ops) == 0:
        return []
    for A, x in zip(ops, inputs):
        shape = A.domain_shape if forward else A.range_shape
        if tuple(x.shape) != tuple(shape):
            raise ValueError(f"Expected input of shape {tuple(shape)}. Got {tuple(x.shape)}")
    if out is None:
        out = [torch.zeros(A.range_shape if forward else A.domain_shape, device=x.device)
               if additive else
               torch.empty(A.range_shape if forward else A.domain_shape, device=x.device)
               for A, x in zip(ops, inputs)]
    dev, plans = _device_plans(ops, list(inputs) + list(out))
    op = _tsx.TSX_OP_FP if forward else _tsx.TSX_OP_BP
    nbytes = _tsx.multi_workspace_size(plans, op)
    with torch.cuda.device_of(inputs[0]):
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        vols = [x.data_ptr() for x in (inputs if forward else out)]
        projs = [y.data_ptr() for y in (out if forward else inputs)]
        _tsx.multi_project(plans, vols, projs, forward, additive, stream, ws.data_ptr(), nbytes)
    return out


def fused_forward(operators, volumes, out=None, additive=False):
    return _run(operators, volumes, out, True, additive)


def fused_backward(operators, projections, out=None, additive=False):
    return _run(operators, projections, out, False, additive)
