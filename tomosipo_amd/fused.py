"""Fused projection of many (volume, geometry) instances in one launch.

``fused_forward([A_1, ..., A_n], [x_1, ..., x_n])`` returns ``[A_k(x_k)]`` and
``fused_backward`` the backprojections, as one fused kernel launch per up to
16 instances (shapes and geometries may differ) instead of n launches; the
results are bit-identical to the per-operator calls. All tensors must be
float32, contiguous and on the same HIP device. There is no counterpart in
tomosipo, which applies one operator per call (tomosipo/Operator.py:176-233).

Each instance may carry its own epilogue (``Epilogue``): overwrite,
accumulate, or one half of a SIRT iteration (``residual``: out = w * (y - A x);
``scaled_add``: out += w * A.T(y)), so that plain projections and SIRT
iterations of different requests share a launch.

``fused_normal(ops, xs)`` returns the projections ``y_k = A_k(x_k)`` and their
backprojections ``b_k = A_k.T(y_k)``, pipelined over groups of instances: the
fused backprojection of one group runs on a second stream while the fused
forward projection of the next runs on the caller's, so that the forward
projector (bound by the vector-memory address path) and the backprojector
(bound by the vector ALU) of different instances share the compute units.
Results are bit-identical to the sequential calls.
"""
import dataclasses

import torch

from . import _tsx


@dataclasses.dataclass
class Epilogue:
    """How an instance's projection r is written to its output o.

    mode: "set" (o = r), "add" (o += r), "residual" (o = w * (y - r); forward
    only) or "scaled_add" (o += w * r). ``y`` and ``w`` have the output's shape.
    """
    mode: str = "set"
    y: object = None
    w: object = None


_MODES = {"set": _tsx.TSX_EP_SET, "add": _tsx.TSX_EP_ADD, "residual": _tsx.TSX_EP_RESIDUAL,
          "scaled_add": _tsx.TSX_EP_SCALED_ADD}


def _device_plans(ops, tensors):
    for k, A in enumerate(ops):
        if A.projector.voxel_supersampling != 1:
            raise ValueError(f"operator {k}: fused launches take voxel_supersampling = 1 only "
                             "(apply operators with voxel supersampling one at a time)")
    dev = tensors[0].device
    for t in tensors:
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.device == dev):
            raise ValueError("fused projection needs torch tensors on one HIP device")
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError("fused projection needs contiguous float32 tensors")
    index = dev.index if dev.index is not None else torch.cuda.current_device()
    return dev, [A.projector.plan(index) for A in ops]


def _check_shapes(tensors, shapes, what):
    for k, (t, shape) in enumerate(zip(tensors, shapes)):
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"{what} {k} must be a torch tensor, got {type(t)}")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"Expected {what} {k} of shape {tuple(shape)}. Got {tuple(t.shape)}")


def _run(ops, inputs, out, forward, additive, epilogues=None):
    if len(ops) != len(inputs):
        raise ValueError("one input per operator is required")
    if len(ops) == 0:
        return []
    in_shapes = [A.domain_shape if forward else A.range_shape for A in ops]
    out_shapes = [A.range_shape if forward else A.domain_shape for A in ops]
    _check_shapes(inputs, in_shapes, "input")
    if out is None:
        zero = additive or (epilogues is not None and any(e.mode in ("add", "scaled_add")
                                                          for e in epilogues))
        out = [torch.zeros(s, device=x.device) if zero else torch.empty(s, device=x.device)
               for s, x in zip(out_shapes, inputs)]
    else:
        out = list(out)
        if len(out) != len(ops):
            raise ValueError(f"one output per operator is required: got {len(out)} for {len(ops)}")
        _check_shapes(out, out_shapes, "output")
    extra = []
    eps = None
    if epilogues is not None:
        epilogues = list(epilogues)
        if len(epilogues) != len(ops):
            raise ValueError("one epilogue per operator is required")
        eps = []
        for k, e in enumerate(epilogues):
            if e.mode not in _MODES:
                raise ValueError(f"epilogue {k}: unknown mode {e.mode!r}")
            if e.mode == "residual" and not forward:
                raise ValueError(f"epilogue {k}: the residual form is for forward projections")
            need = {"residual": ("y", "w"), "scaled_add": ("w",)}.get(e.mode, ())
            ptrs = {}
            for name in ("y", "w"):
                t = getattr(e, name)
                if name in need:
                    if t is None:
                        raise ValueError(f"epilogue {k}: mode {e.mode!r} needs {name}")
                    _check_shapes([t], [out_shapes[k]], f"epilogue {k} {name}")
                    extra.append(t)
                    ptrs[name] = t.data_ptr()
                else:
                    ptrs[name] = None
            eps.append((_MODES[e.mode], ptrs["y"], ptrs["w"]))
    dev, plans = _device_plans(ops, list(inputs) + list(out) + extra)
    op = _tsx.TSX_OP_FP if forward else _tsx.TSX_OP_BP
    nbytes = _tsx.multi_workspace_size(plans, op)
    with torch.cuda.device_of(inputs[0]):
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        vols = [x.data_ptr() for x in (inputs if forward else out)]
        projs = [y.data_ptr() for y in (out if forward else inputs)]
        _tsx.multi_project(plans, vols, projs, forward, additive, stream, ws.data_ptr(), nbytes,
                           epilogues=eps)
    return out


def fused_forward(operators, volumes, out=None, additive=False, epilogues=None):
    return _run(operators, volumes, out, True, additive, epilogues)


def fused_backward(operators, projections, out=None, additive=False, epilogues=None):
    return _run(operators, projections, out, False, additive, epilogues)


_SIDE_STREAMS = {}


def _side_stream(dev):
    """A second stream on `dev` (default priority), created once per device."""
    index = dev.index if dev.index is not None else torch.cuda.current_device()
    if index not in _SIDE_STREAMS:
        _SIDE_STREAMS[index] = torch.cuda.Stream(device=index)
    return _SIDE_STREAMS[index]


def fused_normal(operators, volumes, projections=None, out=None, groups=8, events=None):
    """``(ys, bs)`` with ``ys[k] = A_k(x_k)`` and ``bs[k] = A_k.T(ys[k])``.

    The instances are taken in ``groups`` consecutive groups (at most one per
    instance); group g's fused forward projection runs on the caller's current
    stream, its fused backprojection on a side stream after it, overlapping
    the forward projection of group g + 1. The caller's stream waits for the
    side stream before returning, so later work on it sees every result.
    ``projections`` and ``out`` may hold the output tensors. With ``events`` (a
    list), each group appends ``(first, last, fp0, fp1, bp0, bp1)``: its instance
    range and timing events recorded around its forward launch (caller's
    stream) and its backprojection launch (side stream).
    """
    ops, xs = list(operators), list(volumes)
    n = len(ops)
    if len(xs) != n:
        raise ValueError("one volume per operator is required")
    if n == 0:
        return [], []
    _check_shapes(xs, [A.domain_shape for A in ops], "input")
    dev = xs[0].device
    ys = list(projections) if projections is not None else \
        [torch.empty(A.range_shape, device=dev) for A in ops]
    bs = list(out) if out is not None else [torch.empty(A.domain_shape, device=dev) for A in ops]
    if len(ys) != n or len(bs) != n:
        raise ValueError("one projection and one output per operator are required")
    _check_shapes(ys, [A.range_shape for A in ops], "projection")
    _check_shapes(bs, [A.domain_shape for A in ops], "output")
    _device_plans(ops, xs + ys + bs)
    groups = max(1, min(int(groups), n))
    bounds = [round(g * n / groups) for g in range(groups + 1)]
    main = torch.cuda.current_stream(dev)
    side = _side_stream(dev)
    for t in ys + bs:
        t.record_stream(side)  # (used on the side stream: not reused before its work is done)
    for g in range(groups):
        sl = slice(bounds[g], bounds[g + 1])
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)] if events is not None else None
        if ev:
            ev[0].record(main)
        _run(ops[sl], xs[sl], ys[sl], True, False)
        done = torch.cuda.Event()
        done.record(main)
        if ev:
            ev[1].record(main)
        with torch.cuda.stream(side):
            side.wait_event(done)
            if ev:
                ev[2].record(side)
            _run(ops[sl], ys[sl], bs[sl], False, False)
            if ev:
                ev[3].record(side)
        if ev:
            events.append((bounds[g], bounds[g + 1] - 1, *ev))
    main.wait_stream(side)
    return ys, bs
