"""PyTorch autograd support (tomosipo ``torch_support.py``).

``to_autograd(A, num_extra_dims=0, is_2d=False)`` returns a function whose
forward is ``A`` and whose backward is ``A.T`` (``torch_support.py:24-130``);
``AutogradOperator`` / ``autograd_operator`` wrap an operator with the same
interface (``torch_support.py:205-370``). Outputs are always float32 and the
number of dimensions is asserted as in the reference (``:33-42``).

Where the reference loops over the extra (batch, channel) dimensions with one
synchronous projector call per slice (``torch_support.py:49-53, 70-74``), all
slices of a device tensor go to the kernels in one batched launch. Additive
operators are projected with overwrite semantics here, since the output
buffer is freshly allocated.
"""
import itertools
import warnings

import torch
from torch.autograd import Function

from .Operator import BackprojectionOperator
from .projector import _run


def _base_and_dir(operator):
    """(Operator, True for forward / False for backprojection)."""
    forward = True
    if isinstance(operator, BackprojectionOperator):
        operator, forward = operator.parent, False
    if isinstance(operator, AutogradOperator):
        operator = operator.operator
    return operator, forward


def _apply(operator, inp, out, n_extra):
    """out[...] = operator(inp[...]) over the leading ``n_extra`` dims, in the dtype of ``out``."""
    base, forward = _base_and_dir(operator)
    extra = tuple(inp.shape[:n_extra])
    batch = 1
    for d in extra:
        batch *= int(d)
    if batch == 0:
        return
    if out.dtype == torch.float64:
        if not (inp.is_cuda and out.is_cuda):
            raise RuntimeError("the float64 projector runs on a HIP device only")
    if inp.is_cuda and out.is_cuda and inp.device == out.device and out.is_contiguous():
        x = inp.detach().to(out.dtype).contiguous()
        in_numel = x[(0,) * n_extra].numel() if n_extra else x.numel()
        out_numel = out[(0,) * n_extra].numel() if n_extra else out.numel()
        with torch.cuda.device_of(x):
            if forward:
                _run(base.projector, x, out, True, False, batch, in_numel, out_numel)
            else:
                _run(base.projector, out, x, False, False, batch, out_numel, in_numel)
        return
    # host tensors: one staged call per slice
    fn = base._fp if forward else base._bp
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for sub in itertools.product(*[range(d) for d in extra]):
            res = fn(inp[sub].detach().to(torch.float32).contiguous())
            out[sub].copy_(res)


class OperatorFunction(Function):
    @staticmethod
    def forward(ctx, input, operator, num_extra_dims=0, is_2d=False, dtype=torch.float32):
        extra_dims = input.size()[:num_extra_dims]
        if input.requires_grad:
            ctx.operator = operator
            ctx.extra_dims = extra_dims
            ctx.is_2d = is_2d
            ctx.dtype = dtype
        expected_ndim = (2 if is_2d else 3) + num_extra_dims
        assert input.ndim == expected_ndim, (
            f"Tomosipo autograd operator expected {expected_ndim} dimensions but got {input.ndim}.\n"
            "To add batch and channel dimensions set argument num_extra_dims=2\n"
            "To remove the first operator dimension set argument is_2d=True\n")
        output = input.new_empty(tuple(extra_dims) + tuple(operator.range_shape), dtype=dtype)
        if is_2d:
            input = torch.unsqueeze(input, dim=-3)
        _apply(operator, input, output, num_extra_dims)
        if is_2d:
            output = torch.squeeze(output, dim=-3)
        return output

    @staticmethod
    def backward(ctx, grad_output):
        operator = ctx.operator
        extra_dims = ctx.extra_dims
        is_2d = ctx.is_2d
        grad_input = grad_output.new_empty(tuple(extra_dims) + tuple(operator.domain_shape),
                                           dtype=ctx.dtype)
        if is_2d:
            grad_output = torch.unsqueeze(grad_output, dim=-3)
        _apply(operator.T, grad_output, grad_input, len(extra_dims))
        if is_2d:
            grad_input = torch.squeeze(grad_input, dim=-3)
        return grad_input, None, None, None, None


def to_autograd(operator, num_extra_dims=0, is_2d=False, dtype=None):
    """Autograd function of ``operator`` (forward ``A``, backward ``A.T``).

    ``dtype`` defaults to float32, as in tomosipo (outputs are always float32).
    ``dtype=torch.float64`` selects the FP64 kernels on a HIP device; forward and
    backward are then exact transposes to FP64 rounding, which finite-difference
    gradient checks need. This option has no counterpart in the reference.
    """
    dtype = torch.float32 if dtype is None else dtype
    if dtype not in (torch.float32, torch.float64):
        raise ValueError(f"dtype must be torch.float32 or torch.float64, got {dtype}")

    def f(x):
        return OperatorFunction.apply(x, operator, num_extra_dims, is_2d, dtype)

    return f


class AutogradOperator:
    """An operator with autograd support; additive operators are rejected."""

    def __init__(self, operator, num_extra_dims=0, is_2d=False):
        if operator.additive:
            raise ValueError("Additive operators are not supported")
        self.operator = operator
        self._fp_op = to_autograd(operator, num_extra_dims, is_2d)
        self._bp_op = to_autograd(operator.T, num_extra_dims, is_2d)
        self._transpose = BackprojectionOperator(self)

    def _fp(self, volume, out=None):
        if out is None:
            return self._fp_op(volume)
        out[...] = self._fp_op(volume)
        return out

    def _bp(self, projection, out=None):
        if out is None:
            return self._bp_op(projection)
        out[...] = self._bp_op(projection)
        return out

    def __call__(self, volume, out=None):
        return self._fp(volume, out)

    def transpose(self):
        return self._transpose

    @property
    def T(self):
        return self.transpose()

    @property
    def additive(self):
        return False

    @property
    def astra_compat_vg(self):
        return self.operator.astra_compat_vg

    @property
    def astra_compat_pg(self):
        return self.operator.astra_compat_pg

    @property
    def domain(self):
        return self.operator.domain

    @property
    def range(self):
        return self.operator.range

    @property
    def domain_shape(self):
        return self.operator.domain_shape

    @property
    def range_shape(self):
        return self.operator.range_shape


def autograd_operator(volume_geometry, projection_geometry, voxel_supersampling=1,
                      detector_supersampling=1, num_extra_dims=0, is_2d=False):
    from .Operator import operator

    op = operator(volume_geometry, projection_geometry, voxel_supersampling=voxel_supersampling,
                  detector_supersampling=detector_supersampling)
    return AutogradOperator(operator=op, num_extra_dims=num_extra_dims, is_2d=is_2d)
