"""DLPack link: arrays of other libraries on the HIP device (tomosipo ``links/cupy.py:25-159``).

tomosipo links CuPy arrays with a dedicated backend that hands ASTRA a raw
device pointer. Here any array that implements the DLPack protocol
(``__dlpack__`` / ``__dlpack_device__``: CuPy, JAX, and others) is linked:
the projector reads and writes it through a zero-copy torch view
(``torch.utils.dlpack.from_dlpack``), on the current stream of its device.
Arrays this link creates (``new_zeros``, ``clone``, outputs of ``A(x)``) are of
the input's own type when its module or a package above it has a
``from_dlpack`` function (as CuPy's does), torch tensors otherwise.

As for CuPy in tomosipo, coercions copy with a warning (float32,
C-contiguous) and a wrong shape raises ``ValueError``. Unlike tomosipo's CuPy
link, a DLPack array is compatible with a torch tensor on the same device.
numpy arrays and torch tensors keep their own links, which are registered
first.
"""
import importlib
import warnings
from contextlib import contextmanager

import torch
import torch.utils.dlpack

from .base import Link, backends
from .numpy import NumpyLink
from .torch import TorchLink

# DLPack device types (dlpack.h): CPU 1, CUDA 2, ROCm 10
_DL_CPU, _DL_CUDA, _DL_ROCM = 1, 2, 10


def _producer_from_dlpack(obj):
    """The ``from_dlpack`` of the array's own module or of the nearest package
    above it (cupy.from_dlpack for CuPy arrays), or None."""
    parts = type(obj).__module__.split(".")
    for n in range(len(parts), 0, -1):
        try:
            mod = importlib.import_module(".".join(parts[:n]))
        except ImportError:
            continue
        fn = getattr(mod, "from_dlpack", None)
        if callable(fn):
            return fn
    return None


class DLPackLink(Link):
    def __init__(self, shape, initial_value, _view=None):
        super().__init__(shape, initial_value)
        if not DLPackLink.__accepts__(initial_value):
            raise ValueError(f"Expected initial_value to implement the DLPack protocol. "
                             f"Got {initial_value.__class__}")
        view = _view if _view is not None else torch.utils.dlpack.from_dlpack(initial_value)
        if tuple(shape) != tuple(view.shape):
            raise ValueError(f"Expected initial_value with shape {tuple(shape)}. Got {tuple(view.shape)}")
        self._wrap = _producer_from_dlpack(initial_value)
        if view.dtype != torch.float32 or not view.is_contiguous():
            if view.dtype != torch.float32:
                warnings.warn(
                    f"The parameter initial_value is of type {view.dtype}; expected float32. "
                    "The type has been automatically converted. Convert it to float32 first "
                    "to inhibit this warning. ")
            if not view.is_contiguous():
                warnings.warn(
                    "The parameter initial_value should be C_CONTIGUOUS. It has been automatically "
                    "made contiguous. Make it contiguous first to inhibit this warning. ")
            view = view.to(dtype=torch.float32).contiguous()
            initial_value = self._wrap(view) if self._wrap is not None else view
        self._data = initial_value
        self._view = view

    @staticmethod
    def __accepts__(initial_value):
        if isinstance(initial_value, torch.Tensor):
            return False
        return hasattr(initial_value, "__dlpack__") and hasattr(initial_value, "__dlpack_device__")

    def __compatible_with__(self, other):
        if isinstance(other, (DLPackLink, TorchLink)):
            return self.tensor.device == other.tensor.device if isinstance(other, DLPackLink) \
                else self.tensor.device == other.data.device
        if isinstance(other, NumpyLink):
            return not self.tensor.is_cuda
        return NotImplemented

    @property
    def tensor(self):
        """Zero-copy torch view of the linked array (what the projector uses)."""
        return self._view

    @property
    def data(self):
        return self._data

    @data.setter
    def data(self, val):
        Link.data.fset(self, val)

    @contextmanager
    def context(self):
        if self._view.is_cuda:
            with torch.cuda.device_of(self._view):
                yield
        else:
            yield

    def _new(self, shape, t):
        obj = self._wrap(t) if self._wrap is not None else t
        return DLPackLink(shape, obj, _view=t) if obj is not t else TorchLink(shape, t)

    def new_zeros(self, shape):
        return self._new(shape, self._view.new_zeros(shape))

    def new_full(self, shape, value):
        return self._new(shape, self._view.new_full(shape, value))

    def new_empty(self, shape):
        return self._new(shape, self._view.new_empty(shape))

    def clone(self):
        return self._new(self._view.shape, self._view.clone())


backends.append(DLPackLink)
