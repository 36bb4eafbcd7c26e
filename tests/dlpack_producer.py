"""A minimal DLPack producer for the link tests: an array type of its own
(not a torch tensor, not a numpy array) that exports its memory through
``__dlpack__`` / ``__dlpack_device__``, with a module-level ``from_dlpack``,
as CuPy's arrays do. It stands in for CuPy, which is not installed here."""
import torch
import torch.utils.dlpack


class Array:
    def __init__(self, t):
        self._t = t

    def __dlpack__(self, stream=None, **kw):
        return self._t.__dlpack__(stream=stream) if stream is not None else self._t.__dlpack__()

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()

    @property
    def shape(self):
        return tuple(self._t.shape)

    def to_torch(self):
        return self._t


def from_dlpack(x):
    return Array(torch.utils.dlpack.from_dlpack(x))
