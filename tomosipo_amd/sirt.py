"""SIRT on the HIP projector, with fused elementwise work and HIP-graph replay.

The iteration is the one tomosipo's documentation and benchmark use
(README.md:150-164, doc/intro/fast_reconstruction.rst:92-104,
notebooks/sirt_benchmark.py:116-139):

    R = 1 / A(1),  C = 1 / A.T(1), both clipped at 1 / ts.epsilon,
    x <- x + C * A.T(R * (y - A x)).

With ``fused=True`` one iteration is two library calls: the forward projector
writes ``R * (y - A x)`` directly and the backprojector adds ``C * A.T(.)``
into ``x``. With ``graph=True`` one iteration is captured once as a HIP graph
(torch.cuda.CUDAGraph on ROCm) and replayed, which removes the per-iteration
launch and Python overhead. All buffers, including the projector
workspaces, are allocated before capture.
"""
import torch

from . import _tsx
from .Operator import Operator


class SIRT:
    def __init__(self, A, y, x0=None, *, fused=True):
        if not isinstance(A, Operator):
            raise TypeError("SIRT needs a tomosipo_amd Operator")
        if A.projector.voxel_supersampling != 1:
            raise ValueError("the fused SIRT takes voxel_supersampling = 1 only")
        if not (isinstance(y, torch.Tensor) and y.is_cuda):
            raise ValueError("y must be a torch tensor on a HIP device")
        from . import epsilon

        self.A = A
        self.dev = y.device
        self.y = y.to(torch.float32).contiguous()
        with torch.cuda.device_of(self.y):
            R = A(torch.ones(A.domain_shape, device=self.dev))
            self.R = torch.clamp(1.0 / R, max=1.0 / epsilon)
            C = A.T(torch.ones(A.range_shape, device=self.dev))
            self.C = torch.clamp(1.0 / C, max=1.0 / epsilon)
            self.x = (torch.zeros(A.domain_shape, device=self.dev) if x0 is None
                      else x0.to(device=self.dev, dtype=torch.float32).clone().contiguous())
            self.tmp_y = torch.empty(A.range_shape, device=self.dev)
            self.tmp_x = torch.empty(A.domain_shape, device=self.dev)
            dev_index = self.dev.index if self.dev.index is not None else torch.cuda.current_device()
            self.plan = A.projector.plan(dev_index)
            nf = self.plan.workspace_size(_tsx.TSX_OP_FP)
            nb = self.plan.workspace_size(_tsx.TSX_OP_BP)
            self.ws_fp = torch.empty(max(nf, 1), dtype=torch.uint8, device=self.dev)
            self.ws_bp = torch.empty(max(nb, 1), dtype=torch.uint8, device=self.dev)
            self.nf, self.nb = nf, nb
        self.fused = fused
        self.graph = None

    def _iteration(self):
        stream = torch.cuda.current_stream(self.dev).cuda_stream
        if self.fused:
            self.plan.forward_residual(self.x.data_ptr(), self.y.data_ptr(), self.R.data_ptr(),
                                       self.tmp_y.data_ptr(), stream, workspace=self.ws_fp.data_ptr(),
                                       workspace_bytes=self.nf)
            self.plan.backward_scaled_add(self.x.data_ptr(), self.tmp_y.data_ptr(),
                                          self.C.data_ptr(), stream, workspace=self.ws_bp.data_ptr(),
                                          workspace_bytes=self.nb)
        else:
            self.plan.forward(self.x.data_ptr(), self.tmp_y.data_ptr(), False, stream,
                              workspace=self.ws_fp.data_ptr(), workspace_bytes=self.nf)
            torch.sub(self.y, self.tmp_y, out=self.tmp_y)
            self.tmp_y.mul_(self.R)
            self.plan.backward(self.tmp_x.data_ptr(), self.tmp_y.data_ptr(), False, stream,
                               workspace=self.ws_bp.data_ptr(), workspace_bytes=self.nb)
            self.x.addcmul_(self.C, self.tmp_x)

    def capture(self):
        """Capture one iteration as a graph (runs one eager iteration first as warm-up)."""
        with torch.cuda.device_of(self.x):
            side = torch.cuda.Stream(device=self.dev)
            side.wait_stream(torch.cuda.current_stream(self.dev))
            with torch.cuda.stream(side):
                x_save = self.x.clone()
                self._iteration()
                self.x.copy_(x_save)
            torch.cuda.current_stream(self.dev).wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._iteration()
            self.graph = g
        return g

    def run(self, num_iterations, graph=True):
        with torch.cuda.device_of(self.x):
            if graph:
                if self.graph is None:
                    x_save = self.x.clone()
                    self.capture()
                    # capture does not execute; the warm-up iteration was undone
                    self.x.copy_(x_save)
                for _ in range(num_iterations):
                    self.graph.replay()
            else:
                for _ in range(num_iterations):
                    self._iteration()
        return self.x


def sirt(A, y, num_iterations, x0=None, graph=True, fused=True):
    """``num_iterations`` SIRT iterations of ``A`` for data ``y`` (device tensor)."""
    return SIRT(A, y, x0, fused=fused).run(num_iterations, graph=graph)
