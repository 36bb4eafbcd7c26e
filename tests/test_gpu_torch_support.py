"""The torch link and autograd operators, after tests/test_torch_support.py.

Host tensors go through the same HIP kernels as device tensors (staged
copies), so every case runs on both devices. Results of the plain and the
autograd operator come from the same kernels and must be equal bitwise.
"""
import pytest

import tomosipo_amd as ts
from tomosipo_amd.torch_support import to_autograd, AutogradOperator, autograd_operator

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

N = 64
N_angles = N * 3 // 2
M = N * 3 // 2
devices = ["cpu", "cuda"]


@pytest.mark.parametrize("device", devices)
def test_link_new_tensors(device):
    vg = ts.volume(shape=N)
    pg = ts.parallel(angles=N_angles, shape=(N, M))
    A = ts.operator(vg, pg)
    vol = torch.zeros(A.domain_shape, device=device)
    link = ts.link(vg, vol)
    assert torch.equal(vol, link.new_zeros(A.domain_shape).data)
    assert torch.equal(vol + 1.0, link.new_full(A.domain_shape, 1.0).data)
    assert link.shape == link.new_empty(A.domain_shape).shape
    for t in (link.new_zeros(A.domain_shape), link.new_full(A.domain_shape, 1.0),
              link.new_empty(A.domain_shape)):
        assert t.data.device == vol.device


@pytest.mark.parametrize("device", devices)
def test_fp_bp(device):
    vg = ts.volume(shape=N)
    pg = ts.parallel(angles=N_angles, shape=(N, M))
    A = ts.operator(vg, pg)
    vol = torch.ones(A.domain_shape, device=device)
    sino = A(vol)
    bp = A.T(sino)
    assert sino.device == vol.device and bp.device == vol.device
    assert 1.0 < sino.sum() and 1.0 < bp.sum()
    if device == "cpu":
        assert torch.equal(sino, A(vol.cuda()).cpu())
        assert torch.equal(bp, A.T(sino.cuda()).cpu())


@pytest.mark.parametrize("device", devices)
def test_float64_input(device):
    vg = ts.volume(shape=(1, N, N))
    pg = ts.parallel(angles=N_angles, shape=(1, M))
    A = ts.operator(vg, pg)
    x = torch.ones((1, N, N), dtype=torch.float64, device=device)
    with pytest.warns(UserWarning):
        y = A(x)
    y_ag = to_autograd(A)(x)
    assert y.dtype == y_ag.dtype == torch.float32
    assert torch.equal(y, y_ag)


@pytest.mark.parametrize("device", devices)
def test_autograd(device):
    vg = ts.volume(shape=(1, N, N))
    pg = ts.parallel(angles=N_angles, shape=(1, M))
    A = ts.operator(vg, pg)
    x = torch.ones(A.domain_shape, dtype=torch.float32, device=device, requires_grad=True)
    y = to_autograd(A)(x)
    y.backward(y)
    assert torch.equal(x.grad, A.T(A(x.detach())))


@pytest.mark.parametrize("device", devices)
def test_autograd_shape(device):
    vg = ts.volume(shape=(1, N, N))
    pg = ts.parallel(angles=N_angles, shape=(1, M))
    A = ts.operator(vg, pg)
    A_ag = to_autograd(A, num_extra_dims=2)
    x1 = torch.ones(1, 1, *A.domain_shape, device=device, requires_grad=True)
    x2 = torch.ones(2, 3, *A.domain_shape, device=device, requires_grad=True)
    y1, y2 = A_ag(x1), A_ag(x2)
    y1.backward(y1)
    y2.backward(y2)
    assert tuple(y1.shape) == (1, 1, *A.range_shape)
    assert tuple(y2.shape) == (2, 3, *A.range_shape)
    assert tuple(x1.grad.shape) == (1, 1, *A.domain_shape)
    assert tuple(x2.grad.shape) == (2, 3, *A.domain_shape)
    assert torch.equal(x2.grad[1, 2], A.T(A(x2.detach()[1, 2])))


@pytest.mark.parametrize("device", devices)
def test_autograd_operator(device):
    vg = ts.volume(shape=(1, N, N))
    pg = ts.parallel(angles=N_angles, shape=(1, M))
    A = ts.operator(vg, pg)
    x = torch.zeros(A.domain_shape, dtype=torch.float32, device=device)
    x[:, 10:-10, 10:-10] = 1.0
    x[:, 20:-20, 20:-20] = 0.0
    for A_ag in [AutogradOperator(A), autograd_operator(vg, pg)]:
        y = A(x)
        b = A.T(y)
        y2 = A(b)
        y_ag = A_ag(x)
        b_ag = A_ag.T(y_ag)
        y2_ag = A_ag(b_ag)
        assert torch.equal(y, y_ag) and torch.equal(b, b_ag) and torch.equal(y2, y2_ag)
        y_ag[...] = 0
        b_ag[...] = 0
        y2_ag[...] = 0
        A_ag(x, out=y_ag)
        A_ag.T(y_ag, out=b_ag)
        A_ag(b_ag, out=y2_ag)
        assert torch.equal(y, y_ag) and torch.equal(b, b_ag) and torch.equal(y2, y2_ag)


@pytest.mark.parametrize("device", devices)
def test_autograd_operator_backward(device):
    vg = ts.volume(shape=(1, N, N))
    pg = ts.parallel(angles=N_angles, shape=(1, M))
    A = ts.operator(vg, pg)
    A_ag = AutogradOperator(A)
    x = torch.ones(A_ag.domain_shape, dtype=torch.float32, device=device, requires_grad=True)
    y = A_ag(x)
    y.backward(y)
    assert torch.equal(x.grad, A.T(A(x.detach())))
    p = torch.ones(A_ag.range_shape, dtype=torch.float32, device=device, requires_grad=True)
    q = A_ag.T(p)
    q.backward(q)
    assert torch.equal(p.grad, A(A.T(p.detach())))
