"""Checks at the full size of the throughput workload (cone_vec 512^3, 720 angles,
512 x 768 detector, helical source shift +-0.25 and jitter 1e-3).

* parity with the FP64 oracle on every 90th angle (8 angles of the full-size
  geometry): FP and BP relative L2 <= 1e-5;
* size-independent properties on all 720 angles: matched adjoint
  |<Ax, y> - <x, A^T y>| / (|Ax| |y|) <= 1e-5, linearity (rel. L2 <= 1e-5),
  and a fused two-instance launch bitwise equal to single launches.
"""
import numpy as np
import pytest

import oracle
import tomosipo_amd as ts
from tests.conftest import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

N, NA = 512, 720


def _instance(i):
    base = ts.cone(angles=NA, shape=(512, 768), size=(np.sqrt(2), 1.5 * np.sqrt(2)),
                   cone_angle=0.5).to_vec()
    src = base.src_pos
    src[:, 0] += np.linspace(-0.25, 0.25, NA)
    src += np.random.default_rng(1000 + i).normal(0.0, 1e-3, size=src.shape)
    pg = ts.cone_vec(shape=base.det_shape, src_pos=src, det_pos=base.det_pos, det_v=base.det_v,
                     det_u=base.det_u)
    return ts.volume(shape=N, size=1), pg


def _phantom(seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randn((N,) * 3, generator=g, device="cuda") * 0.01
    x[128:384, 128:384, 128:384] += 1.0
    return x


def test_fullsize_parity_on_an_angle_subset():
    vg, pg = _instance(0)
    sub = pg[::90]
    print(f"[fullsize] parity on {sub.num_angles} angles", flush=True)
    A = ts.operator(vg, sub)
    x = _phantom(2000)
    y = A(x)
    g = oracle.oracle_geometry(vg, sub)
    xh = x.double().cpu().numpy()
    assert rel_l2(y.cpu().numpy(), oracle.fp(g, xh)) <= 1e-5
    yh = y.cpu().numpy()
    b = A.T(y)
    assert rel_l2(b.cpu().numpy(), oracle.bp(g, yh.astype(np.float64))) <= 1e-5


def test_fullsize_adjoint_linearity_and_fused():
    from tomosipo_amd.fused import fused_forward, fused_backward

    vg, pg = _instance(0)
    A = ts.operator(vg, pg)
    print("[fullsize] adjoint / linearity / fused on 720 angles", flush=True)
    x = _phantom(2000)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    y = torch.rand(A.range_shape, generator=g, device="cuda")
    Ax = A(x)
    ATy = A.T(y)
    lhs = torch.sum(Ax.double() * y.double())
    rhs = torch.sum(x.double() * ATy.double())
    adj = abs((lhs - rhs).item()) / (torch.linalg.norm(Ax.double()) * torch.linalg.norm(y.double())).item()
    print(f"[fullsize] adjoint {adj:.3e}", flush=True)
    assert adj <= 1e-5
    x2 = _phantom(2001)
    lin = A(x + 2 * x2) - (Ax + 2 * A(x2))
    assert (torch.linalg.norm(lin) / torch.linalg.norm(Ax)).item() <= 1e-5
    del lin
    vg1, pg1 = _instance(1)
    B = ts.operator(vg1, pg1)
    ys = fused_forward([A, B], [x, x2])
    assert torch.equal(ys[0], Ax) and torch.equal(ys[1], B(x2))
    bs = fused_backward([A, B], [y, ys[1]])
    assert torch.equal(bs[0], ATy) and torch.equal(bs[1], B.T(ys[1]))
