"""Legacy composite projection ts.astra.project / forward / backward (GPU).

tomosipo's astra.py:221-371 projects every volume dataset onto every
projection dataset (ASTRA's composite call, astra.py:279-287): forward
projection datasets become the sum over the volumes' projections (SET) or
gain it (ADD); backprojection likewise into the volumes. Checked against
per-pair operator calls (bitwise for one pair, 1e-6 for sums, whose
summation order differs) and against the FP64 oracle (1e-5), on device
tensors (fused launches) and numpy arrays (staged per pair). The reference's
own test (tests/test_astra.py:46-60) runs forward and backward with
detector and voxel supersampling 1 and 2; here each option is also checked:
detector supersampling against the oracle, voxel supersampling k of the
backprojection as the exact transpose of projecting the k-times refined
volume.
"""
import numpy as np
import pytest

import oracle
import tomosipo_amd as ts
from tests.conftest import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda"


def _geoms():
    vgs = [ts.volume(shape=24, size=1, pos=(0.0, 0.0, -0.3)),
           ts.volume(shape=(20, 24, 28), size=(0.8, 1.0, 1.1), pos=(0.1, 0.2, 0.25))]
    pgs = [ts.cone(angles=30, shape=(32, 48), size=(1.6, 2.4), cone_angle=0.5),
           ts.parallel(angles=25, shape=(30, 40), size=(1.5, 2.0))]
    return vgs, pgs


@pytest.mark.parametrize("on_device", [True, False])
def test_composite_forward_and_backward(on_device):
    vgs, pgs = _geoms()
    rng = np.random.default_rng(3)
    xs = [rng.random(vg.shape, dtype=np.float32) for vg in vgs]
    ys = [rng.random(ts.operator(vgs[0], pg).range_shape, dtype=np.float32) for pg in pgs]
    conv = (lambda a: torch.from_numpy(a).to(DEV)) if on_device else (lambda a: a.copy())
    vd = [ts.data(vg, conv(x)) for vg, x in zip(vgs, xs)]
    pd = [ts.data(pg, conv(np.zeros_like(y))) for pg, y in zip(pgs, ys)]
    print(f"[legacy] forward, device={on_device}", flush=True)
    ts.astra.forward(*vd, *pd)
    for j, pg in enumerate(pgs):
        ref = sum(ts.operator(vg, pg)(x) for vg, x in zip(vgs, xs))
        got = pd[j].data.cpu().numpy() if on_device else pd[j].data
        assert rel_l2(got, ref) <= 1e-6, j
        o = sum(oracle.fp(oracle.oracle_geometry(vg, pg), x.astype(np.float64)) for vg, x in zip(vgs, xs))
        assert rel_l2(got, o) <= 1e-5, j
    # additive: a second forward adds the same sums
    before = [d.data.clone() if on_device else d.data.copy() for d in pd]
    ts.astra.forward(*vd, *pd, additive=True)
    for d, b in zip(pd, before):
        got = d.data.cpu().numpy() if on_device else d.data
        bb = b.cpu().numpy() if on_device else b
        assert rel_l2(got, 2 * bb) <= 1e-6
    # backward: every volume becomes the sum of the datasets' backprojections
    pd = [ts.data(pg, conv(y)) for pg, y in zip(pgs, ys)]
    vd = [ts.data(vg, conv(np.full(vg.shape, 5.0, dtype=np.float32))) for vg in vgs]
    print(f"[legacy] backward, device={on_device}", flush=True)
    ts.astra.backward(*pd, *vd)
    for i, vg in enumerate(vgs):
        ref = sum(ts.operator(vg, pg).T(y) for pg, y in zip(pgs, ys))
        got = vd[i].data.cpu().numpy() if on_device else vd[i].data
        assert rel_l2(got, ref) <= 1e-6, i
        o = sum(oracle.bp(oracle.oracle_geometry(vg, pg), y.astype(np.float64)) for pg, y in zip(pgs, ys))
        assert rel_l2(got, o) <= 1e-5, i


def test_single_pair_equals_operator_bitwise():
    vg = ts.volume(shape=24, size=1)
    pg = ts.cone(angles=30, shape=(32, 48), size=(1.6, 2.4), cone_angle=0.5)
    A = ts.operator(vg, pg)
    x = torch.rand(A.domain_shape, device=DEV)
    pd = ts.data(pg, torch.zeros(A.range_shape, device=DEV))
    ts.astra.forward(ts.data(vg, x), pd, projector=A.astra_projector)
    assert torch.equal(pd.data, A(x))
    vd = ts.data(vg, torch.zeros(A.domain_shape, device=DEV))
    ts.astra.backward(vd, pd)
    assert torch.equal(vd.data, A.T(pd.data))


def test_supersampling_options():
    """The reference's test_forward_backward (tests/test_astra.py:46-60) with
    values checked: detector supersampling 2 against the oracle; voxel
    supersampling 2 of the backprojection is the transpose of projecting the
    refined volume (oracle on the 2x refined grid), and does not change the
    forward projection."""
    pg = ts.cone(size=np.sqrt(2), cone_angle=1 / 2, shape=10, angles=12)
    vg = ts.volume(shape=10)
    x = ts.phantom.ellipsoids(10, 4, seed=1)
    rng = np.random.default_rng(2)
    y = rng.random((10, 12, 10), dtype=np.float32)
    for dss, vss in ((1, 1), (2, 2), (1, 2), (2, 1)):
        print(f"[legacy] detector_supersampling={dss} voxel_supersampling={vss}", flush=True)
        pd = ts.data(pg, np.zeros((10, 12, 10), dtype=np.float32))
        ts.astra.forward(ts.data(vg, x.copy()), pd, detector_supersampling=dss, voxel_supersampling=vss)
        g = oracle.oracle_geometry(vg, pg, det_supersampling=dss)
        assert rel_l2(pd.data, oracle.fp(g, x.astype(np.float64))) <= 1e-5
        vd = ts.data(vg, np.zeros((10, 10, 10), dtype=np.float32))
        ts.astra.backward(vd, ts.data(pg, y.copy()), detector_supersampling=dss, voxel_supersampling=vss)
        fine = ts.volume(shape=10 * vss, size=vg.size, pos=tuple(vg.pos[0]))
        gf = oracle.oracle_geometry(fine, pg, det_supersampling=dss)
        bf = oracle.bp(gf, y.astype(np.float64))
        ref = bf.reshape(10, vss, 10, vss, 10, vss).sum(axis=(1, 3, 5))
        assert rel_l2(vd.data, ref) <= 1e-5


def test_operator_voxel_supersampling_is_an_exact_transpose_pair():
    from tomosipo_amd.torch_support import to_autograd

    vg = ts.volume(shape=12, size=1)
    pg = ts.parallel(angles=10, shape=(16, 18), size=(1.2, 1.4))
    A = ts.operator(vg, pg, voxel_supersampling=3)
    x = torch.rand(A.domain_shape, device=DEV, dtype=torch.float64)
    y = torch.rand(A.range_shape, device=DEV, dtype=torch.float64)
    f = to_autograd(A, dtype=torch.float64)
    g = to_autograd(A.T, dtype=torch.float64)
    Ax, ATy = f(x), g(y)
    adj = abs(float((Ax * y).sum() - (x * ATy).sum())) / float(torch.linalg.norm(Ax) * torch.linalg.norm(y))
    assert adj < 1e-12
    fine = ts.volume(shape=36, size=1)
    go = oracle.oracle_geometry(fine, pg)
    xr = x.cpu().numpy().repeat(3, 0).repeat(3, 1).repeat(3, 2)
    assert rel_l2(Ax.cpu().numpy(), oracle.fp(go, xr)) < 1e-12
    assert rel_l2(A(x.float()).cpu().numpy(), oracle.fp(go, xr)) <= 1e-5
