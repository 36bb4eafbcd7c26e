"""The benchmark configurations at their own sizes, checked against the oracle (GPU).

* SIRT, cone 256^3, 360 angles, 256 x 384 detector: per-operation parity on
  every 30th angle at the full volume size; 50 iterations replayed as a HIP
  graph bitwise equal to the eager loop; the final iterate of 50 iterations
  against the FP64 oracle's SIRT on the same geometry scaled to 64^3 / 90
  angles (printed and bounded).
* autograd, batch 8 x 128^3, parallel 180 angles, 128 x 192: the gradient of
  ||f(x) - y||^2 equals the oracle's 2 A^T (A x - y) per batch element
  (relative L2 <= 1e-5), as tomosipo's tests/test_torch_support.py:91-100
  assert grad = A^T A x.
* the throughput workload exactly as bench.py runs it (16 x cone_vec 512^3,
  720 angles, 512 x 768, one fused forward and one fused backprojection
  launch): every instance equal to its single-instance launch bitwise, and
  instance 0 against the oracle on every 90th angle (forward: those angles of
  its output; backprojection: the fused launch with instance 0's sinogram
  zero except on those angles).
* the drop-in ctypes binding given in the repository's integration notes,
  run as written (only the library path is filled in), SET and ADD, FP and BP, against the oracle; and its
  FDK binding against ts.fdk.

Tolerances: FP and BP relative L2 <= 1e-5 against the FP64 oracle.
"""
import os
import re
import types

import numpy as np
import pytest

import oracle
import tomosipo_amd as ts
from tests.conftest import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEV = "cuda"
TOL = 1e-5


def _c2_geometry(n=256, angles=360):
    vg = ts.volume(shape=n, size=1)
    pg = ts.cone(angles=angles, shape=(n, 3 * n // 2), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    return vg, pg


def _c2_phantom(n, seed):
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    x = torch.randn((n,) * 3, generator=g, device=DEV) * 0.01
    x[n // 4:3 * n // 4, n // 4:3 * n // 4, n // 4:3 * n // 4] += 1.0
    return x


def test_c2_per_operation_parity_on_an_angle_subset():
    vg, pg = _c2_geometry()
    sub = pg.to_vec()[::30]
    print(f"[c2] 256^3 cone, {sub.num_angles} of 360 angles", flush=True)
    A = ts.operator(vg, sub)
    x = _c2_phantom(256, 3000)
    y = A(x)
    g = oracle.oracle_geometry(vg, sub)
    e_fp = rel_l2(y.cpu().numpy(), oracle.fp(g, x.double().cpu().numpy()))
    yh = y.double().cpu().numpy()
    e_bp = rel_l2(A.T(y).cpu().numpy(), oracle.bp(g, yh))
    print(f"[c2]   fp {e_fp:.2e}  bp {e_bp:.2e}", flush=True)
    assert e_fp <= TOL and e_bp <= TOL


def test_c2_sirt50_graph_replay_equals_eager():
    from tomosipo_amd.sirt import SIRT

    vg, pg = _c2_geometry()
    A = ts.operator(vg, pg)
    y = A(_c2_phantom(256, 3000))
    eager, graphed = SIRT(A, y), SIRT(A, y)
    print("[c2] SIRT-50 eager vs graph", flush=True)
    for k in range(5):
        eager.run(10, graph=False)
        graphed.run(10, graph=True)
        torch.cuda.synchronize()
        _ = (eager.x * 2).sum().item()  # unrelated eager work between replays
        assert torch.equal(eager.x, graphed.x), f"after {10 * (k + 1)} iterations"
    r = (torch.linalg.norm(A(eager.x) - y) / torch.linalg.norm(y)).item()
    print(f"[c2]   relative residual after 50 iterations {r:.3e}", flush=True)
    assert r < 0.2


def _oracle_sirt(g, y, iters):
    R = oracle.fp(g, np.ones(g.vol_shape))
    R = np.minimum(1 / np.maximum(R, 1e-300), 1e8)
    C = oracle.bp(g, np.ones(g.sino_shape))
    C = np.minimum(1 / np.maximum(C, 1e-300), 1e8)
    x = np.zeros(g.vol_shape)
    for _ in range(iters):
        x += C * oracle.bp(g, R * (y - oracle.fp(g, x)))
    return x


def test_c2_sirt50_final_iterate_against_oracle_sirt():
    """The SIRT-50 geometry scaled to 64^3 / 90 angles / 64 x 96 (the FP64
    oracle takes a few seconds there); graph replay with the fused epilogues."""
    from tomosipo_amd.sirt import sirt

    vg, pg = _c2_geometry(64, 90)
    A = ts.operator(vg, pg)
    g = oracle.oracle_geometry(vg, pg)
    x = _c2_phantom(64, 3001).double().cpu().numpy()
    y = oracle.fp(g, x)
    x_gpu = sirt(A, torch.from_numpy(y.astype(np.float32)).to(DEV), 50).cpu().numpy()
    x_ref = _oracle_sirt(g, y, 50)
    err = rel_l2(x_gpu, x_ref)
    print(f"[c2] SIRT-50 at 64^3 / 90 angles: final iterate rel L2 vs FP64 oracle {err:.2e}", flush=True)
    assert err < 1e-4


def test_c4_batched_gradient_matches_oracle():
    from tomosipo_amd.torch_support import to_autograd

    vg = ts.volume(shape=128, size=1)
    pg = ts.parallel(angles=180, shape=(128, 192), size=(1, 1.5))
    A = ts.operator(vg, pg)
    g = oracle.oracle_geometry(vg, pg)
    print("[c4] batch 8 x 128^3, parallel 180 angles", flush=True)
    xs = np.stack([ts.phantom.ellipsoids(128, 20, seed=4000 + s) for s in range(8)])
    ys = np.stack([oracle.fp(g, ts.phantom.ellipsoids(128, 20, seed=4100 + s)).astype(np.float32)
                   for s in range(8)])
    x = torch.tensor(xs, device=DEV, requires_grad=True)
    y = torch.tensor(ys, device=DEV)
    loss = ((to_autograd(A, num_extra_dims=1)(x) - y) ** 2).sum()
    loss.backward()
    worst = 0.0
    for k in range(8):
        ref = 2 * oracle.bp(g, oracle.fp(g, xs[k].astype(np.float64)) - ys[k])
        worst = max(worst, rel_l2(x.grad[k].cpu().numpy(), ref))
    print(f"[c4]   worst gradient rel L2 {worst:.2e}", flush=True)
    assert worst <= TOL


def test_c3_as_benchmarked():
    import bench
    from tomosipo_amd.fused import fused_forward, fused_backward

    dev = torch.device(DEV, 0)
    ops, xs, geos = [], [], []
    for i in range(16):
        vg, pg = bench.instance_geometry(i, ts)
        geos.append((vg, pg))
        ops.append(ts.operator(vg, pg))
        xs.append(bench.phantom(i, torch, dev))
    ys = [torch.empty(A.range_shape, device=dev) for A in ops]
    bs = [torch.empty(A.domain_shape, device=dev) for A in ops]
    print("[c3] 16 instances: fused FP, fused BP", flush=True)
    fused_forward(ops, xs, out=ys)
    fused_backward(ops, ys, out=bs)
    torch.cuda.synchronize()
    for k, A in enumerate(ops):
        assert torch.equal(ys[k], A(xs[k])), f"forward, instance {k}"
        assert torch.equal(bs[k], A.T(ys[k])), f"backprojection, instance {k}"
    print("[c3]   every instance equals its single launch", flush=True)
    # instance 0 against the oracle on every 90th angle
    vg, pg = geos[0]
    sub = pg[::90]
    g = oracle.oracle_geometry(vg, sub)
    x0 = xs[0].double().cpu().numpy()
    e_fp = rel_l2(ys[0][:, ::90, :].cpu().numpy(), oracle.fp(g, x0))
    mask = torch.zeros(720, device=dev)
    mask[::90] = 1.0
    ys[0] *= mask[None, :, None]
    fused_backward(ops, ys, out=bs)
    e_bp = rel_l2(bs[0].cpu().numpy(), oracle.bp(g, ys[0][:, ::90, :].double().cpu().numpy()))
    print(f"[c3]   instance 0, 8 angles: fp {e_fp:.2e}  bp {e_bp:.2e}", flush=True)
    assert e_fp <= TOL and e_bp <= TOL


def _integration_snippets():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```python\n(.*?)```", text, flags=re.S)
    assert len(blocks) >= 2
    from tomosipo_amd import _tsx

    ns = {}
    exec(blocks[0].replace('"libtsx.so"', repr(_tsx.LIB_PATH)), ns)
    exec(blocks[1], ns)
    return ns


@pytest.mark.parametrize("kind", ["cone", "parallel"])
def test_integration_binding_direct_fpbp3d(kind):
    ns = _integration_snippets()
    vg = ts.volume(shape=48, size=1)
    if kind == "cone":
        pg = ts.cone(angles=40, shape=(48, 72), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    else:
        pg = ts.parallel(angles=40, shape=(48, 72), size=(1, 1.5))
    proj_id = ns["create_astra_projector"](vg, pg)
    g = oracle.oracle_geometry(vg, pg)
    x = ts.phantom.ellipsoids(48, 8, seed=21)
    xd = torch.from_numpy(x).to(DEV)
    y = torch.empty(g.sino_shape, device=DEV)
    link = types.SimpleNamespace
    print(f"[binding] {kind}: tsx_direct_fpbp3d FP/BP, SET/ADD", flush=True)
    ns["direct_project"](proj_id, link(data=xd), link(data=y), forward=True)
    y_ref = oracle.fp(g, x.astype(np.float64))
    assert rel_l2(y.cpu().numpy(), y_ref) <= TOL
    ns["direct_project"](proj_id, link(data=xd), link(data=y), forward=True, additive=True)
    assert rel_l2(y.cpu().numpy(), 2 * y_ref) <= TOL
    yin = torch.from_numpy(y_ref.astype(np.float32)).to(DEV)
    b = torch.full(g.vol_shape, 7.0, device=DEV)
    ns["direct_project"](proj_id, link(data=b), link(data=yin), forward=False)
    b_ref = oracle.bp(g, y_ref.astype(np.float32).astype(np.float64))
    assert rel_l2(b.cpu().numpy(), b_ref) <= TOL
    ns["direct_project"](proj_id, link(data=b), link(data=yin), forward=False, additive=True)
    assert rel_l2(b.cpu().numpy(), 2 * b_ref) <= TOL
    # the binding's results are the library's own (SET then ADD of the same
    # backprojection doubles it exactly)
    A = ts.operator(vg, pg)
    assert torch.equal(b, 2 * A.T(yin))


def test_integration_binding_fdk():
    ns = _integration_snippets()
    vg = ts.volume(shape=48, size=1)
    pg = ts.cone(angles=60, shape=(48, 72), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    A = ts.operator(vg, pg)
    y = A(torch.from_numpy(ts.phantom.ellipsoids(48, 8, seed=22)).to(DEV))
    vol = ts.data(vg, torch.zeros(A.domain_shape, device=DEV))
    ns["fdk"](vol, ts.data(pg, y))
    ref = ts.fdk(A, y)
    torch.cuda.synchronize()
    assert torch.equal(vol.data, ref)
