"""Forward projection in z-slabs and backprojection in ranges of angles (GPU).

The fast kernels address their scratch layouts with 32-bit byte offsets. A
volume whose padded layouts reach 4 GiB is forward-projected slab by slab and
projections whose transposed copy reaches 4 GiB are backprojected in ranges of
angles, the chunks' sums added in the output. TSX_CHUNK_BYTES lowers the
4 GiB bound when a plan is created, so that the small cases below take the
same chunked paths as the 1024^3 ones.

Tolerances: FP32 kernels against the FP64 oracle, relative L2 <= 1e-5 (as in
the other parity tests); chunked against unchunked results (the same kernels,
extra float32 additions), relative L2 <= 1e-6; matched adjoint
|<Ax, y> - <x, A^T y>| / (|Ax| |y|) <= 1e-5.
"""
import numpy as np
import pytest

import oracle
import tomosipo_amd as ts
from tests.conftest import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DEV = "cuda"
TOL = 1e-5


def _rel(a, b):
    return (torch.linalg.norm((a - b).double()) / torch.linalg.norm(b.double())).item()


def _chunked_operator(monkeypatch, vg, pg, limit):
    """An operator whose device plan is created with TSX_CHUNK_BYTES = limit."""
    monkeypatch.setenv("TSX_CHUNK_BYTES", str(limit))
    A = ts.operator(vg, pg)
    plan = A.projector.plan(0)
    monkeypatch.delenv("TSX_CHUNK_BYTES")
    return A, plan.info()


def _cases():
    vg = ts.volume(shape=48, size=1)
    cone = ts.cone(angles=40, shape=(48, 72), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    par = ts.parallel(angles=40, shape=(48, 72), size=(1, 1.5))
    # rotated about y by 1.2 rad: some angles step along z (slabs cut their
    # slice range) and the backprojector is the general one
    tilted = ts.rotate(pos=0, axis=(0, 1, 0), angles=1.2) * cone.to_vec()
    # turned 90 degrees about y: both projectors run on the plan with x and z
    # swapped, whose slabs and angle ranges are cut by the same bound
    turned = ts.rotate(pos=0, axis=(0, 1, 0), angles=np.pi / 2) * cone.to_vec()
    return {"cone": (vg, cone), "parallel": (vg, par), "tilted": (vg, tilted), "turned": (vg, turned)}


@pytest.mark.parametrize("case", ["cone", "parallel", "tilted", "turned"])
def test_slabs_and_angle_ranges_match_oracle(monkeypatch, case):
    vg, pg = _cases()[case]
    A, info = _chunked_operator(monkeypatch, vg, pg, 400000)
    print(f"[chunked] {case}: {info['fp_slabs']} slabs, bp {info['bp_kernel']} "
          f"{info['bp_launch_angles']} angles per launch, axis swap {info['fp_axis_swap']}", flush=True)
    assert info["fp_kernel"] == "padded" and info["fp_slabs"] == 3
    assert info["fp_axis_swap"] == info["bp_axis_swap"] == (1 if case == "turned" else 0)
    if case != "tilted":
        assert info["bp_kernel"] == "separable" and info["bp_launch_angles"] == 25
    x = ts.phantom.box_noise((48, 48, 48), 12, 30, 0.01, seed=11)
    g = oracle.oracle_geometry(vg, pg)
    xd = torch.from_numpy(x).to(DEV)
    y = A(xd)
    assert rel_l2(y.cpu().numpy(), oracle.fp(g, x)) <= TOL
    yr = np.random.default_rng(12).random(A.range_shape).astype(np.float32)
    b = A.T(torch.from_numpy(yr).to(DEV))
    assert rel_l2(b.cpu().numpy(), oracle.bp(g, yr)) <= TOL
    # against the same geometry on unchunked plans
    B = ts.operator(vg, pg)
    assert B.projector.plan(0).info()["fp_slabs"] == 1
    assert _rel(y, B(xd)) <= 1e-6
    assert _rel(b, B.T(torch.from_numpy(yr).to(DEV))) <= 1e-6
    # matched adjoint of the chunked pair
    adj = abs(torch.sum(y.double() * torch.from_numpy(yr).to(DEV).double()).item()
              - torch.sum(xd.double() * b.double()).item())
    adj /= torch.linalg.norm(y.double()).item() * np.linalg.norm(yr.astype(np.float64))
    assert adj <= TOL


def test_chunked_epilogues_and_modes(monkeypatch):
    """ADD mode, the SIRT residual r = w (y - Ax) and the scaled update
    x += c A^T r on chunked plans equal the unchunked plans' results."""
    vg, pg = _cases()["cone"]
    A, _ = _chunked_operator(monkeypatch, vg, pg, 400000)
    B = ts.operator(vg, pg)
    g = torch.Generator(device=DEV)
    g.manual_seed(13)
    x = torch.rand(A.domain_shape, generator=g, device=DEV)
    y = torch.rand(A.range_shape, generator=g, device=DEV)
    w = torch.rand(A.range_shape, generator=g, device=DEV)
    c = torch.rand(A.domain_shape, generator=g, device=DEV)
    acc = torch.rand(A.range_shape, generator=g, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    added = []
    for op in (A, B):
        o = acc.clone()
        op.projector.plan(0).forward(x.data_ptr(), o.data_ptr(), True, s)
        added.append(o)
    assert _rel(added[0] - acc, added[1] - acc) <= 1e-6
    res = []
    for op in (A, B):
        p = op.projector.plan(0)
        r = torch.empty(op.range_shape, device=DEV)
        ws = torch.empty(max(p.workspace_size(0), 1), dtype=torch.uint8, device=DEV)
        p.forward_residual(x.data_ptr(), y.data_ptr(), w.data_ptr(), r.data_ptr(), s,
                           workspace=ws.data_ptr(), workspace_bytes=p.workspace_size(0))
        v = x.clone()
        wb = torch.empty(max(p.workspace_size(1), 1), dtype=torch.uint8, device=DEV)
        p.backward_scaled_add(v.data_ptr(), r.data_ptr(), c.data_ptr(), s,
                              workspace=wb.data_ptr(), workspace_bytes=p.workspace_size(1))
        res.append((r, v))
    assert _rel(res[0][0], res[1][0]) <= 1e-6
    # the updated volumes; their increments carry the rounding of the sums
    # x + c A^T r (about 6e-8 |x|) relative to the smaller increment
    assert _rel(res[0][1], res[1][1]) <= 1e-6
    assert _rel(res[0][1] - x, res[1][1] - x) <= TOL
    # the residual written over its own data term (out == y)
    r2 = y.clone()
    p = A.projector.plan(0)
    ws = torch.empty(p.workspace_size(0), dtype=torch.uint8, device=DEV)
    p.forward_residual(x.data_ptr(), r2.data_ptr(), w.data_ptr(), r2.data_ptr(), s,
                       workspace=ws.data_ptr(), workspace_bytes=p.workspace_size(0))
    assert torch.equal(r2, res[0][0])


def test_fused_launch_with_a_chunked_instance(monkeypatch):
    """A chunked instance in a fused launch takes its own launches; every
    instance equals its per-operator call bitwise."""
    from tomosipo_amd.fused import fused_forward, fused_backward

    vg, pg = _cases()["cone"]
    A, _ = _chunked_operator(monkeypatch, vg, pg, 400000)
    ops = [ts.operator(ts.volume(shape=24, size=1),
                       ts.cone(angles=30, shape=(24, 36), size=(1.4, 2.1), cone_angle=0.5)), A,
           ts.operator(ts.volume(shape=32, size=1),
                       ts.cone(angles=20, shape=(32, 48), size=(1.4, 2.1), cone_angle=0.5))]
    g = torch.Generator(device=DEV)
    g.manual_seed(14)
    xs = [torch.rand(op.domain_shape, generator=g, device=DEV) for op in ops]
    ys = fused_forward(ops, xs)
    for op, x, y in zip(ops, xs, ys):
        assert torch.equal(y, op(x))
    bs = fused_backward(ops, ys)
    for op, y, b in zip(ops, ys, bs):
        assert torch.equal(b, op.T(y))


def test_1024_cubed_forward_in_slabs_matches_oracle():
    """A 1024^3 volume (4.3 GB of float32; row-paired layouts of 8.6 GB per
    stepping axis) is forward-projected in 3 z-slabs without any lowered
    bound: two angles of a 1024 x 1024 detector against the oracle on the
    whole detector (the cone's rays cross both slab faces), and their
    backprojection against the oracle on a 128-slice subvolume."""
    n = 1024
    vg = ts.volume(shape=n, size=1)
    pg = ts.cone(angles=8, shape=(n, n), size=(np.sqrt(2), np.sqrt(2)), cone_angle=0.5).to_vec()[::4]
    A = ts.operator(vg, pg)
    info = A.projector.plan(0).info()
    print(f"[1024] fp slabs {info['fp_slabs']}, bp {info['bp_kernel']} "
          f"{info['bp_launch_angles']} angles per launch", flush=True)
    assert info["fp_kernel"] == "padded" and info["fp_slabs"] == 3
    g = torch.Generator(device=DEV)
    g.manual_seed(15)
    x = torch.rand((n, n, n), generator=g, device=DEV)
    y = A(x)
    xh = x.cpu().numpy().astype(np.float64)
    del x
    ref = oracle.fp(oracle.oracle_geometry(vg, pg), xh)
    del xh
    e = rel_l2(y.cpu().numpy(), ref)
    print(f"[1024] fp rel L2 {e:.3e}", flush=True)
    assert e <= TOL
    torch.cuda.empty_cache()
    b = A.T(y)
    z0, z1 = 280, 408
    sub = ts.volume(shape=(z1 - z0, n, n), pos=((z0 + z1) / 2 / n - 0.5, 0, 0), size=((z1 - z0) / n, 1, 1))
    ref = oracle.bp(oracle.oracle_geometry(sub, pg), y.cpu().numpy())
    e = rel_l2(b[z0:z1].cpu().numpy(), ref)
    print(f"[1024] bp slices {z0}..{z1 - 1}: rel L2 {e:.3e}", flush=True)
    assert e <= TOL


def test_1024_cubed_backprojection_in_angle_ranges():
    """1100 angles of a 1024 x 1024 detector: the transposed projections
    (4.6 GB) are backprojected in two ranges of sorted angles (1018 angles of
    4.2 MB stay below 4 GiB). The result equals the sum of the
    backprojections of two operators on the halves of the scan, which each
    fit one launch; the chunked pair (slabs, ranges) is a matched adjoint."""
    n, na = 1024, 1100
    vg = ts.volume(shape=n, size=1)
    pg = ts.cone(angles=na, shape=(n, n), size=(np.sqrt(2), np.sqrt(2)), cone_angle=0.5).to_vec()
    A = ts.operator(vg, pg)
    info = A.projector.plan(0).info()
    print(f"[1024] {na} angles: bp {info['bp_kernel']} {info['bp_launch_angles']} angles per launch",
          flush=True)
    assert info["bp_kernel"] == "separable" and info["bp_launch_angles"] == (2 ** 32 - 1) // ((n + 6) * n * 4)
    g = torch.Generator(device=DEV)
    g.manual_seed(16)
    y = torch.rand(A.range_shape, generator=g, device=DEV)
    b = A.T(y)
    h = na // 2
    parts = torch.zeros_like(b)
    for lo, hi in ((0, h), (h, na)):
        P = ts.operator(vg, pg[lo:hi])
        assert P.projector.plan(0).info()["bp_launch_angles"] == hi - lo
        parts.add_(P.T(y[:, lo:hi].contiguous()))
        del P
    e = _rel(b, parts)
    print(f"[1024] chunked vs halves: rel L2 {e:.3e}", flush=True)
    assert e <= 1e-6
    del parts
    x = torch.rand((n, n, n), generator=g, device=DEV)
    Ax = A(x)
    adj = abs(torch.sum(Ax.double() * y.double()).item() - torch.sum(x.double() * b.double()).item())
    adj /= torch.linalg.norm(Ax.double()).item() * torch.linalg.norm(y.double()).item()
    print(f"[1024] adjoint {adj:.3e}", flush=True)
    assert adj <= TOL
