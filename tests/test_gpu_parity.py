"""Parity of the HIP projector with the FP64 oracle (GPU).

Tolerances (floating point, FP32 kernels vs FP64 oracle):
  forward and backprojection: relative L2 error <= 1e-5;
  matched adjoint: |<Ax, y> - <x, A^T y>| / (|Ax| |y|) <= 1e-5.
"""
import numpy as np
import pytest

import oracle
import tomosipo_amd as ts
from tomosipo_amd import _tsx
from tests.conftest import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5
DEV = "cuda"


def _cone_vec_helix(n_angles, rows, cols, size_v, size_u, shift, jitter, seed):
    base = ts.cone(angles=n_angles, shape=(rows, cols), size=(size_v, size_u), cone_angle=0.5).to_vec()
    src = base.src_pos
    src[:, 0] += np.linspace(-shift, shift, n_angles)
    src += np.random.default_rng(seed).normal(0.0, jitter, size=src.shape)
    return ts.cone_vec(shape=base.det_shape, src_pos=src, det_pos=base.det_pos, det_v=base.det_v,
                       det_u=base.det_u)


def _check_fp_bp(vg, pg, x, ss=1, tol=TOL):
    A = ts.operator(vg, pg, detector_supersampling=ss)
    g = oracle.oracle_geometry(vg, pg, det_supersampling=ss)
    print(f"[parity] vol={vg.shape} det={pg.det_shape} angles={pg.num_angles} ss={ss}", flush=True)
    xd = torch.from_numpy(x).to(DEV)
    y = A(xd)
    y_ref = oracle.fp(g, x)
    e_fp = rel_l2(y.cpu().numpy(), y_ref)
    yin = y_ref.astype(np.float32)
    b = A.T(torch.from_numpy(yin).to(DEV))
    b_ref = oracle.bp(g, yin)
    e_bp = rel_l2(b.cpu().numpy(), b_ref)
    # matched adjoint on the GPU results, with an independent y
    yr = np.random.default_rng(7).random(A.range_shape).astype(np.float32)
    bt = A.T(torch.from_numpy(yr).to(DEV)).double().cpu().numpy()
    yt = y.double().cpu().numpy()
    adj = abs(np.sum(yt * yr) - np.sum(x.astype(np.float64) * bt)) / (np.linalg.norm(yt) * np.linalg.norm(yr))
    print(f"[parity]   fp={e_fp:.3e} bp={e_bp:.3e} adjoint={adj:.3e}", flush=True)
    assert e_fp <= tol
    assert e_bp <= tol
    assert adj <= tol


def test_parallel3d_vec_64():
    vg = ts.volume(shape=64, size=1)
    pg = ts.parallel(angles=96, shape=(64, 96), size=(1, 1.5)).to_vec()
    x = ts.phantom.box_noise((64, 64, 64), 16, 40, 0.01, seed=1)
    _check_fp_bp(vg, pg, x)


def test_cone_small():
    vg = ts.volume(shape=48, size=1)
    pg = ts.cone(angles=40, shape=(48, 72), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    x = ts.phantom.box_noise((48, 48, 48), 12, 30, 0.01, seed=2)
    _check_fp_bp(vg, pg, x)


def test_cone_vec_helix_jitter():
    vg = ts.volume(shape=40, size=1)
    pg = _cone_vec_helix(36, 40, 60, np.sqrt(2), 1.5 * np.sqrt(2), 0.25, 1e-3, seed=3)
    x = ts.phantom.ellipsoids(40, count=8, seed=3)
    _check_fp_bp(vg, pg, x)


@pytest.mark.parametrize("kind", ["parallel", "cone"])
def test_truncated_detector(kind):
    """A detector much narrower than the volume: for about a third of the
    (brick, angle) pairs of the separable backprojector no detector column
    meets the brick, and some bricks meet none at one angle and some at the
    next (the slice ranges must still advance through such angles)."""
    vg = ts.volume(shape=(20, 48, 48), size=(20 / 48, 1, 1))
    if kind == "parallel":
        pg = ts.parallel(angles=36, shape=(24, 6), size=(1.0, 0.3)).to_vec()
    else:
        pg = ts.cone(angles=36, shape=(24, 8), size=(1.5, 0.4), src_orig_dist=3,
                     src_det_dist=5).to_vec()
    assert ts.operator(vg, pg).projector.plan(0).info()["bp_kernel"] == "separable"
    x = ts.phantom.box_noise((20, 48, 48), 8, 40, 0.01, seed=4)
    _check_fp_bp(vg, pg, x)


@pytest.mark.parametrize("rows", [64, 39, 38, 24, 18])
def test_backprojector_row_pitch(rows):
    """The separable backprojector splats detector rows onto the z-grid in
    three ways: row by row (row step up to 1.01 voxels), with the pending cell
    carried in a register (1.01 to 1.99), or rows two apart together (from
    1.99). Parallel beam, 32^3 unit volume, detector 1.2 high: the row step is
    38.4 / rows voxels (0.6, 0.985, 1.0105, 1.6, 2.13)."""
    vg = ts.volume(shape=32, size=1)
    pg = ts.parallel(angles=24, shape=(rows, 48), size=(1.2, 1.5)).to_vec()
    assert ts.operator(vg, pg).projector.plan(0).info()["bp_kernel"] == "separable"
    x = ts.phantom.ellipsoids(32, count=6, seed=rows)
    _check_fp_bp(vg, pg, x)


@pytest.mark.parametrize("kind", ["parallel", "cone"])
def test_backprojector_short_detector(kind):
    """Row step 1.5 voxels (the splat that carries the pending cell in a
    register), with a detector shorter than the volume along z: the rows of
    the top and bottom bricks are cut off by the detector's edges, and those
    columns must still read zero where no row reaches."""
    vg = ts.volume(shape=32, size=1)
    if kind == "parallel":
        pg = ts.parallel(angles=20, shape=(16, 48), size=(0.75, 1.5)).to_vec()
    else:
        pg = ts.cone(angles=20, shape=(14, 48), size=(1.0, 2.4), src_orig_dist=4, src_det_dist=6).to_vec()
    assert ts.operator(vg, pg).projector.plan(0).info()["bp_kernel"] == "separable"
    x = ts.phantom.ellipsoids(32, count=6, seed=11)
    _check_fp_bp(vg, pg, x)


@pytest.mark.parametrize("fp_variant,brick_z", [("2", "32"), ("3", "16"), ("4", "32"), ("5", "32")])
def test_autotuner_alternatives_match_oracle(monkeypatch, fp_variant, brick_z):
    """The kernel variants the autotuner may choose besides the defaults: the
    per-lane forward march, the plain and row-paired layouts, the LDS-staged
    march and 16-deep backprojector bricks (forced through the environment
    overrides), on a helical cone_vec
    geometry with a volume depth that is not a multiple of the brick depth."""
    monkeypatch.setenv("TSX_FP_VARIANT", fp_variant)
    monkeypatch.setenv("TSX_BP_BRICK_Z", brick_z)
    vg = ts.volume(shape=(44, 40, 36), size=1)
    pg = _cone_vec_helix(30, 44, 60, np.sqrt(2), 1.5 * np.sqrt(2), 0.25, 1e-3, seed=7)
    x = ts.phantom.ellipsoids(44, count=8, seed=7)[:, :40, :36].copy()
    A = ts.operator(vg, pg)
    info = A.projector.plan(0).info()
    assert info["fp_variant"] == int(fp_variant) and info["bp_brick_z"] == int(brick_z)
    _check_fp_bp(vg, pg, x)


def _fp_with_variant(monkeypatch, A, x, variant):
    monkeypatch.setenv("TSX_FP_VARIANT", variant)
    y = A(x)
    torch.cuda.synchronize()
    return y


@pytest.mark.parametrize("case", ["cone", "parallel", "coarse", "supersampled", "fullsize"])
def test_lds_staged_march_equals_plain_march_bitwise(monkeypatch, case):
    """The LDS-staged forward march (variant 5) computes the plain march's
    (variant 3) samples, weights and sums in the same order, so the two agree
    bit for bit: on a helical cone_vec scan, a parallel scan, detector pixels
    four voxels wide (the block's slice window exceeds the LDS window and the
    kernel falls back to the plain march), detector supersampling 2 (four
    marches per block) and the 512^3 / 512 x 768 throughput geometry on 36 of
    its angles. Fused launches of variant 5 equal single ones."""
    from tomosipo_amd.fused import fused_forward

    if case == "cone":
        vg = ts.volume(shape=(44, 40, 36), size=1)
        pg = _cone_vec_helix(30, 44, 60, np.sqrt(2), 1.5 * np.sqrt(2), 0.25, 1e-3, seed=7)
    elif case == "parallel":
        vg = ts.volume(shape=(40, 36, 44), size=(1.0, 0.9, 1.1))
        pg = ts.parallel(angles=45, shape=(40, 60), size=(1.2, 1.8))
    elif case == "coarse":
        vg = ts.volume(shape=64, size=1)
        pg = ts.cone(angles=24, shape=(16, 24), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    elif case == "supersampled":
        vg = ts.volume(shape=40, size=1)
        pg = ts.cone(angles=20, shape=(40, 60), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    else:
        vg = ts.volume(shape=512, size=1)
        pg = ts.cone(angles=720, shape=(512, 768), size=(np.sqrt(2), 1.5 * np.sqrt(2)),
                     cone_angle=0.5).to_vec()[::20]
    A = ts.operator(vg, pg, detector_supersampling=2 if case == "supersampled" else 1)
    g = torch.Generator(device=DEV)
    g.manual_seed(5)
    x = torch.rand(A.domain_shape, generator=g, device=DEV)
    print(f"[lds] {case}: vol {vg.shape}, det {pg.det_shape}, {pg.num_angles} angles", flush=True)
    y3 = _fp_with_variant(monkeypatch, A, x, "3")
    y5 = _fp_with_variant(monkeypatch, A, x, "5")
    assert torch.equal(y3, y5), (y3 - y5).abs().max().item()
    if case in ("cone", "parallel"):
        ys = fused_forward([A, A], [x, x.flip(0).contiguous()])
        assert torch.equal(ys[0], y5)
        assert torch.equal(ys[1], A(x.flip(0).contiguous()))
    if case != "fullsize":
        go = oracle.oracle_geometry(vg, pg, det_supersampling=2 if case == "supersampled" else 1)
        assert rel_l2(y5.cpu().numpy(), oracle.fp(go, x.double().cpu().numpy())) <= TOL


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_random_geometries(seed):
    rng = np.random.default_rng(100 + seed)
    shape = tuple(int(s) for s in rng.integers(8, 24, size=3))
    vg = ts.volume(shape=shape, pos=rng.normal(0, 0.2, 3), size=rng.uniform(0.5, 1.5, 3))
    T = ts.rotate(pos=0, axis=rng.normal(size=3), angles=rng.normal())
    rows, cols = (int(s) for s in rng.integers(8, 30, size=2))
    if seed % 2 == 0:
        pg = T * ts.cone(angles=rng.uniform(0, 2 * np.pi, 17), shape=(rows, cols),
                         size=(rng.uniform(1, 3), rng.uniform(1, 3)), src_orig_dist=rng.uniform(2, 5),
                         src_det_dist=rng.uniform(5, 9)).to_vec()
    else:
        pg = T * ts.parallel(angles=rng.uniform(0, np.pi, 17), shape=(rows, cols),
                             size=(rng.uniform(1, 3), rng.uniform(1, 3))).to_vec()
    x = rng.random(vg.shape).astype(np.float32)
    _check_fp_bp(vg, pg, x)


@pytest.mark.parametrize("ss", [2, 3, 6])
def test_detector_supersampling(ss):
    vg = ts.volume(shape=24, size=1)
    pg = ts.cone(angles=12, shape=(16, 20), size=(1.4, 2.0), cone_angle=0.5)
    x = ts.phantom.ellipsoids(24, count=5, seed=ss)
    _check_fp_bp(vg, pg, x, ss=ss)


@pytest.mark.parametrize("case", ["cone2", "parallel3", "turned2", "cone5"])
def test_refined_detector_projectors(monkeypatch, case):
    """Detector supersampling ss: both projectors run on the plan of the
    detector refined ss times -- the forward projection as the mean of each
    pixel's ss x ss fine pixels, the backprojection with each fine pixel
    taking its coarse pixel's value over ss^2. Parity with the oracle and the
    matched adjoint (1e-5), agreement with the general kernels of the same
    plan (forward 1e-5, backward 1e-6), ADD, residual and scaled-add
    epilogues, a batch of two, a fused launch holding the instance, and the
    refined plan's angle ranges under a lowered chunk bound."""
    import os
    from tomosipo_amd.fused import fused_backward

    ss = int(case[-1])
    vg = ts.volume(shape=(20, 24, 28), size=(20 / 24, 1, 28 / 24))
    if case.startswith("parallel"):
        pg = ts.parallel(angles=16, shape=(10, 18), size=(1.0, 1.6)).to_vec()
    else:
        pg = ts.cone(angles=16, shape=(12, 18), size=(1.4, 2.1), cone_angle=0.5).to_vec()
    if case.startswith("turned"):
        pg = ts.rotate(pos=0, axis=(0, 1, 0), angles=np.pi / 2) * pg
    A = ts.operator(vg, pg, detector_supersampling=ss)
    plan = A.projector.plan(0)
    info = plan.info()
    print(f"[refined] {case}: {info}", flush=True)
    assert (info["bp_kernel"], info["bp_refined_detector"]) == ("separable", ss)
    assert (info["fp_kernel"], info["fp_refined_detector"]) == ("padded", ss)
    assert info["bp_axis_swap"] == info["fp_axis_swap"] == (1 if case.startswith("turned") else 0)
    x = ts.phantom.ellipsoids(28, count=5, seed=ss)[:20, :24, :28].copy()
    _check_fp_bp(vg, pg, x, ss=ss)
    g = torch.Generator(device=DEV)
    g.manual_seed(31)
    y = torch.rand(A.range_shape, generator=g, device=DEV)
    xd = torch.from_numpy(x).to(DEV)
    fx = A(xd)
    b = A.T(y)
    os.environ["TSX_FORCE_GENERAL"] = "1"
    try:
        assert plan.info()["bp_kernel"] == "general"
        fx_gen = A(xd)
        b_gen = A.T(y)
    finally:
        del os.environ["TSX_FORCE_GENERAL"]
    assert (torch.linalg.norm(fx - fx_gen) / torch.linalg.norm(fx_gen)).item() <= 1e-5
    assert (torch.linalg.norm(b - b_gen) / torch.linalg.norm(b_gen)).item() <= 1e-6
    s = torch.cuda.current_stream().cuda_stream
    accp = torch.rand(A.range_shape, generator=g, device=DEV)
    op_ = accp.clone()
    plan.forward(xd.data_ptr(), op_.data_ptr(), True, s)
    assert (torch.linalg.norm(op_ - accp - fx) / torch.linalg.norm(fx)).item() <= 1e-6
    w = torch.rand(A.range_shape, generator=g, device=DEV)
    r = torch.empty(A.range_shape, device=DEV)
    plan.forward_residual(xd.data_ptr(), y.data_ptr(), w.data_ptr(), r.data_ptr(), s)
    assert (torch.linalg.norm(r - w * (y - fx)) / torch.linalg.norm(w * (y - fx))).item() <= 1e-6
    xs = torch.rand((2,) + tuple(A.domain_shape), generator=g, device=DEV)
    fs = torch.empty((2,) + tuple(A.range_shape), device=DEV)
    plan.forward(xs.data_ptr(), fs.data_ptr(), False, s, batch=2, vol_stride=xs[0].numel(),
                 proj_stride=fs[0].numel())
    for k in range(2):
        assert torch.equal(fs[k], A(xs[k]))
    acc = torch.rand(A.domain_shape, generator=g, device=DEV)
    o = acc.clone()
    plan.backward(o.data_ptr(), y.data_ptr(), True, s)
    assert (torch.linalg.norm(o - acc - b) / torch.linalg.norm(b)).item() <= 1e-6
    c = torch.rand(A.domain_shape, generator=g, device=DEV)
    v = acc.clone()
    plan.backward_scaled_add(v.data_ptr(), y.data_ptr(), c.data_ptr(), s)
    assert (torch.linalg.norm(v - (acc + c * b)) / torch.linalg.norm(c * b)).item() <= 1e-6
    ys = torch.rand((2,) + tuple(A.range_shape), generator=g, device=DEV)
    bs = torch.empty((2,) + tuple(A.domain_shape), device=DEV)
    plan.backward(bs.data_ptr(), ys.data_ptr(), False, s, batch=2, vol_stride=bs[0].numel(),
                  proj_stride=ys[0].numel())
    for k in range(2):
        assert torch.equal(bs[k], A.T(ys[k]))
    other = ts.operator(ts.volume(shape=16, size=1),
                        ts.cone(angles=10, shape=(16, 24), size=(1.4, 2.1), cone_angle=0.5))
    ops = [other, A, other]
    ins = [torch.rand(op.range_shape, generator=g, device=DEV) for op in ops]
    outs = fused_backward(ops, ins)
    for op, yi, bi in zip(ops, ins, outs):
        assert torch.equal(bi, op.T(yi))
    from tomosipo_amd.fused import fused_forward
    vins = [torch.rand(op.domain_shape, generator=g, device=DEV) for op in ops]
    fouts = fused_forward(ops, vins)
    for op, xi, yi in zip(ops, vins, fouts):
        assert torch.equal(yi, op(xi))
    # the refined plan's transposed projections in ranges of 5 angles
    per_angle = (12 * ss + 6) * 18 * ss * 4 if not case.startswith("parallel") else (10 * ss + 6) * 18 * ss * 4
    monkeypatch.setenv("TSX_CHUNK_BYTES", str(5 * per_angle + 1))
    C = ts.operator(vg, pg, detector_supersampling=ss)
    cinfo = C.projector.plan(0).info()
    monkeypatch.delenv("TSX_CHUNK_BYTES")
    assert (cinfo["bp_refined_detector"], cinfo["bp_launch_angles"]) == (ss, 5)
    assert (torch.linalg.norm(C.T(y) - b) / torch.linalg.norm(b)).item() <= 1e-6


@pytest.mark.parametrize("kind", ["cone", "parallel"])
def test_sheared_forward_tiles(monkeypatch, kind):
    """A detector rolled about its normal: the forward tiles follow the
    detector direction that keeps a wave's rays on one row of the slices,
    each tile column's rows taken cyclically. Parity with the oracle (1e-5)
    on a detector whose rows are not a multiple of the tile rows, sheared
    and plain tiles (TSX_FP_SHEAR) equal up to the order of their sums
    (1e-6, every pixel covered once), tile rows 4 and 8, and the fused
    launch equal to single calls."""
    from tomosipo_amd.fused import fused_forward

    vg = ts.volume(shape=(36, 40, 44), size=(36 / 40, 1, 44 / 40))
    if kind == "cone":
        base = ts.cone(angles=24, shape=(45, 70), size=(1.4, 2.1), cone_angle=0.5)
    else:
        base = ts.parallel(angles=24, shape=(45, 70), size=(1.0, 1.5))
    pg = ts.rotate(pos=0, axis=(0, 1, 1), angles=0.5) * base.to_vec()
    A = ts.operator(vg, pg)
    info = A.projector.plan(0).info()
    print(f"[shear] {kind}: {info['fp_shear_rows']} rows", flush=True)
    assert info["fp_kernel"] == "padded" and info["fp_shear_rows"] > 0
    x = ts.phantom.ellipsoids(44, count=6, seed=41)[:36, :40, :44].copy()
    _check_fp_bp(vg, pg, x)
    xd = torch.from_numpy(x).to(DEV)
    outs = {}
    for shear in ("1", "0"):
        for rows in ("8", "4"):
            monkeypatch.setenv("TSX_FP_SHEAR", shear)
            monkeypatch.setenv("TSX_FP_TILE_ROWS", rows)
            assert (A.projector.plan(0).info()["fp_shear_rows"] > 0) == (shear == "1")
            outs[shear, rows] = A(xd)
            torch.cuda.synchronize()
    monkeypatch.delenv("TSX_FP_SHEAR")
    monkeypatch.delenv("TSX_FP_TILE_ROWS")
    ref = outs["0", "8"]
    for k, y in outs.items():
        assert (torch.linalg.norm(y - ref) / torch.linalg.norm(ref)).item() <= 1e-6, k
    other = ts.operator(ts.volume(shape=16, size=1),
                        ts.cone(angles=10, shape=(16, 24), size=(1.4, 2.1), cone_angle=0.5))
    ops = [other, A, other]
    g = torch.Generator(device=DEV)
    g.manual_seed(42)
    ins = [torch.rand(op.domain_shape, generator=g, device=DEV) for op in ops]
    ys = fused_forward(ops, ins)
    for op, xi, yi in zip(ops, ins, ys):
        assert torch.equal(yi, op(xi))


@pytest.mark.parametrize("case", ["rolled_cone_ss2", "rolled_parallel_ss3", "turned_x_cone_ss2",
                                  "turned_y_parallel_ss2", "rolled_turned_cone"])
def test_geometry_path_combinations(case):
    """The plans a plan holds, combined: a rolled detector with supersampling
    (forward on the refined plan with sheared tiles, backward on the general
    kernel with sub-rays), a turned scan with supersampling (both projectors
    on the refined plan's axis-swapped plan) and a turned scan whose detector
    is also rolled a little (no swap: waves along detector columns, sheared;
    general backprojector).
    Parity with the oracle and the matched adjoint at 1e-5."""
    vg = ts.volume(shape=(24, 28, 20), size=(24 / 28, 1, 20 / 28))
    kind, ss = ("parallel" if "parallel" in case else "cone"), (int(case[-1]) if "_ss" in case else 1)
    # (detector rows enough for the shear of a 0.5 rad roll: at most half the rows)
    if kind == "cone":
        base = ts.cone(angles=14, shape=(36, 24), size=(1.4, 1.4), cone_angle=0.5).to_vec()
    else:
        base = ts.parallel(angles=14, shape=(36, 24), size=(1.0, 1.0)).to_vec()
    pg = base
    if "turned_x" in case:
        pg = ts.rotate(pos=0, axis=(0, 0, 1), angles=np.pi / 2) * pg
    if "turned_y" in case or case == "rolled_turned_cone":
        pg = ts.rotate(pos=0, axis=(0, 1, 0), angles=np.pi / 2) * pg
    if "rolled" in case:
        pg = ts.rotate(pos=0, axis=(0, 1, 1), angles=0.5 if case != "rolled_turned_cone" else 0.15) * pg
    info = ts.operator(vg, pg, detector_supersampling=ss).projector.plan(0).info()
    print(f"[combo] {case}: {info}", flush=True)
    if "turned" in case and "rolled" not in case:
        assert info["fp_axis_swap"] and info["bp_axis_swap"]
    if ss > 1:
        assert info["fp_refined_detector"] == ss
        assert info["bp_refined_detector"] == (ss if "turned" in case else 0)
    if "rolled" in case:
        assert info["bp_kernel"] == "general" and info["fp_shear_rows"] > 0
    # on the turned scan rolled a little the detector's rows run along the
    # slices' rows: waves along its columns, sheared in columns
    assert info["fp_column_waves"] == (case == "rolled_turned_cone")
    x = ts.phantom.ellipsoids(28, count=5, seed=len(case))[:24, :28, :20].copy()
    _check_fp_bp(vg, pg, x, ss=ss)


@pytest.mark.parametrize("kind", ["parallel", "cone"])
def test_wide_slices(kind):
    """Slices of 2048 x 2048: the padded layouts' rows (33.6 MB, row-paired)
    exceed 24-bit row offsets, and the forward projector forms them with
    32-bit multiplies. Parity with the oracle and the matched adjoint (1e-5),
    agreement with the general forward kernel (1e-5)."""
    import os
    vg = ts.volume(shape=(8, 2048, 2048), size=(8 / 2048, 1, 1))
    if kind == "parallel":
        pg = ts.parallel(angles=4, shape=(8, 2048), size=(8 / 2048, 1.5)).to_vec()
    else:
        pg = ts.cone(angles=4, shape=(8, 2048), size=(16 / 2048, 2.0), src_orig_dist=4,
                     src_det_dist=8).to_vec()
    A = ts.operator(vg, pg)
    assert A.projector.plan(0).info()["fp_kernel"] == "padded"
    x = np.random.default_rng(51).random(vg.shape).astype(np.float32)
    _check_fp_bp(vg, pg, x)
    xd = torch.from_numpy(x).to(DEV)
    y = A(xd)
    os.environ["TSX_FORCE_GENERAL"] = "1"
    try:
        y_gen = A(xd)
    finally:
        del os.environ["TSX_FORCE_GENERAL"]
    assert (torch.linalg.norm(y - y_gen) / torch.linalg.norm(y_gen)).item() <= 1e-5


def test_additive_and_out():
    vg = ts.volume(shape=32, size=1)
    pg = ts.parallel(angles=20, shape=(32, 48), size=(1, 1.5))
    A = ts.operator(vg, pg)
    B = ts.operator(vg, pg, additive=True)
    x = torch.from_numpy(ts.phantom.box_noise((32, 32, 32), 8, 20, 0.01, seed=4)).to(DEV)
    y = torch.zeros(A.range_shape, device=DEV)
    B(x, out=y)
    B(x, out=y)
    torch.testing.assert_close(y, 2 * A(x), rtol=1e-6, atol=1e-6)
    b = torch.zeros(A.domain_shape, device=DEV)
    B.T(y, out=b)
    B.T(y, out=b)
    torch.testing.assert_close(b, 2 * A.T(y), rtol=1e-6, atol=1e-6)


def test_numpy_and_data_paths():
    vg = ts.volume(shape=32, size=1)
    pg = ts.cone(angles=20, shape=(32, 48), size=(1.4, 2.1), cone_angle=0.5)
    A = ts.operator(vg, pg)
    x = ts.phantom.box_noise((32, 32, 32), 8, 20, 0.01, seed=5)
    y_np = A(x)
    assert isinstance(y_np, np.ndarray) and y_np.dtype == np.float32
    y_t = A(torch.from_numpy(x).to(DEV)).cpu().numpy()
    np.testing.assert_array_equal(y_np, y_t)
    vd = ts.data(vg, x.copy())
    yd = A(vd)
    assert isinstance(yd, ts.Data.Data)
    np.testing.assert_array_equal(yd.data, y_np)
    np.testing.assert_array_equal(A.T(yd).data, A.T(y_np))
    # CPU torch tensors stay on the CPU
    y_cpu = A(torch.from_numpy(x))
    assert y_cpu.device.type == "cpu"
    np.testing.assert_array_equal(y_cpu.numpy(), y_np)


def test_batched_equals_single():
    from tomosipo_amd.torch_support import to_autograd

    vg = ts.volume(shape=24, size=1)
    pg = ts.parallel(angles=30, shape=(24, 36), size=(1, 1.5))
    A = ts.operator(vg, pg)
    f = to_autograd(A, num_extra_dims=2)
    xs = torch.from_numpy(np.stack([ts.phantom.ellipsoids(24, 6, seed=s) for s in range(6)])).to(DEV)
    xs = xs.reshape(2, 3, 24, 24, 24)
    ys = f(xs)
    for a in range(2):
        for c in range(3):
            assert torch.equal(ys[a, c], A(xs[a, c].contiguous()))
    g = to_autograd(A.T, num_extra_dims=2)
    bs = g(ys)
    for a in range(2):
        for c in range(3):
            assert torch.equal(bs[a, c], A.T(ys[a, c].contiguous()))


def test_autograd_grad_equals_ATAx():
    from tomosipo_amd.torch_support import to_autograd, AutogradOperator

    vg = ts.volume(shape=(1, 64, 64))
    pg = ts.parallel(angles=96, shape=(1, 96))
    A = ts.operator(vg, pg)
    x = torch.ones(A.domain_shape, dtype=torch.float32, device=DEV, requires_grad=True)
    y = to_autograd(A)(x)
    y.backward(y)
    torch.testing.assert_close(x.grad, A.T(A(x.detach())))
    A_ag = AutogradOperator(A)
    p = torch.ones(A.range_shape, dtype=torch.float32, device=DEV, requires_grad=True)
    q = A_ag.T(p)
    q.backward(q)
    torch.testing.assert_close(p.grad, A(A.T(p.detach())))


@pytest.mark.parametrize("kind", ["parallel", "cone"])
def test_fast_and_general_kernels_agree(kind, monkeypatch):
    """The padded FP / separable BP and the general kernels compute the same operator."""
    vg = ts.volume(shape=(40, 36, 44), size=(1.0, 0.9, 1.1))
    if kind == "parallel":
        pg = ts.parallel(angles=45, shape=(40, 60), size=(1.2, 1.8))
    else:
        pg = _cone_vec_helix(45, 40, 60, 1.2, 1.8, 0.2, 1e-3, seed=9)
    A = ts.operator(vg, pg)
    info = A.projector._host_plan.info()
    assert info["fp_kernel"] == "padded" and info["bp_kernel"] == "separable", info
    x = torch.from_numpy(ts.phantom.ellipsoids(44, 6, seed=1)[:40, :36, :44].copy()).to(DEV)
    y_fast, b_fast = A(x), A.T(A(x))
    monkeypatch.setenv("TSX_FORCE_GENERAL", "1")
    y_gen, b_gen = A(x), A.T(A(x))
    assert rel_l2(y_fast.cpu().numpy(), y_gen.cpu().numpy()) <= TOL
    assert rel_l2(b_fast.cpu().numpy(), b_gen.cpu().numpy()) <= TOL
    g = oracle.oracle_geometry(vg, pg)
    y_ref = oracle.fp(g, x.cpu().numpy())
    assert rel_l2(y_fast.cpu().numpy(), y_ref) <= TOL
    assert rel_l2(y_gen.cpu().numpy(), y_ref) <= TOL


def test_fp64_kernels_match_oracle_and_gradcheck():
    """FP64 FP/BP agree with the FP64 oracle to ~1e-12; FD gradient check in FP64
    at 16^3, 24 angles (loss ||Ax - y||^2 through the autograd wrapper)."""
    from tomosipo_amd.torch_support import to_autograd

    vg = ts.volume(shape=16, size=1)
    pg = ts.parallel(angles=24, shape=(16, 24), size=(1, 1.5))
    A = ts.operator(vg, pg)
    g = oracle.oracle_geometry(vg, pg)
    rng = np.random.default_rng(11)
    x = rng.random(A.domain_shape)
    y = rng.random(A.range_shape)
    f = to_autograd(A, dtype=torch.float64)
    xd = torch.tensor(x, device=DEV, dtype=torch.float64, requires_grad=True)
    yd = torch.tensor(y, device=DEV, dtype=torch.float64)
    Ax = f(xd)
    assert rel_l2(Ax.detach().cpu().numpy(), oracle.fp(g, x)) < 1e-12
    loss = ((Ax - yd) ** 2).sum()
    loss.backward()
    grad_ref = 2 * oracle.bp(g, oracle.fp(g, x) - y)
    assert rel_l2(xd.grad.cpu().numpy(), grad_ref) < 1e-12
    # finite differences of the loss in FP64
    xs = torch.tensor(rng.random(A.domain_shape) * 0.1, device=DEV, dtype=torch.float64,
                      requires_grad=True)
    assert torch.autograd.gradcheck(lambda v: ((f(v) - yd) ** 2).sum(), (xs,), eps=1e-6,
                                    atol=1e-6, rtol=1e-5, fast_mode=True)
    # the float32 path keeps the reference contract: float32 out for float64 in
    assert to_autograd(A)(xd.detach()).dtype == torch.float32


def test_batched_gradient_matches_oracle():
    """Batch of 8 volumes through to_autograd(num_extra_dims=1): the gradient of
    ||Ax - y||^2 equals the oracle's 2 A^T (Ax - y) within 1e-5 (reduced size)."""
    from tomosipo_amd.torch_support import to_autograd

    vg = ts.volume(shape=32, size=1)
    pg = ts.parallel(angles=45, shape=(32, 48), size=(1, 1.5))
    A = ts.operator(vg, pg)
    g = oracle.oracle_geometry(vg, pg)
    xs = np.stack([ts.phantom.ellipsoids(32, 20, seed=s) for s in range(8)])
    ys = np.stack([oracle.fp(g, ts.phantom.ellipsoids(32, 20, seed=100 + s)) for s in range(8)])
    x = torch.tensor(xs, device=DEV, requires_grad=True)
    y = torch.tensor(ys.astype(np.float32), device=DEV)
    loss = ((to_autograd(A, num_extra_dims=1)(x) - y) ** 2).sum()
    loss.backward()
    for k in range(8):
        ref = 2 * oracle.bp(g, oracle.fp(g, xs[k]) - ys[k])
        assert rel_l2(x.grad[k].cpu().numpy(), ref) <= 1e-5


def test_backprojector_angle_parts(monkeypatch):
    """Volumes of few bricks (64^3: 32) sum the angles of the separable
    backprojector in parts of consecutive angles: a single launch runs the
    parts in separate blocks and adds their sums in order, a fused launch runs
    them in turn in one block. Single, batched and fused launches give the
    same sums bitwise (plain, additive and SIRT scaled-add epilogues); they
    agree with one part (TSX_BP_PARTS=1) to FP32 rounding and with the
    oracle."""
    from tomosipo_amd.fused import fused_backward, Epilogue

    vg1, pg1 = ts.volume(shape=64, size=1), ts.parallel(angles=96, shape=(64, 96), size=(1, 1.5)).to_vec()
    vg2 = ts.volume(shape=(40, 48, 44), size=1)
    pg2 = _cone_vec_helix(60, 40, 60, 1.4, 2.0, 0.1, 1e-3, seed=3)
    A, B = ts.operator(vg1, pg1), ts.operator(vg2, pg2)
    plan = A.projector.plan(0)
    ve = 64 ** 3
    # the workspace holds the part sums (24 parts of 4 angles)
    assert plan.workspace_size(_tsx.TSX_OP_BP) >= 24 * ve * 4
    g = torch.Generator(device=DEV)
    g.manual_seed(5)
    ya = torch.rand(A.range_shape, generator=g, device=DEV)
    yb = torch.rand(B.range_shape, generator=g, device=DEV)
    ba, bb = A.T(ya), B.T(yb)
    fa, fb = fused_backward([A, B], [ya, yb])
    assert torch.equal(fa, ba) and torch.equal(fb, bb)
    # batched: three projection sets of A in one call
    ys = torch.stack([ya, ya * 2, ya + 1])
    out = torch.empty((3,) + A.domain_shape, device=DEV)
    plan.backward(out.data_ptr(), ys.data_ptr(), False, torch.cuda.current_stream().cuda_stream, batch=3,
                  vol_stride=ve, proj_stride=ya.numel())
    for k in range(3):
        assert torch.equal(out[k], A.T(ys[k])), k
    # additive and SIRT scaled add through the reduction, single and fused
    x0 = torch.rand(A.domain_shape, generator=g, device=DEV)
    c = torch.rand(A.domain_shape, generator=g, device=DEV)
    add_s, add_f = x0.clone(), x0.clone()
    plan.backward(add_s.data_ptr(), ya.data_ptr(), True, torch.cuda.current_stream().cuda_stream)
    fused_backward([A], [ya], out=[add_f], additive=True)
    assert torch.equal(add_s, add_f) and torch.equal(add_s, x0 + ba)
    sa_s, sa_f = x0.clone(), x0.clone()
    plan.backward_scaled_add(sa_s.data_ptr(), ya.data_ptr(), c.data_ptr(), torch.cuda.current_stream().cuda_stream)
    fused_backward([A], [ya], out=[sa_f], epilogues=[Epilogue("scaled_add", w=c)])
    assert torch.equal(sa_s, sa_f)
    assert rel_l2(sa_s.cpu().numpy(), (x0 + c * ba).cpu().numpy()) <= 1e-6
    # one part: the same sums to FP32 rounding
    monkeypatch.setenv("TSX_BP_PARTS", "1")
    assert plan.workspace_size(_tsx.TSX_OP_BP) < 12 * ve * 4
    ba1, bb1 = A.T(ya), B.T(yb)
    monkeypatch.delenv("TSX_BP_PARTS")
    print(f"[parts] one part vs parts: {rel_l2(ba.cpu().numpy(), ba1.cpu().numpy()):.2e}, "
          f"{rel_l2(bb.cpu().numpy(), bb1.cpu().numpy()):.2e}", flush=True)
    assert rel_l2(ba.cpu().numpy(), ba1.cpu().numpy()) <= 1e-6
    assert rel_l2(bb.cpu().numpy(), bb1.cpu().numpy()) <= 1e-6
    for vg, pg, y, b in ((vg1, pg1, ya, ba), (vg2, pg2, yb, bb)):
        geo = oracle.oracle_geometry(vg, pg)
        assert rel_l2(b.cpu().numpy(), oracle.bp(geo, y.double().cpu().numpy())) <= TOL


def test_fused_multi_instance_equals_single():
    """Fused launches over instances of different shapes and geometries are
    bit-identical to per-operator calls."""
    from tomosipo_amd.fused import fused_forward, fused_backward

    ops, xs = [], []
    for k in range(5):
        n = 16 + 4 * k
        vg = ts.volume(shape=(n, n + 2, n - 2), size=1)
        if k % 2:
            pg = ts.parallel(angles=10 + k, shape=(n, n + 8), size=(1.2, 1.6))
        else:
            pg = _cone_vec_helix(12 + k, n, n + 10, 1.4, 2.0, 0.1, 1e-3, seed=k)
        ops.append(ts.operator(vg, pg))
        xs.append(torch.from_numpy(ts.phantom.ellipsoids(n + 2, 5, seed=k)[:n, :n + 2, :n - 2].copy()).to(DEV))
    # one instance that needs the general kernels (tilted detector)
    vg = ts.volume(shape=20, size=1)
    ops.append(ts.operator(vg, ts.rotate(pos=0, axis=(0, 1, 1), angles=0.3) *
                           ts.cone(angles=9, shape=(20, 30), size=(1.4, 2.1), cone_angle=0.5).to_vec()))
    xs.append(torch.rand(20, 20, 20, device=DEV))
    ys = fused_forward(ops, xs)
    for A, x, y in zip(ops, xs, ys):
        assert torch.equal(y, A(x))
    bs = fused_backward(ops, ys)
    for A, y, b in zip(ops, ys, bs):
        assert torch.equal(b, A.T(y))


def test_fused_normal_pipeline_equals_sequential():
    """fused_normal over instance groups (each group's backprojection on a side
    stream, overlapping the next group's forward projection) gives the results
    of one fused forward call followed by one fused backward call, bitwise, for
    1, 3 and more groups than instances; work queued on the caller's stream
    afterwards sees every result; one instance is checked against the oracle."""
    from tomosipo_amd.fused import fused_forward, fused_backward, fused_normal

    ops, xs, geos = [], [], []
    for k in range(6):
        n = 24 + 8 * k
        vg = ts.volume(shape=(n, n + 2, n - 2), size=1)
        if k % 2:
            pg = ts.parallel(angles=20 + k, shape=(n, n + 8), size=(1.2, 1.6))
        else:
            pg = _cone_vec_helix(24 + k, n, n + 10, 1.4, 2.0, 0.1, 1e-3, seed=k)
        ops.append(ts.operator(vg, pg))
        geos.append((vg, pg))
        xs.append(torch.rand((n, n + 2, n - 2), device=DEV))
    vg = ts.volume(shape=20, size=1)
    ops.append(ts.operator(vg, ts.rotate(pos=0, axis=(0, 1, 1), angles=0.3) *
                           ts.cone(angles=9, shape=(20, 30), size=(1.4, 2.1), cone_angle=0.5).to_vec()))
    xs.append(torch.rand(20, 20, 20, device=DEV))
    ref_y = fused_forward(ops, xs)
    ref_b = fused_backward(ops, ref_y)
    for groups in (1, 3, 20):
        print(f"[fused_normal] groups {groups}", flush=True)
        events = []
        ys, bs = fused_normal(ops, xs, groups=groups, events=events)
        assert len(events) == min(groups, len(ops))
        assert [e[0] for e in events][0] == 0 and events[-1][1] == len(ops) - 1
        sums = torch.stack([b.sum() for b in bs])      # queued on the caller's stream
        for k in range(len(ops)):
            assert torch.equal(ys[k], ref_y[k]), (groups, k)
            assert torch.equal(bs[k], ref_b[k]), (groups, k)
        assert torch.equal(sums, torch.stack([b.sum() for b in ref_b]))
    # outputs given by the caller are filled in place
    ys = [torch.full(A.range_shape, 7.0, device=DEV) for A in ops]
    bs = [torch.full(A.domain_shape, 7.0, device=DEV) for A in ops]
    fused_normal(ops, xs, ys, bs, groups=4)
    for k in range(len(ops)):
        assert torch.equal(ys[k], ref_y[k]) and torch.equal(bs[k], ref_b[k])
    g = oracle.oracle_geometry(*geos[2])
    y64 = oracle.fp(g, xs[2].double().cpu().numpy())
    assert rel_l2(ys[2].cpu().numpy(), y64) <= TOL
    assert rel_l2(bs[2].cpu().numpy(), oracle.bp(g, ys[2].double().cpu().numpy())) <= TOL


def test_fused_more_instances_than_one_launch_takes():
    """20 instances of assorted shapes: more than one fused launch holds (16).
    Two of them are 128^3 scans with 200 angles (9600 forward blocks each), so
    the first fused forward launch has more than 16384 blocks and takes chunks
    of 4 in its order over the XCDs; the small ones (tens to a few thousand
    blocks) keep chunks of 1. Fused results equal per-operator calls bitwise,
    and every instance is checked against the oracle."""
    from tomosipo_amd.fused import fused_forward, fused_backward

    rng = np.random.default_rng(11)
    ops, xs, geos = [], [], []
    for k in range(20):
        n = 128 if k in (3, 7) else int(rng.integers(8, 40))
        a = 200 if k in (3, 7) else int(rng.integers(3, 30))
        vg = ts.volume(shape=(n, n + 1, n + 3), size=1)
        if k % 3 == 0:
            pg = ts.parallel(angles=a, shape=(n, 3 * n // 2), size=(1.2, 1.6))
        else:
            pg = _cone_vec_helix(a, n + 2, 3 * n // 2, 1.4, 2.0, 0.1, 1e-3, seed=k)
        ops.append(ts.operator(vg, pg))
        geos.append((vg, pg))
        xs.append(torch.from_numpy(rng.random((n, n + 1, n + 3), dtype=np.float32)).to(DEV))
    print("[fused] 20 instances", flush=True)
    ys = fused_forward(ops, xs)
    for A, x, y in zip(ops, xs, ys):
        assert torch.equal(y, A(x))
    bs = fused_backward(ops, ys)
    for A, y, b in zip(ops, ys, bs):
        assert torch.equal(b, A.T(y))
    for k in range(20):
        g = oracle.oracle_geometry(*geos[k])
        assert rel_l2(ys[k].cpu().numpy(), oracle.fp(g, xs[k].cpu().numpy())) <= TOL, k
        assert rel_l2(bs[k].cpu().numpy(), oracle.bp(g, ys[k].double().cpu().numpy())) <= TOL, k


def test_fused_epilogues_equal_single_calls():
    """Per-instance epilogues in one fused launch: a plain projection, an
    accumulation and the two SIRT half-steps side by side, each equal to the
    single-operator call with the same epilogue (tsx_forward_residual,
    tsx_backward_scaled_add) bitwise."""
    from tomosipo_amd.fused import fused_forward, fused_backward, Epilogue

    ops = [ts.operator(ts.volume(shape=n, size=1),
                       ts.cone(angles=20 + n, shape=(n, 3 * n // 2), size=(1.4, 2.1), cone_angle=0.5))
           for n in (16, 24, 32)]
    g = torch.Generator(device=DEV)
    g.manual_seed(8)
    xs = [torch.rand(A.domain_shape, generator=g, device=DEV) for A in ops]
    ys = [torch.rand(A.range_shape, generator=g, device=DEV) for A in ops]
    ws = [torch.rand(A.range_shape, generator=g, device=DEV) for A in ops]
    acc = torch.rand(ops[1].range_shape, generator=g, device=DEV)
    out = [torch.empty(ops[0].range_shape, device=DEV), acc.clone(), torch.empty(ops[2].range_shape, device=DEV)]
    fused_forward(ops, xs, out=out,
                  epilogues=[Epilogue("set"), Epilogue("add"), Epilogue("residual", y=ys[2], w=ws[2])])
    assert torch.equal(out[0], ops[0](xs[0]))
    assert torch.equal(out[1], acc + ops[1](xs[1]))
    p2 = ops[2].projector.plan(0)
    ref = torch.empty(ops[2].range_shape, device=DEV)
    p2.forward_residual(xs[2].data_ptr(), ys[2].data_ptr(), ws[2].data_ptr(), ref.data_ptr(),
                        torch.cuda.current_stream().cuda_stream,
                        workspace=torch.empty(p2.workspace_size(0), dtype=torch.uint8, device=DEV).data_ptr(),
                        workspace_bytes=p2.workspace_size(0))
    assert torch.equal(out[2], ref)
    cs = [torch.rand(A.domain_shape, generator=g, device=DEV) for A in ops]
    vols = [x.clone() for x in xs]
    fused_backward(ops, ys, out=vols, epilogues=[Epilogue("scaled_add", w=c) for c in cs])
    for A, x, y, c, v in zip(ops, xs, ys, cs, vols):
        p = A.projector.plan(0)
        r = x.clone()
        nb = p.workspace_size(1)
        p.backward_scaled_add(r.data_ptr(), y.data_ptr(), c.data_ptr(), torch.cuda.current_stream().cuda_stream,
                              workspace=torch.empty(max(nb, 1), dtype=torch.uint8, device=DEV).data_ptr(),
                              workspace_bytes=nb)
        assert torch.equal(v, r)


def test_autotune_keeps_parity_and_caches_choices():
    """Autotuning picks kernel variants per plan; results stay within the
    oracle tolerance, a second plan of the same shape takes the cached
    choices, and fused launches over differently tuned plans still equal
    per-operator calls."""
    from tomosipo_amd.fused import fused_forward, fused_backward

    vg = ts.volume(shape=(48, 64, 56), size=1)
    pg = _cone_vec_helix(40, 48, 72, 1.4, 2.0, 0.1, 1e-3, seed=3)
    x = ts.phantom.ellipsoids(64, 6, seed=3)[:48, :64, :56].copy()
    A = ts.operator(vg, pg)
    assert not A.projector.autotune           # small: not tuned automatically
    xd = torch.from_numpy(x).to(DEV)
    y0 = A(xd)
    b0 = A.T(y0)
    plan = A.projector.plan(0)
    assert plan.info()["tuned"] is False
    plan.autotune(torch.cuda.current_stream().cuda_stream)
    info = plan.info()
    print(f"[autotune] {info}", flush=True)
    assert info["tuned"] is True
    assert info["fp_variant"] in (2, 3, 4)
    assert info["bp_brick_z"] in (16, 32)
    y1 = A(xd)
    b1 = A.T(y0)
    g = oracle.oracle_geometry(vg, pg)
    assert rel_l2(y1.cpu().numpy(), oracle.fp(g, x)) <= TOL
    assert rel_l2(b1.cpu().numpy(), oracle.bp(g, y0.cpu().numpy())) <= TOL
    assert rel_l2(y1.cpu().numpy(), y0.cpu().numpy()) <= 1e-6
    assert rel_l2(b1.cpu().numpy(), b0.cpu().numpy()) <= 1e-6
    # same shape, new plan: cached choices, no timing
    B = ts.operator(vg, _cone_vec_helix(40, 48, 72, 1.4, 2.0, 0.1, 1e-3, seed=4))
    pb = B.projector.plan(0)
    pb.autotune(torch.cuda.current_stream().cuda_stream)
    ib = pb.info()
    assert (ib["fp_variant"], ib["bp_brick_z"]) == (info["fp_variant"], info["bp_brick_z"])
    # fused over a tuned and an untuned plan
    C = ts.operator(vg, _cone_vec_helix(40, 48, 72, 1.4, 2.0, 0.1, 1e-3, seed=5))
    ops = [A, C, B]
    xs = [xd, xd * 0.5, xd + 1.0]
    ys = fused_forward(ops, xs)
    for op, xi, yi in zip(ops, xs, ys):
        assert torch.equal(yi, op(xi))
    bs = fused_backward(ops, ys)
    for op, yi, bi in zip(ops, ys, bs):
        assert torch.equal(bi, op.T(yi))


@pytest.mark.parametrize("case", ["rolled", "turned_rolled", "turned", "supersampled"])
def test_autotune_plans_of_every_kind(case):
    """Autotuning the plans that take the newer paths -- sheared tiles (the
    shear timed against plain tiles), waves along detector columns, the
    axis-swapped plan, the refined-detector plan: the tuned plan reports
    tuned, and its results stay within the oracle tolerance and within 1e-6
    of the untuned ones."""
    vg = ts.volume(shape=(40, 48, 44), size=(40 / 48, 1, 44 / 48))
    base = ts.cone(angles=30, shape=(44, 60), size=(1.4, 1.9), cone_angle=0.5).to_vec()
    turned = ts.rotate(pos=0, axis=(0, 1, 0), angles=np.pi / 2) * base
    pg = {"rolled": ts.rotate(pos=0, axis=(0, 1, 1), angles=0.5) * base,
          "turned_rolled": ts.rotate(pos=0, axis=(0, 1, 1), angles=0.15) * turned,
          "turned": turned, "supersampled": base}[case]
    ss = 2 if case == "supersampled" else 1
    A = ts.operator(vg, pg, detector_supersampling=ss)
    plan = A.projector.plan(0)
    x = ts.phantom.ellipsoids(48, count=6, seed=61)[:40, :48, :44].copy()
    xd = torch.from_numpy(x).to(DEV)
    y0 = A(xd)
    b0 = A.T(y0)
    plan.autotune(torch.cuda.current_stream().cuda_stream)
    info = plan.info()
    print(f"[autotune] {case}: {info}", flush=True)
    assert info["tuned"] is True
    y1 = A(xd)
    b1 = A.T(y0)
    g = oracle.oracle_geometry(vg, pg, det_supersampling=ss)
    assert rel_l2(y1.cpu().numpy(), oracle.fp(g, x)) <= TOL
    assert rel_l2(b1.cpu().numpy(), oracle.bp(g, y0.cpu().numpy())) <= TOL
    assert rel_l2(y1.cpu().numpy(), y0.cpu().numpy()) <= 1e-6
    assert rel_l2(b1.cpu().numpy(), b0.cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("swap", [1, 2])
def test_swapped_axes_projectors(swap):
    """Detector v parallel to x (swap 1: the scan turned 90 degrees about y)
    or to y (swap 2: about x): both projectors run on the plan with that axis
    and z swapped -- the forward projector on the volume moved into that
    plan's layout, the separable backprojector's result moved back through
    the epilogue. Parity with the oracle and the matched adjoint (1e-5),
    agreement with the general kernels of the same plan (forward 1e-5: other
    sample order; backward 1e-6), ADD, residual and scaled-add epilogues, a
    batch of two, and a fused launch that holds the instance."""
    import os
    from tomosipo_amd.fused import fused_forward, fused_backward

    vg = ts.volume(shape=(40, 48, 56), size=(40 / 48, 1, 56 / 48))
    base = ts.cone(angles=30, shape=(48, 72), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5).to_vec()
    axis = (0, 1, 0) if swap == 1 else (0, 0, 1)
    pg = ts.rotate(pos=0, axis=axis, angles=np.pi / 2) * base
    A = ts.operator(vg, pg)
    plan = A.projector.plan(0)
    info = plan.info()
    assert info["bp_axis_swap"] == swap and info["fp_axis_swap"] == swap and info["fp_kernel"] == "padded"
    x = ts.phantom.box_noise(vg.shape, 10, 30, 0.01, seed=21)
    _check_fp_bp(vg, pg, x)
    g = torch.Generator(device=DEV)
    g.manual_seed(22)
    y = torch.rand(A.range_shape, generator=g, device=DEV)
    xd = torch.from_numpy(x).to(DEV)
    fx = A(xd)
    b = A.T(y)
    os.environ["TSX_FORCE_GENERAL"] = "1"
    try:
        assert plan.info()["fp_axis_swap"] == 0
        fx_gen = A(xd)
        b_gen = A.T(y)
    finally:
        del os.environ["TSX_FORCE_GENERAL"]
    assert (torch.linalg.norm(fx - fx_gen) / torch.linalg.norm(fx_gen)).item() <= 1e-5
    assert (torch.linalg.norm(b - b_gen) / torch.linalg.norm(b_gen)).item() <= 1e-6
    s = torch.cuda.current_stream().cuda_stream
    accp = torch.rand(A.range_shape, generator=g, device=DEV)
    op_ = accp.clone()
    plan.forward(xd.data_ptr(), op_.data_ptr(), True, s)
    assert (torch.linalg.norm(op_ - accp - fx) / torch.linalg.norm(fx)).item() <= 1e-6
    w = torch.rand(A.range_shape, generator=g, device=DEV)
    r = torch.empty(A.range_shape, device=DEV)
    plan.forward_residual(xd.data_ptr(), y.data_ptr(), w.data_ptr(), r.data_ptr(), s)
    assert (torch.linalg.norm(r - w * (y - fx)) / torch.linalg.norm(w * (y - fx))).item() <= 1e-6
    xs = torch.rand((2,) + tuple(A.domain_shape), generator=g, device=DEV)
    fs = torch.empty((2,) + tuple(A.range_shape), device=DEV)
    plan.forward(xs.data_ptr(), fs.data_ptr(), False, s, batch=2, vol_stride=xs[0].numel(),
                 proj_stride=fs[0].numel())
    for k in range(2):
        assert torch.equal(fs[k], A(xs[k]))
    other = ts.operator(ts.volume(shape=24, size=1),
                        ts.cone(angles=20, shape=(24, 36), size=(1.4, 2.1), cone_angle=0.5))
    ops = [other, A, other]
    ins = [torch.rand(op.domain_shape, generator=g, device=DEV) for op in ops]
    outs = fused_forward(ops, ins)
    for op, xi, yi in zip(ops, ins, outs):
        assert torch.equal(yi, op(xi))
    backs = fused_backward(ops, outs)
    for op, yi, bi in zip(ops, outs, backs):
        assert torch.equal(bi, op.T(yi))
    s = torch.cuda.current_stream().cuda_stream
    acc = torch.rand(A.domain_shape, generator=g, device=DEV)
    o = acc.clone()
    plan.backward(o.data_ptr(), y.data_ptr(), True, s)
    assert (torch.linalg.norm(o - acc - b) / torch.linalg.norm(b)).item() <= 1e-6
    c = torch.rand(A.domain_shape, generator=g, device=DEV)
    v = acc.clone()
    plan.backward_scaled_add(v.data_ptr(), y.data_ptr(), c.data_ptr(), s)
    assert (torch.linalg.norm(v - (acc + c * b)) / torch.linalg.norm(c * b)).item() <= 1e-6
    ys = torch.rand((2,) + tuple(A.range_shape), generator=g, device=DEV)
    bs = torch.empty((2,) + tuple(A.domain_shape), device=DEV)
    plan.backward(bs.data_ptr(), ys.data_ptr(), False, s, batch=2, vol_stride=bs[0].numel(),
                  proj_stride=ys[0].numel())
    for k in range(2):
        assert torch.equal(bs[k], A.T(ys[k]))
