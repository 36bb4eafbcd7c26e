"""C-ABI library: loads, exports what include/tsx.h declares, validates geometry (CPU).

No compute calls here: without a HIP device only host-only plans (device = -1)
are created.
"""
import ctypes
import os
import re
import warnings

import numpy as np
import pytest

import oracle
import tomosipo_amd as ts
from tomosipo_amd import _tsx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "tsx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tsx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _tsx.load()
    declared = _declared_functions()
    assert len(declared) >= 10
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_tsx.EXPORTED)
    assert _tsx.version().startswith("tsx")


def _plan(vg, pg, **kw):
    return _tsx.Plan(vg.to_astra(), pg.to_vec().to_astra(), device=-1, **kw)


def test_host_plan_info_and_axes():
    vg = ts.volume(shape=64, size=1)
    pg = ts.parallel(angles=96, shape=(64, 96), size=(1, 1.5))
    p = _plan(vg, pg)
    info = p.info()
    assert info["device"] == -1 and info["plan_bytes"] == 0
    assert info["vol_shape"] == (64, 64, 64)
    assert (info["n_angles"], info["det_rows"], info["det_cols"]) == (96, 64, 96)
    np.testing.assert_array_equal(p.axes(), oracle.axes(oracle.oracle_geometry(vg, pg)))


def test_kernel_selection():
    vg = ts.volume(shape=512, size=1)
    base = ts.cone(angles=720, shape=(512, 768), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    info = _plan(vg, base).info()
    assert info["fp_kernel"] == "padded" and info["bp_kernel"] == "separable"
    tilted = ts.rotate(pos=0, axis=(0, 1, 1), angles=0.2) * base.to_vec()
    assert _plan(vg, tilted).info()["bp_kernel"] == "general"
    # detector supersampling: the separable backprojector on the refined detector
    info = _plan(vg, base, det_supersampling=2).info()
    assert (info["bp_kernel"], info["bp_refined_detector"]) == ("separable", 2)
    assert _plan(vg, tilted, det_supersampling=2).info()["bp_kernel"] == "general"
    # a c1 row of the row-paired x-stepping layout of 2048^2 slices is 33.6 MB,
    # beyond 24-bit row offsets: the padded kernel with 32-bit row offsets
    wide = ts.volume(shape=(8, 2048, 2048), size=(0.004, 1, 1))
    par = ts.parallel(angles=4, shape=(8, 2048), size=(0.004, 1.5))
    assert _plan(wide, par).info()["fp_kernel"] == "padded"
    assert _plan(ts.volume(shape=(8, 1024, 1024), size=(0.008, 1, 1)),
                 ts.parallel(angles=4, shape=(8, 1024), size=(0.008, 1.5))).info()["fp_kernel"] == "padded"


def test_fast_kernels_at_1024_cubed():
    """A 1024^3 volume's row-paired layouts (8.6 GB per stepping axis) and the
    transposed projections of 720 angles of a 1024 x 1536 detector (4.6 GB)
    exceed the 32-bit offsets of the fast kernels; the plan keeps both on the
    fast kernels, the forward projector in z-slabs (the deepest that fit,
    balanced: 3 of 342 slices) and the backprojector in ranges of angles
    (678 angles of 6.3 MB each stay below 4 GiB)."""
    vg = ts.volume(shape=1024, size=1)
    pg = ts.cone(angles=720, shape=(1024, 1536), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    info = _plan(vg, pg).info()
    assert info["fp_kernel"] == "padded" and info["bp_kernel"] == "separable"
    # (d + 4) * 1024 * 1028 * 8 bytes < 2^32  ->  d <= 506  ->  3 slabs of 342
    assert info["fp_slabs"] == 3
    assert (2 ** 32 - 1) // ((1024 + 6) * 1536 * 4) == 678
    assert info["bp_launch_angles"] == 678
    # the 512^3 workload fits one slab and one launch
    small = _plan(ts.volume(shape=512, size=1),
                  ts.cone(angles=720, shape=(512, 768), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5))
    assert small.info()["fp_slabs"] == 1 and small.info()["bp_launch_angles"] == 720


def test_swapped_axes_separable_backprojector():
    """A scan whose detector v axis is parallel to x (the usual scan turned
    by 90 degrees about y) or to y (turned about x) keeps the separable
    backprojector and the padded forward projector, both run on the plan
    with that axis and z swapped; a tilted detector keeps the general
    backprojector and its own forward layouts."""
    vg = ts.volume(shape=(40, 48, 56), size=(40 / 48, 1, 56 / 48))
    pg = ts.cone(angles=30, shape=(48, 72), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5).to_vec()
    info = _plan(vg, pg).info()
    assert info["bp_axis_swap"] == info["fp_axis_swap"] == 0
    vx = ts.rotate(pos=0, axis=(0, 1, 0), angles=np.pi / 2) * pg
    info = _plan(vg, vx).info()
    assert (info["bp_kernel"], info["bp_axis_swap"], info["fp_kernel"], info["fp_axis_swap"]) == \
        ("separable", 1, "padded", 1)
    vy = ts.rotate(pos=0, axis=(0, 0, 1), angles=np.pi / 2) * pg
    info = _plan(vg, vy).info()
    assert (info["bp_kernel"], info["bp_axis_swap"], info["fp_kernel"], info["fp_axis_swap"]) == \
        ("separable", 2, "padded", 2)
    tilted = ts.rotate(pos=0, axis=(0, 1, 0), angles=1.0) * pg
    info = _plan(vg, tilted).info()
    assert (info["bp_kernel"], info["fp_kernel"], info["fp_axis_swap"]) == ("general", "padded", 0)


def test_sheared_forward_tiles():
    """Forward tiles follow the detector direction that keeps a wave's rays
    on one row of the slices: none for an upright detector (or one moved to
    the upright plan by the axis swap), rows added at the tile columns' ends
    for a detector rolled about its normal."""
    vg = ts.volume(shape=(40, 48, 56), size=(40 / 48, 1, 56 / 48))
    pg = ts.cone(angles=30, shape=(48, 72), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5).to_vec()
    assert _plan(vg, pg).info()["fp_shear_rows"] == 0
    turned = ts.rotate(pos=0, axis=(0, 1, 0), angles=np.pi / 2) * pg
    assert _plan(vg, turned).info()["fp_shear_rows"] == 0
    rolled = ts.rotate(pos=0, axis=(0, 1, 1), angles=0.2) * pg
    info = _plan(vg, rolled).info()
    # |slope| about tan(0.2)-ish over a wave of 64: a few rows each side
    assert 2 <= info["fp_shear_rows"] <= 16 and not info["fp_column_waves"], info
    # turned 90 degrees about y and rolled a little: no axis swap (v is not
    # parallel to x), the detector's rows run along the slices' rows, so the
    # waves run along its columns, sheared by a few columns
    both = ts.rotate(pos=0, axis=(0, 1, 1), angles=0.15) * turned
    info = _plan(vg, both).info()
    assert info["fp_axis_swap"] == 0 and info["fp_column_waves"] and 1 <= info["fp_shear_rows"] <= 16, info


def test_refined_detector_workspace():
    """Detector supersampling ss runs both projectors on the plan of the
    detector refined ss times: the forward workspace holds the refined
    projections (rounded up to 256 bytes) and that plan's own workspace, the
    backward workspace is that plan's (a regression check: the forward
    requirement was once the coarse plan's layouts, smaller than what the
    refined path writes)."""
    vg = ts.volume(shape=(20, 24, 28), size=(20 / 24, 1, 28 / 24))
    for ss in (2, 3):
        pg = ts.cone(angles=16, shape=(12, 18), size=(1.4, 2.1), cone_angle=0.5).to_vec()
        fine = ts.cone(angles=16, shape=(12 * ss, 18 * ss), size=(1.4, 2.1), cone_angle=0.5).to_vec()
        p, q = _plan(vg, pg, det_supersampling=ss), _plan(vg, fine)
        info = p.info()
        assert (info["fp_refined_detector"], info["bp_refined_detector"]) == (ss, ss)
        for batch in (1, 3):
            sino = 12 * ss * 16 * 18 * ss * 4 * batch
            need = (sino + 255) // 256 * 256 + q.workspace_size(_tsx.TSX_OP_FP, batch)
            assert p.workspace_size(_tsx.TSX_OP_FP, batch) == need
            assert p.workspace_size(_tsx.TSX_OP_BP, batch) == q.workspace_size(_tsx.TSX_OP_BP, batch)


def test_backprojector_angle_parts_workspace(monkeypatch):
    """The separable backprojector of a volume of fewer than 512 bricks
    (16 x 16 x 32 voxels each) sums its angles in parts of at least 4 angles,
    about 1024 blocks per launch; its workspace adds the part sums to the
    transposed projections (rounded up to 256 bytes). 64^3 x 96 angles: 32
    bricks, 24 parts; 128^3 x 180: 256 bricks, parts of 45 angles, 4 parts;
    256^3: 2048 bricks, one part. TSX_BP_PARTS=1 keeps one part."""
    cases = ((64, 96, 24), (128, 180, 4), (256, 40, 1))
    for n, na, parts in cases:
        p = _plan(ts.volume(shape=n, size=1), ts.parallel(angles=na, shape=(n, n + n // 2), size=(1, 1.5)))
        for batch in (1, 2):
            monkeypatch.setenv("TSX_BP_PARTS", "1")
            one = p.workspace_size(_tsx.TSX_OP_BP, batch)
            monkeypatch.delenv("TSX_BP_PARTS")
            ws = p.workspace_size(_tsx.TSX_OP_BP, batch)
            if parts == 1:
                assert ws == one, n
            else:
                assert ws == (one + 255) // 256 * 256 + parts * n ** 3 * 4 * batch, (n, batch)


def test_refined_forward_bound():
    """The refined forward path keeps ss^2 times the projections in its
    workspace: a 512 x 768 detector with 720 angles (1.1 GB of projections)
    takes it up to ss = 3 (10 GB) and keeps the native supersampled kernel at
    ss = 4 (18 GB of refined projections, beyond 16 GiB)."""
    vg = ts.volume(shape=64, size=1)
    pg = ts.cone(angles=720, shape=(512, 768), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    for ss, refined in ((3, 3), (4, 0)):
        info = _plan(vg, pg, det_supersampling=ss).info()
        assert info["fp_refined_detector"] == refined and info["fp_kernel"] == "padded", (ss, info)


def test_chunk_limit_override(monkeypatch):
    """TSX_CHUNK_BYTES lowers the bound at plan creation: a 48^3 cone plan
    then takes 3 slabs of 16 (row-paired x layout of a 16-slice slab:
    20 * 48 * 52 * 8 = 399360 bytes < 400000) and launches of 25 angles
    (54 * 72 * 4 = 15552 bytes per angle)."""
    vg = ts.volume(shape=48, size=1)
    pg = ts.cone(angles=40, shape=(48, 72), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    monkeypatch.setenv("TSX_CHUNK_BYTES", "400000")
    info = _plan(vg, pg).info()
    assert (info["fp_kernel"], info["fp_slabs"]) == ("padded", 3)
    assert (info["bp_kernel"], info["bp_launch_angles"]) == ("separable", 25)
    # one angle's projections beyond the bound: the general backprojector
    monkeypatch.setenv("TSX_CHUNK_BYTES", "15000")
    info = _plan(vg, pg).info()
    assert info["bp_kernel"] == "general" and info["bp_launch_angles"] == 0
    monkeypatch.delenv("TSX_CHUNK_BYTES")
    info = _plan(vg, pg).info()
    assert (info["fp_slabs"], info["bp_launch_angles"]) == (1, 40)


@pytest.mark.parametrize("bad", ["nan", "zero_u", "parallel_uv", "zero_ray"])
def test_invalid_geometry_raises_value_error(bad):
    vg = ts.volume(shape=8, size=1)
    kw = dict(shape=4, det_pos=(0, 0, 0), det_v=(1, 0, 0), det_u=(0, 0, 1))
    if bad == "nan":
        kw["det_pos"] = (np.nan, 0, 0)
    elif bad == "zero_u":
        kw["det_u"] = (0, 0, 0)
    elif bad == "parallel_uv":
        kw["det_u"] = (2, 0, 0)
    pg = ts.parallel_vec(ray_dir=(0, 0, 0) if bad == "zero_ray" else (0, 1, 0), **kw)
    with pytest.raises(ValueError):
        _plan(vg, pg)


def test_unsupported_options():
    vg = ts.volume(shape=8, size=1)
    pg = ts.parallel(angles=4, shape=8)
    with pytest.raises(ValueError, match="voxel supersampling"):
        _plan(vg, pg, vox_supersampling=2)
    with pytest.raises(ValueError, match="supersampling"):
        _plan(vg, pg, det_supersampling=0)
    with pytest.raises(ValueError, match="vector geometry"):
        _tsx.Plan(vg.to_astra(), pg.to_astra(), device=-1)


def test_compute_on_host_plan_reports_no_device():
    vg = ts.volume(shape=8, size=1)
    pg = ts.parallel(angles=4, shape=8)
    p = _plan(vg, pg)
    with pytest.raises(_tsx.TsxError) as e:
        p.forward(1, 1, False, None)
    assert e.value.code == -4
    with pytest.raises(_tsx.TsxError) as e:
        p.autotune(None)
    assert e.value.code == -4
    assert p.info()["tuned"] is False
    lib = _tsx.load()
    assert lib.tsx_direct_fpbp3d(p.handle, None, None, 1, 7, None) != 0
    # the no-allocation entry point: host plans fail before any device work
    assert lib.tsx_direct_fpbp3d_noalloc(p.handle, 1, 1, 1, 0, None) == -4
    assert lib.tsx_direct_fpbp3d_noalloc(None, 1, 1, 1, 0, None) == -1


def test_null_arguments():
    lib = _tsx.load()
    out = ctypes.c_void_p()
    assert lib.tsx_plan_create(None, -1, ctypes.byref(out)) == -1
    assert b"NULL" in lib.tsx_last_error()
    assert lib.tsx_plan_destroy(None) == 0


def test_operator_shapes_and_transpose_identity():
    vg = ts.volume(shape=(3, 10, 11))
    pg = ts.parallel(angles=7, shape=(3, 12))
    A = ts.operator(vg, pg)
    assert A.domain_shape == (3, 10, 11) and A.range_shape == (3, 7, 12)
    assert A.T is A.T.T.T
    assert A.T.domain_shape == A.range_shape and A.T.range_shape == A.domain_shape
    assert A.domain is vg and A.range is pg
    # doc/intro/lab_frame.rst: a (1, 10, 1) range
    B = ts.operator(ts.volume(shape=(1, 10, 10)), ts.parallel(angles=10, shape=(1, 1)))
    assert B.range_shape == (1, 10, 1)


def test_links_validate_shape_and_warn_on_coercion():
    vg = ts.volume(shape=4)
    with pytest.raises(ValueError):
        ts.link(vg, np.zeros((4, 4, 5), dtype=np.float32))
    with pytest.warns(UserWarning):
        ts.link(vg, np.zeros((4, 4, 4), dtype=np.float64))
    with pytest.warns(UserWarning):
        ts.link(vg, np.zeros((4, 4, 8), dtype=np.float32)[:, :, ::2])
    lk = ts.link(vg, 2.0)
    assert lk.data.shape == (4, 4, 4) and np.all(lk.data == 2.0)
    import torch

    with pytest.raises(ValueError):
        ts.link(vg, torch.zeros(4, 4, 5))
    with pytest.warns(UserWarning):
        ts.link(vg, torch.zeros(4, 4, 4, dtype=torch.float64))
    tl = ts.link(vg, torch.zeros(4, 4, 4))
    assert torch.equal(tl.new_full((4, 4, 4), 1.0).data, torch.ones(4, 4, 4))
    assert tl.new_empty((4, 4, 4)).shape == (4, 4, 4)


def test_no_device_raises_instead_of_falling_back():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    A = ts.operator(ts.volume(shape=8, size=1), ts.parallel(angles=4, shape=8))
    with pytest.raises(RuntimeError, match="HIP"):
        A(np.ones(A.domain_shape, dtype=np.float32))


def test_autograd_asserts_dimensions():
    import torch
    from tomosipo_amd.torch_support import to_autograd, AutogradOperator

    A = ts.operator(ts.volume(shape=8, size=1), ts.parallel(angles=4, shape=8))
    with pytest.raises(AssertionError):
        to_autograd(A)(torch.ones(1, 8, 8, 8))
    with pytest.raises(ValueError):
        AutogradOperator(ts.operator(ts.volume(shape=8), ts.parallel(angles=4, shape=8), additive=True))


def test_data_container():
    vg = ts.volume(shape=4)
    d = ts.data(vg, np.ones((4, 4, 4), dtype=np.float32))
    assert d.is_volume() and not d.is_projection()
    assert ts.data(vg, d) is d
    with pytest.raises(ValueError):
        ts.data(ts.volume(shape=5), d)
    c = d.clone()
    c.data[...] = 3
    assert np.all(d.data == 1)
    with pytest.raises(AttributeError):
        d.data = np.zeros((4, 4, 4))


def test_fused_calls_check_output_lists_and_shapes():
    """fused_forward / fused_backward reject an output list of the wrong length
    or with a wrongly shaped tensor before anything reaches the library."""
    import torch
    from tomosipo_amd.fused import fused_forward, fused_backward, Epilogue

    A = ts.operator(ts.volume(shape=8, size=1), ts.parallel(angles=4, shape=(8, 12)))
    B = ts.operator(ts.volume(shape=6, size=1), ts.parallel(angles=3, shape=(6, 9)))
    xa, xb = torch.zeros(A.domain_shape), torch.zeros(B.domain_shape)
    with pytest.raises(ValueError, match="one output per operator"):
        fused_forward([A, B], [xa, xb], out=[torch.zeros(A.range_shape)])
    with pytest.raises(ValueError, match="output 1"):
        fused_forward([A, B], [xa, xb], out=[torch.zeros(A.range_shape), torch.zeros(A.range_shape)])
    with pytest.raises(ValueError, match="input 0"):
        fused_backward([A], [xa])
    with pytest.raises(ValueError, match="output 0"):
        fused_backward([A], [torch.zeros(A.range_shape)], out=[torch.zeros(8, 8, 7)])
    with pytest.raises(ValueError, match="needs w"):
        fused_backward([A], [torch.zeros(A.range_shape)], epilogues=[Epilogue("scaled_add")])
    with pytest.raises(ValueError, match="forward projections"):
        fused_backward([A], [torch.zeros(A.range_shape)],
                       epilogues=[Epilogue("residual", y=xa, w=xa)])
    with pytest.raises(ValueError, match="epilogue 0 y"):
        fused_forward([A], [xa], epilogues=[Epilogue("residual", y=xa, w=torch.zeros(A.range_shape))])


def test_fused_normal_checks_arguments():
    """fused_normal rejects mismatched lists and wrongly shaped tensors before
    anything reaches the library or a stream."""
    import torch
    from tomosipo_amd.fused import fused_normal

    A = ts.operator(ts.volume(shape=8, size=1), ts.parallel(angles=4, shape=(8, 12)))
    B = ts.operator(ts.volume(shape=6, size=1), ts.parallel(angles=3, shape=(6, 9)))
    xa, xb = torch.zeros(A.domain_shape), torch.zeros(B.domain_shape)
    assert fused_normal([], []) == ([], [])
    with pytest.raises(ValueError, match="one volume per operator"):
        fused_normal([A, B], [xa])
    with pytest.raises(ValueError, match="input 1"):
        fused_normal([A, B], [xa, xa])
    with pytest.raises(ValueError, match="one projection and one output"):
        fused_normal([A, B], [xa, xb], projections=[torch.zeros(A.range_shape)])
    with pytest.raises(ValueError, match="projection 1"):
        fused_normal([A, B], [xa, xb], projections=[torch.zeros(A.range_shape)] * 2)
    with pytest.raises(ValueError, match="output 0"):
        fused_normal([A, B], [xa, xb], out=[xb, xb])
    with pytest.raises(ValueError, match="one HIP device"):
        fused_normal([A, B], [xa, xb])


def test_fused_c_abi_rejects_bad_epilogues_and_pointer_lists():
    lib = _tsx.load()
    vg, pg = ts.volume(shape=8, size=1), ts.parallel(angles=4, shape=(8, 12))
    p = _plan(vg, pg)
    hs = _tsx._ptr_array([p.handle.value])
    one = _tsx._ptr_array([1])
    eps = (_tsx.TsxEpilogue * 1)()
    eps[0].mode = _tsx.TSX_EP_RESIDUAL
    # host-only plans are rejected first; then NULL epilogue arrays
    assert lib.tsx_multi_forward_ep(hs, one, one, eps, 1, None, 0, None) == -4
    assert lib.tsx_multi_forward_ep(hs, one, one, None, 1, None, 0, None) == -1
    with pytest.raises(ValueError):
        _tsx.multi_project([p], [1, 2], [1], True, False, None)


def test_library_links_the_roctx_tracer():
    """The entry points open roctx ranges: the library links the roctx runtime."""
    import subprocess

    out = subprocess.run(["ldd", _tsx.LIB_PATH], capture_output=True, text=True).stdout
    assert "roctx" in out


def test_legacy_projection_argument_errors():
    """ts.astra.project / forward / backward raise as tomosipo's astra.py:262-268."""
    vd = ts.data(ts.volume(shape=4))
    pd = ts.data(ts.parallel(angles=3, shape=4))
    with pytest.raises(ValueError, match="forward argument"):
        ts.astra.project(vd, pd)
    with pytest.raises(ValueError, match="at least one projection dataset"):
        ts.astra.forward(vd)
    with pytest.raises(ValueError, match="at least one projection dataset"):
        ts.astra.backward(pd, pd)
    with pytest.raises(TypeError):
        ts.astra.forward(vd, pd, projector=object())


def test_dlpack_link_on_host_arrays():
    """A foreign DLPack array (tests/dlpack_producer.py, standing in for CuPy)
    gets the DLPack link: shape errors, float32 coercion with a warning, new
    arrays of the producer's own type sharing memory with their torch view,
    compatibility with torch tensors and numpy arrays on the same device."""
    import torch
    from tests import dlpack_producer as dp
    from tomosipo_amd.links.dlpack import DLPackLink
    from tomosipo_amd.links.base import are_compatible

    vg = ts.volume(shape=4)
    a = dp.Array(torch.zeros(4, 4, 4))
    lk = ts.link(vg, a)
    assert isinstance(lk, DLPackLink) and lk.data is a
    with pytest.raises(ValueError):
        ts.link(vg, dp.Array(torch.zeros(4, 4, 5)))
    with pytest.warns(UserWarning):
        c = ts.link(vg, dp.Array(torch.zeros(4, 4, 4, dtype=torch.float64)))
    assert isinstance(c.data, dp.Array) and c.tensor.dtype == torch.float32
    z = lk.new_full((2, 3, 4), 2.0)
    assert isinstance(z, DLPackLink) and isinstance(z.data, dp.Array)
    z.tensor[0, 0, 0] = 5.0
    assert z.data.to_torch()[0, 0, 0].item() == 5.0
    assert are_compatible(lk, ts.link(vg, torch.zeros(4, 4, 4)))
    assert are_compatible(lk, ts.link(vg, np.zeros((4, 4, 4), dtype=np.float32)))
    # numpy arrays and torch tensors keep their own links
    assert type(ts.link(vg, np.zeros((4, 4, 4), dtype=np.float32))).__name__ == "NumpyLink"
    assert type(ts.link(vg, torch.zeros(4, 4, 4))).__name__ == "TorchLink"
