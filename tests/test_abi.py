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
    assert _plan(vg, base, det_supersampling=2).info()["bp_kernel"] == "general"
    # a c1 row of the x-stepping layout of 2048^2 slices is 16.8 MB: beyond the
    # 24-bit row offsets of the padded kernel, so the general kernel is chosen
    wide = ts.volume(shape=(8, 2048, 2048), size=(0.004, 1, 1))
    par = ts.parallel(angles=4, shape=(8, 2048), size=(0.004, 1.5))
    assert _plan(wide, par).info()["fp_kernel"] == "general"
    assert _plan(ts.volume(shape=(8, 1024, 1024), size=(0.008, 1, 1)),
                 ts.parallel(angles=4, shape=(8, 1024), size=(0.008, 1.5))).info()["fp_kernel"] == "padded"


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
