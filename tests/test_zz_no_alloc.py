"""tsx_direct_fpbp3d_noalloc against tsx_direct_fpbp3d (GPU, through the C ABI).

The no-allocation entry point runs the same kernels with the workspace taken
from the library's static device buffer, so its results must equal those of
tsx_direct_fpbp3d bit for bit. Two calls in a row on different plans, the
second with smaller inputs (the buffer then holds the first call's larger
workspace), FP and BP, SET and ADD; the plans cover the padded forward
projector, the separable backprojector and the general backprojector on
length-weighted projections (tilted detector). A plan whose workspace
exceeds the buffer is rejected before anything is launched.
"""
import ctypes

import numpy as np
import pytest

import tomosipo_amd as ts
from tomosipo_amd import _tsx

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda"
SET, ADD = _tsx.TSX_MODE_SET, _tsx.TSX_MODE_ADD
FP, BP = _tsx.TSX_OP_FP, _tsx.TSX_OP_BP


def _cone(n, angles):
    vg = ts.volume(shape=n, size=1)
    pg = ts.cone(angles=angles, shape=(n, 3 * n // 2), size=(np.sqrt(2), 1.5 * np.sqrt(2)), cone_angle=0.5)
    return vg, pg


def _tilted(n, angles):
    vg = ts.volume(shape=n, size=1)
    pg = ts.cone(angles=angles, shape=(n, 3 * n // 2), size=(np.sqrt(2), 1.5 * np.sqrt(2)),
                 cone_angle=0.5).to_vec()
    R = ts.rotate(pos=0, axis=(0, 1, 0), angles=0.2)
    return vg, R * pg


def _calls(plan, lib, x, y0, b0, stream):
    """(FP SET, FP ADD, BP SET, BP ADD) of both entry points on copies of the same inputs."""
    out = {}
    for name, fn in (("existing", lib.tsx_direct_fpbp3d), ("noalloc", lib.tsx_direct_fpbp3d_noalloc)):
        res = []
        for op, mode in ((FP, SET), (FP, ADD), (BP, SET), (BP, ADD)):
            if op == FP:
                vol, proj = x.clone(), y0.clone()
                o = proj
            else:
                vol, proj = b0.clone(), y0.clone()
                o = vol
            rc = fn(plan.handle, ctypes.c_void_p(vol.data_ptr()), ctypes.c_void_p(proj.data_ptr()),
                    mode, op, ctypes.c_void_p(stream))
            assert rc == 0, (name, op, mode, lib.tsx_last_error().decode())
            torch.cuda.synchronize()
            res.append(o)
        out[name] = res
    return out


@pytest.mark.parametrize("shapes", [("cone", 64, 90, 32, 24), ("tilted", 40, 36, 24, 20)])
def test_noalloc_equals_direct_on_two_calls_second_smaller(shapes):
    kind, n1, a1, n2, a2 = shapes
    make = _cone if kind == "cone" else _tilted
    lib = _tsx.load()
    stream = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device=DEV)
    g.manual_seed(77)
    for n, angles in ((n1, a1), (n2, a2)):
        vg, pg = make(n, angles)
        A = ts.operator(vg, pg)
        plan = A.projector.plan(torch.cuda.current_device())
        info = plan.info()
        print(f"[noalloc] {kind} {n}^3, {angles} angles: fp_kernel {info['fp_kernel']}, "
              f"bp_kernel {info['bp_kernel']}, workspace FP {plan.workspace_size(FP)} B, "
              f"BP {plan.workspace_size(BP)} B", flush=True)
        assert 0 < max(plan.workspace_size(FP), plan.workspace_size(BP)) <= _tsx.NOALLOC_SCRATCH_BYTES
        x = torch.rand(A.domain_shape, generator=g, device=DEV)
        y0 = torch.rand(A.range_shape, generator=g, device=DEV)
        b0 = torch.rand(A.domain_shape, generator=g, device=DEV)
        out = _calls(plan, lib, x, y0, b0, stream)
        for k, (e, a) in enumerate(zip(out["existing"], out["noalloc"])):
            assert torch.equal(e, a), f"{kind} {n}^3 call {k}: max diff {(e - a).abs().max().item()}"
        # and the results are projections (not left-over inputs)
        assert not torch.equal(out["noalloc"][0], y0)
        assert not torch.equal(out["noalloc"][2], b0)


def test_noalloc_rejects_plans_beyond_the_scratch_buffer():
    vg, pg = _cone(320, 8)
    A = ts.operator(vg, pg)
    plan = A.projector.plan(torch.cuda.current_device())
    need = plan.workspace_size(FP)
    assert need > _tsx.NOALLOC_SCRATCH_BYTES
    lib = _tsx.load()
    x = torch.zeros(1, device=DEV)
    rc = lib.tsx_direct_fpbp3d_noalloc(plan.handle, ctypes.c_void_p(x.data_ptr()),
                                       ctypes.c_void_p(x.data_ptr()), SET, FP,
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -2  # TSX_ERR_UNSUPPORTED
    assert "scratch" in lib.tsx_last_error().decode()
