"""FDK on the GPU (tsx_fdk) against the FP64 oracle (oracle/fdk.py).

Tolerance: relative L2 error <= 1e-5 against the oracle (FP32 filter and
accumulation, FP64 geometry). The reference's own FDK test
(tests/test_astra.py:64-84) only runs the reconstruction; its scenario is
repeated here with the accumulate semantics checked.
"""
import numpy as np
import pytest

import oracle
from oracle import fdk as F
import tomosipo_amd as ts
from tests.conftest import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _helix(n_angles, rows, cols, seed):
    base = ts.cone(angles=n_angles, shape=(rows, cols), size=(1.4, 1.8), cone_angle=0.4).to_vec()
    src = base.src_pos
    src[:, 0] += np.linspace(-0.1, 0.1, n_angles)
    src += np.random.default_rng(seed).normal(0.0, 1e-3, size=src.shape)
    return ts.cone_vec(shape=base.det_shape, src_pos=src, det_pos=base.det_pos, det_v=base.det_v,
                       det_u=base.det_u)


@pytest.mark.parametrize("kind", ["circular", "helix", "tilted"])
def test_fdk_matches_oracle(kind):
    """circular and helix take the upright-detector kernel, tilted the general one."""
    vg = ts.volume(shape=(28, 32, 30), size=1)
    circ = ts.cone(angles=40, shape=(36, 44), size=(1.4, 1.8), cone_angle=0.4)
    pg = {"circular": circ, "helix": _helix(40, 36, 44, seed=5),
          "tilted": ts.rotate(pos=0, axis=(0, 1, 1), angles=0.2) * circ.to_vec()}[kind]
    print(f"[fdk] {kind}", flush=True)
    A = ts.operator(vg, pg)
    g = oracle.oracle_geometry(vg, pg)
    x = ts.phantom.ellipsoids(32, 6, seed=2)[:28, :32, :30].copy()
    y = oracle.fp(g, x).astype(np.float32)
    ref = F.fdk(g, y)
    out = ts.fdk(A, torch.from_numpy(y).cuda())
    assert rel_l2(out.cpu().numpy(), ref) <= 1e-5
    # numpy in -> numpy out
    out_np = ts.fdk(A, y)
    assert isinstance(out_np, np.ndarray) and rel_l2(out_np, ref) <= 1e-5


@pytest.mark.parametrize("cols", [787, 1536])
def test_fdk_wide_rows(cols):
    """Detector rows wider than one pass of the register-tiled ramp filter
    (768 outputs per pass), one of them of odd width."""
    vg = ts.volume(shape=(8, 20, 20), size=(0.4, 1, 1))
    pg = ts.cone(angles=12, shape=(10, cols), size=(0.6, 1.8), cone_angle=0.3)
    A = ts.operator(vg, pg)
    g = oracle.oracle_geometry(vg, pg)
    x = ts.phantom.box_noise((8, 20, 20), 2, 16, 0.01, seed=cols)
    y = oracle.fp(g, x).astype(np.float32)
    ref = F.fdk(g, y)
    out = ts.fdk(A, torch.from_numpy(y).cuda())
    print(f"[fdk] cols={cols} rel L2 {rel_l2(out.cpu().numpy(), ref):.2e}", flush=True)
    assert rel_l2(out.cpu().numpy(), ref) <= 1e-5


def test_fdk_reconstructs_a_ball():
    N = 64
    vg = ts.volume(shape=N, size=1)
    pg = ts.cone(angles=120, shape=(96, 96), size=1.5, cone_angle=0.25)
    A = ts.operator(vg, pg)
    c = torch.linspace(-0.5 + 0.5 / N, 0.5 - 0.5 / N, N, device="cuda")
    zz, yy, xx = torch.meshgrid(c, c, c, indexing="ij")
    rr = xx ** 2 + yy ** 2 + zz ** 2
    f = ts.fdk(A, A((rr < 0.3 ** 2).float()))
    assert abs(f[rr < 0.2 ** 2].mean().item() - 1.0) < 0.01
    assert abs(f[rr > 0.4 ** 2].mean().item()) < 0.01


def test_astra_fdk_accumulates_reference_scenario():
    # tests/test_astra.py:64-84: hollow box, cone(size=sqrt 2, cone_angle=1/2, 100 angles, 100 px)
    pg = ts.cone(size=np.sqrt(2), cone_angle=1 / 2, angles=100, shape=100)
    vg = ts.volume(shape=100)
    vd, pd = ts.data(vg), ts.data(pg)
    ts.phantom.hollow_box(vd)
    A = ts.operator(vg, pg)
    pd.data[...] = A(vd.data)
    before = vd.data.copy()
    ts.astra.fdk(vd, pd)
    rec = ts.fdk(A, pd.data)
    assert np.all(np.isfinite(vd.data))
    assert rel_l2(vd.data - before, rec) <= 1e-6


def test_fdk_out_is_overwritten():
    vg = ts.volume(shape=24, size=1)
    pg = ts.cone(angles=30, shape=(32, 32), size=1.5, cone_angle=0.5)
    A = ts.operator(vg, pg)
    y = torch.rand(A.range_shape, device="cuda")
    out = torch.full(A.domain_shape, 7.0, device="cuda")
    r = ts.fdk(A, y, out=out)
    assert r.data_ptr() == out.data_ptr()
    assert torch.allclose(out, ts.fdk(A, y))
