"""FDK: the FP64 oracle (oracle/fdk.py) pinned by closed forms, and the API on the CPU.

The reference's FDK is ASTRA's accumulate_FDK (tomosipo/astra.py:374-406), not
available here: bit-level parity with it is unpinned. The oracle's algorithm is
pinned by:
  * the sampled Ram-Lak kernel: g[0] = 1/4, g[+-1] = -1/pi^2, even lags 0, and
    its DC response sum_m g[m] -> 0 (sum over odd m of 1/m^2 = pi^2/8);
  * uniform circular scans: every angular step is 2 pi / N;
  * reconstruction of a uniform ball (projected with the Joseph oracle): interior
    value 1 within 1 %, exterior 0, both near the parallel-beam limit (FBP) and
    at a cone angle of 0.25;
  * linearity.
"""
import numpy as np
import pytest

import oracle
from oracle import fdk as F
import tomosipo_amd as ts
from tomosipo_amd import _tsx


def test_ramp_kernel_values():
    g = F.ramp_kernel(8)
    c = 7  # lag 0
    assert g[c] == 0.25
    assert np.isclose(g[c + 1], -1 / np.pi ** 2) and np.isclose(g[c - 1], -1 / np.pi ** 2)
    assert g[c + 2] == 0 and g[c - 2] == 0
    assert np.isclose(g[c + 3], -1 / (9 * np.pi ** 2))
    big = F.ramp_kernel(20001)
    assert abs(big.sum()) < 1e-4          # DC response of the infinite kernel is 0


def test_angle_steps_uniform_circle():
    n = 37
    t = np.linspace(0, 2 * np.pi, n, endpoint=False)
    src = np.stack([3 * np.cos(t), 3 * np.sin(t), np.zeros(n)], 1)
    assert np.allclose(F.angle_steps(src, np.zeros(3)), 2 * np.pi / n)


def _ball(N, r):
    c = np.linspace(-0.5 + 0.5 / N, 0.5 - 0.5 / N, N)
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    rr = xx ** 2 + yy ** 2 + zz ** 2
    return (rr < r ** 2).astype(np.float64), rr


@pytest.mark.parametrize("cone_angle", [0.01, 0.25])
def test_oracle_fdk_reconstructs_a_uniform_ball(cone_angle):
    N = 40
    vg = ts.volume(shape=N, size=1)
    pg = ts.cone(angles=80, shape=(60, 60), size=1.5, cone_angle=cone_angle)
    g = oracle.oracle_geometry(vg, pg)
    ball, rr = _ball(N, 0.3)
    f = F.fdk(g, oracle.fp(g, ball))
    assert abs(f[rr < 0.2 ** 2].mean() - 1.0) < 0.01
    assert abs(f[rr > 0.4 ** 2].mean()) < 0.01


def test_oracle_fdk_is_linear():
    vg = ts.volume(shape=12, size=1)
    pg = ts.cone(angles=10, shape=(14, 16), size=1.5, cone_angle=0.4)
    g = oracle.oracle_geometry(vg, pg)
    rng = np.random.default_rng(0)
    y1, y2 = rng.random(g.sino_shape), rng.random(g.sino_shape)
    assert np.allclose(F.fdk(g, 2 * y1 - 3 * y2), 2 * F.fdk(g, y1) - 3 * F.fdk(g, y2), atol=1e-10)


def test_fdk_api_without_device():
    vg = ts.volume(shape=8, size=1)
    pg = ts.cone(angles=4, shape=8, size=1.5, cone_angle=0.5)
    A = ts.operator(vg, pg)
    with pytest.raises(RuntimeError):
        ts.fdk(A, np.zeros(A.range_shape, dtype=np.float32))
    with pytest.raises(ValueError):
        ts.astra.fdk(ts.data(vg), ts.data(pg), voxel_supersampling=2)
    p = _tsx.Plan(vg.to_astra(), pg.to_vec().to_astra(), device=-1)
    with pytest.raises(_tsx.TsxError) as e:
        p.fdk(None, None, False, None)
    assert e.value.code == -4
    # an upright cone: row-paired filtered projections, 8 bytes per pixel
    assert p.fdk_workspace_size() == 8 * 4 * 8 * 8
