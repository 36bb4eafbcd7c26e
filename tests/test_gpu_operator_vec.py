"""Operators on volume vector geometries (GPU), after tests/test_Operator.py:50-108.

For a rigid motion T of the volume (with a common scaling S of both
geometries), ts.operator(T * vg, pg) and ts.operator(vg, T^-1 * pg) compute
the same function, forward and back. The two operators reach the kernels
through different geometry arithmetic, so their FP32 outputs agree to
rounding: relative L2 difference <= 1e-5 (the reference asserts np.allclose
on ASTRA's FP32 outputs). One case is also checked against the FP64 oracle.
"""
import itertools

import numpy as np
import pytest

import oracle
import tomosipo_amd as ts
from tests.conftest import rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5
_rng = np.random.default_rng(17)

# the transform families of tests/geometry/test_transform.py:23-52, seeded
translations = [
    ts.translate((0, 0, 0)),
    ts.translate(tuple(_rng.normal(0, 1, size=3))),
]
rotations = [
    ts.rotate(pos=0, axis=(1, 0, 0), angles=0),
    ts.rotate(pos=0, axis=(1, 0, 0), angles=1.0),
    ts.rotate(pos=0, axis=_rng.normal(size=3), angles=_rng.normal()),
    ts.rotate(pos=0, axis=_rng.normal(size=3), angles=_rng.normal(), right_handed=False),
]
scalings = [
    ts.scale(1),
    ts.scale(abs(_rng.normal())),
    ts.scale(tuple(abs(_rng.normal(size=3)))),
]


def _pair(A_v, A_p, x):
    xd = torch.from_numpy(x).cuda()
    y_v, y_p = A_v(xd), A_p(xd)
    e_fp = rel_l2(y_v.cpu().numpy(), y_p.cpu().numpy())
    b_v, b_p = A_v.T(y_v), A_p.T(y_p)
    e_bp = rel_l2(b_v.cpu().numpy(), b_p.cpu().numpy())
    return e_fp, e_bp


@pytest.mark.parametrize("si, ti, ri", itertools.product(range(len(scalings)), range(len(translations)),
                                                         range(len(rotations))))
def test_operator_volume_vector(si, ti, ri):
    S, T, R = scalings[si], translations[ti], rotations[ri]
    pg = S * ts.parallel(angles=64, shape=96, size=1.5).to_vec()
    vg = S * ts.volume(shape=64, size=1)
    x = np.random.default_rng(100 + 8 * si + 4 * ti + ri).normal(size=(64, 64, 64)).astype(np.float32)
    A_v = ts.operator(T * R * vg.to_vec(), pg)
    A_p = ts.operator(vg, (T * R).inv * pg)
    print(f"[volume_vec] S{si} T{ti} R{ri}", flush=True)
    e_fp, e_bp = _pair(A_v, A_p, x)
    print(f"[volume_vec] S{si} T{ti} R{ri}: fp {e_fp:.2e} bp {e_bp:.2e}", flush=True)
    assert e_fp <= TOL and e_bp <= TOL


def test_volume_vector_with_many_steps():
    """A volume vector geometry of 48 steps against a one-step projection
    geometry: each step of the volume pairs with the single detector pose."""
    pg = ts.parallel(angles=1, shape=48).to_vec()
    vg = ts.volume(shape=32)
    angles = np.linspace(0, np.pi, 48, endpoint=False)
    R = ts.rotate(pos=0, axis=(1, 0, 0), angles=angles)
    A_v = ts.operator(R * vg.to_vec(), pg)
    A_p = ts.operator(vg, R.inv * pg)
    assert A_v.range_shape == A_p.range_shape == (48, 48, 48)
    x = np.random.default_rng(3).normal(size=A_v.domain_shape).astype(np.float32)
    e_fp, e_bp = _pair(A_v, A_p, x)
    print(f"[volume_vec] 48 steps: fp {e_fp:.2e} bp {e_bp:.2e}", flush=True)
    assert e_fp <= TOL and e_bp <= TOL
    # the rotated volume seen by a fixed detector is the fixed volume seen by
    # a rotating one: the oracle on the projection-side geometry
    g = oracle.oracle_geometry(vg, R.inv * pg)
    y_ref = oracle.fp(g, x)
    y = A_v(torch.from_numpy(x).cuda()).cpu().numpy()
    assert rel_l2(y, y_ref) <= TOL
