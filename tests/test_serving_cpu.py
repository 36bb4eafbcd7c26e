"""Request-stream and admission logic of the server (CPU)."""
import numpy as np
import pytest

from tomosipo_amd import serving


def test_stream_is_seeded_poisson():
    a = serving.make_stream(500, rate=200.0, seed=3)
    b = serving.make_stream(500, rate=200.0, seed=3)
    assert [r.__dict__ for r in a] == [r.__dict__ for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 1 / 200.0) < 0.2 / 200.0
    assert {r.side for r in a} == {64, 128, 192, 256}
    assert {r.op for r in a} == {"fp", "bp", "sirt"}
    assert {r.kind for r in a} == {"cone", "parallel"}
    assert all(60 <= r.n_angles <= 720 for r in a)


def test_request_geometry_shapes():
    r = serving.Request(rid=0, arrival=0.0, kind="cone", side=64, n_angles=90, op="fp")
    vg, pg = serving.request_geometry(r)
    assert vg.shape == (64, 64, 64) and pg.det_shape == (64, 96) and pg.num_angles == 90
    assert serving.request_bytes(r) > 4 * 64 ** 3


def test_budget_smaller_than_a_request_is_rejected():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        # the check happens before any device work
        srv = serving.Server.__new__(serving.Server)
        srv.torch, srv.budget, srv.max_active, srv.keep_results, srv._ops = torch, 1000, 4, False, {}
        srv.dev = None
        r = serving.Request(rid=0, arrival=0.0, kind="cone", side=64, n_angles=90, op="fp")
        with pytest.raises(ValueError, match="HBM budget"):
            srv.run([r])
