"""Failure handling on the GPU, forced from the host.

TSX_FAIL_LAUNCH names entry-point calls that fail before they enqueue
anything (the library's fault injector), so every kernel still runs
normally. A failed call raises with the injected message and later calls
work; in the request server a fused launch that fails is retried one request
at a time, and a request whose own launch fails again is reported failed
while the other requests finish with results bit-identical to serving them
alone.
"""
import numpy as np
import pytest

import tomosipo_amd as ts
from tomosipo_amd import _tsx, serving

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_injected_failure_raises_and_the_next_call_works(monkeypatch):
    A = ts.operator(ts.volume(shape=16, size=1), ts.parallel(angles=8, shape=(16, 24), size=(1, 1.5)))
    x = torch.rand(A.domain_shape, device="cuda")
    ref = A(x)
    monkeypatch.setenv("TSX_FAIL_LAUNCH", "2")
    y1 = A(x)                                    # call 1
    with pytest.raises(_tsx.TsxError, match="injected failure of call 2"):
        A(x)                                     # call 2
    y3 = A(x)                                    # call 3
    monkeypatch.delenv("TSX_FAIL_LAUNCH")
    assert torch.equal(y1, ref) and torch.equal(y3, ref)


def test_server_isolates_a_failed_request(monkeypatch):
    """Three FP requests and one BP request arrive together. Call 1 (the fused
    forward launch of the three) and call 2 (the retry of the first request
    alone) fail: that request is reported failed; the other two forward
    projections (calls 3, 4) and the backprojection (call 5) complete."""
    reqs = [serving.Request(rid=k, arrival=0.0, kind=("cone", "parallel")[k % 2], side=32,
                            n_angles=20 + k, op="fp", seed=k) for k in range(3)]
    reqs.append(serving.Request(rid=3, arrival=0.0, kind="cone", side=32, n_angles=24, op="bp", seed=3))
    server = serving.Server(keep_results=True)
    for r in reqs:  # device plans exist before the calls are counted
        server._operator(r).projector.plan(0)
    monkeypatch.setenv("TSX_FAIL_LAUNCH", "1,2")
    out = server.run(reqs)
    monkeypatch.delenv("TSX_FAIL_LAUNCH")
    assert out["failed"] == 1 and out["completed"] == 3
    assert "injected failure" in out["records"][0]["error"]
    assert 0 not in out["results"]
    for r in reqs[1:]:
        alone = serving.Server(keep_results=True).run([r])["results"][r.rid]
        assert torch.equal(out["results"][r.rid], alone), r


def test_server_reports_a_request_that_cannot_be_set_up():
    good = serving.Request(rid=0, arrival=0.0, kind="parallel", side=32, n_angles=20, op="fp", seed=0)
    bad = serving.Request(rid=1, arrival=0.0, kind="parallel", side=32, n_angles=0, op="fp", seed=1)
    out = serving.Server(keep_results=True).run([good, bad])
    assert out["completed"] == 1 and out["failed"] == 1
    assert "error" in out["records"][1]
    assert torch.equal(out["results"][0], serving.Server(keep_results=True).run([good])["results"][0])


def test_batches_beyond_one_launch_are_split():
    """rows x batch > 65535 (separable backprojection and the length weighting)
    and depth x batch > 65535 (the general backprojector, tilted detector):
    one batched call equals per-entry calls bitwise."""
    from tomosipo_amd.torch_support import to_autograd

    vg = ts.volume(shape=(16, 8, 8), size=(1, 0.5, 0.5))
    cases = [(ts.parallel(angles=3, shape=(512, 8), size=(1.2, 0.8)), 130),
             (ts.rotate(pos=0, axis=(1, 0, 0), angles=0.3) *
              ts.cone(angles=3, shape=(16, 8), size=(1.4, 0.9), src_orig_dist=4, src_det_dist=6).to_vec(),
              4100)]
    for pg, batch in cases:
        A = ts.operator(vg, pg)
        print(f"[split] bp kernel {A.projector.plan(0).info()['bp_kernel']}, batch {batch}", flush=True)
        y = torch.rand((batch,) + A.range_shape, device="cuda")
        b = to_autograd(A.T, num_extra_dims=1)(y)
        for k in np.linspace(0, batch - 1, 7).astype(int):
            assert torch.equal(b[k], A.T(y[k].contiguous())), k
        x = torch.rand((batch,) + A.domain_shape, device="cuda")
        f = to_autograd(A, num_extra_dims=1)(x)
        for k in np.linspace(0, batch - 1, 7).astype(int):
            assert torch.equal(f[k], A(x[k].contiguous())), k


def test_entry_point_timing(monkeypatch, capfd):
    """TSX_TIMING=1: an entry point times the work it enqueued with HIP events
    and writes one line per call to stderr; results are unchanged."""
    import re
    import tomosipo_amd as ts
    vg = ts.volume(shape=32, size=1)
    pg = ts.cone(angles=20, shape=(32, 48), size=(1.4, 2.1), cone_angle=0.5)
    A = ts.operator(vg, pg)
    x = torch.rand(A.domain_shape, device="cuda")
    y0 = A(x)
    torch.cuda.synchronize()
    capfd.readouterr()
    monkeypatch.setenv("TSX_TIMING", "1")
    y1 = A(x)
    b = A.T(y1)
    torch.cuda.synchronize()
    err = capfd.readouterr().err
    monkeypatch.delenv("TSX_TIMING")
    assert re.search(r"tsx: tsx_forward_batched \d+\.\d+ ms", err), err
    assert re.search(r"tsx: tsx_backward_batched \d+\.\d+ ms", err), err
    assert torch.equal(y0, y1)
