"""Continuous-batching server on the GPU.

Every request's result in the mix (fused launches with other requests) is
bit-identical to the same server serving that request alone; plain
projections are also bit-identical to tomosipo-style one-call-per-launch
serving, and SIRT results agree with it to 1e-5 relative L2 (the server fuses
SIRT's elementwise work into the kernels, whose multiply-add rounds once). A
small HBM budget forces deferred admission without changing results. On the
request mix of the serving benchmark (sides 64-256, 60-720 angles, cone or
parallel, FP / BP / SIRT-20), one request of each kind is also checked
against the FP64 oracle.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from tests.conftest import rel_l2
from tomosipo_amd import serving

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _stream(n=24):
    return serving.make_stream(n, rate=2000.0, seed=5, sides=(32, 48, 64), angles=(20, 60),
                               sirt_iterations=3)


def _alone(reqs):
    out = {}
    for r in reqs:
        res = serving.Server(keep_results=True).run([dataclasses.replace(r, arrival=0.0)])
        out[r.rid] = res["results"][r.rid]
    return out


def _check_mix(reqs, mixed, sirt_tol=1e-5):
    alone = _alone(reqs)
    ref = serving.serve_one_at_a_time(reqs, keep_results=True)
    assert mixed["completed"] == ref["completed"] == len(reqs)
    for r in reqs:
        a = mixed["results"][r.rid]
        assert torch.equal(a, alone[r.rid]), f"request {r.rid} ({r.op}, {r.kind}, {r.side}^3)"
        b = ref["results"][r.rid]
        if r.op == "sirt":
            assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) <= sirt_tol, r
        else:
            assert torch.equal(a, b), f"request {r.rid} ({r.op}, {r.kind}, {r.side}^3)"


@pytest.mark.parametrize("overlap", [True, False])
def test_results_identical_alone_and_in_the_mix(overlap):
    """Phases with the forward and backprojection steps on two streams
    (overlap) and ticks of one forward then one backprojection launch give
    the same results."""
    reqs = _stream()
    mixed = serving.Server(keep_results=True, overlap=overlap).run(reqs)
    _check_mix(reqs, mixed)
    print(f"[serving] mixed {mixed['req_per_s']:.1f} req/s", flush=True)


def test_small_budget_defers_admission():
    reqs = _stream(16)
    budget = 2 * max(serving.request_bytes(r) for r in reqs)
    out = serving.Server(hbm_budget_bytes=budget, keep_results=True).run(reqs)
    assert out["completed"] == len(reqs)
    assert out["deferred_admissions"] > 0
    _check_mix(reqs, out)


def test_warm_plans_tunes_each_shape_once():
    # two large requests (128^3 x 600 angles) are autotuned, the rest are not
    big = serving.make_stream(2, rate=2000.0, seed=6, sides=(128,), angles=(600, 600), ops=("fp",))
    reqs = _stream(12) + [dataclasses.replace(r, rid=100 + r.rid) for r in big]
    shapes = {(r.kind, r.side, r.n_angles) for r in reqs}
    assert serving.warm_plans(reqs) == len(shapes)
    out = serving.Server(keep_results=True).run(reqs)
    assert out["completed"] == len(reqs)


def _oracle_result(r):
    vg, pg = serving.request_geometry(r)
    g = oracle.oracle_geometry(vg, pg)
    dev = torch.device("cuda", 0)
    x = serving._input_data(r, torch, dev).double().cpu().numpy()
    if r.op == "fp":
        return oracle.fp(g, x)
    if r.op == "bp":
        return oracle.bp(g, x)
    y = oracle.fp(g, x)
    R = np.minimum(1 / np.maximum(oracle.fp(g, np.ones(g.vol_shape)), 1e-300), 1e8)
    C = np.minimum(1 / np.maximum(oracle.bp(g, np.ones(g.sino_shape)), 1e-300), 1e8)
    xr = np.zeros(g.vol_shape)
    for _ in range(r.iterations):
        xr += C * oracle.bp(g, R * (y - oracle.fp(g, xr)))
    return xr


def test_serving_mix_at_benchmark_sizes():
    """A seeded stream with the benchmark's parameters (sides 64-256, 60-720
    angles, cone or parallel, FP / BP / SIRT-20) plus one small request of each
    kind: mixed == alone bitwise; the small ones against the oracle."""
    reqs = serving.make_stream(12, rate=200.0, seed=77, sides=(64, 128, 192, 256), angles=(60, 720),
                               sirt_iterations=20)
    small = [serving.Request(rid=100 + k, arrival=reqs[-1].arrival, kind=("cone", "parallel")[k % 2],
                             side=64, n_angles=60 + 20 * k, op=op, iterations=20, seed=500 + k)
             for k, op in enumerate(("fp", "bp", "sirt"))]
    reqs = reqs + small
    print(f"[serving] {len(reqs)} requests at benchmark sizes", flush=True)
    serving.warm_plans(reqs)
    mixed = serving.Server(hbm_budget_bytes=24 << 30, keep_results=True).run(reqs)
    _check_mix(reqs, mixed)
    for r in small:
        err = rel_l2(mixed["results"][r.rid].cpu().numpy(), _oracle_result(r))
        print(f"[serving]   {r.op} {r.kind} 64^3 x {r.n_angles}: rel L2 vs oracle {err:.2e}", flush=True)
        assert err <= (1e-4 if r.op == "sirt" else 1e-5)
