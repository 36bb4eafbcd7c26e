"""Continuous-batching server on the GPU.

Every request's result is bit-identical alone (one request per launch) and in
the mix (fused launches with other requests); a small HBM budget forces
deferred admission without changing results.
"""
import dataclasses

import pytest

from tomosipo_amd import serving

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _stream(n=24):
    return serving.make_stream(n, rate=2000.0, seed=5, sides=(32, 48, 64), angles=(20, 60),
                               sirt_iterations=3)


def test_results_identical_alone_and_in_the_mix():
    reqs = _stream()
    mixed = serving.Server(keep_results=True).run(reqs)
    alone = serving.serve_one_at_a_time(reqs, keep_results=True)
    assert mixed["completed"] == alone["completed"] == len(reqs)
    for r in reqs:
        a, b = mixed["results"][r.rid], alone["results"][r.rid]
        assert torch.equal(a, b), f"request {r.rid} ({r.op}, {r.kind}, {r.side}^3)"
    print(f"[serving] mixed {mixed['req_per_s']:.1f} req/s, alone {alone['req_per_s']:.1f} req/s")


def test_small_budget_defers_admission():
    reqs = _stream(16)
    budget = 2 * max(serving.request_bytes(r) for r in reqs)
    out = serving.Server(hbm_budget_bytes=budget, keep_results=True).run(reqs)
    ref = serving.serve_one_at_a_time(reqs, keep_results=True)
    assert out["completed"] == len(reqs)
    assert out["deferred_admissions"] > 0
    for r in reqs:
        assert torch.equal(out["results"][r.rid], ref["results"][r.rid])


def test_warm_plans_tunes_each_shape_once():
    # two large requests (128^3 x 600 angles) are autotuned, the rest are not
    big = serving.make_stream(2, rate=2000.0, seed=6, sides=(128,), angles=(600, 600), ops=("fp",))
    reqs = _stream(12) + [dataclasses.replace(r, rid=100 + r.rid) for r in big]
    shapes = {(r.kind, r.side, r.n_angles) for r in reqs}
    assert serving.warm_plans(reqs) == len(shapes)
    out = serving.Server(keep_results=True).run(reqs)
    assert out["completed"] == len(reqs)
