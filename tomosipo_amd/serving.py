"""Continuous-batching projection server.

Requests (forward projection, backprojection or an n-iteration SIRT
reconstruction, each on its own volume and geometry) arrive over time. The
server admits a request as soon as it has arrived and its device memory fits
the HBM budget, keeps its volume and projections resident, and advances every
admitted request once per tick: all requests whose next step is a forward
projection go into one fused launch, then all whose next step is a
backprojection (a SIRT request's, right after its forward step in the same
tick) into another (``tomosipo_amd.fused``). A SIRT step rides in
those launches with its elementwise work fused into the kernels' epilogues
(forward: ``r = R * (y - A x)``; backprojection: ``x += C * A.T(r)``), and the
SIRT weights R = 1 / A(1), C = 1 / A.T(1) are computed once per geometry and
shared by every request of that geometry. Requests that arrive while others
are on the device join at the next tick, so their work overlaps on the
device. Results are bit-identical to running each request alone, because the
fused kernels are deterministic and compute exactly what the per-operator
kernels compute.

Failures stay with the request that caused them: a request whose set-up
fails is reported failed at admission; if a fused launch fails, its requests
are retried one launch each, and those whose own launch fails again are
reported failed (``records[rid]["error"]``) and release their memory, while
the others go on.

``serve_one_at_a_time`` is the reference point: the same requests in arrival
order, one request at a time, each processed as tomosipo's own SIRT loop does
it (``notebooks/sirt_benchmark.py:116-139``): its own R and C, one operator
call per projection, elementwise torch operations in between.

The reference has no server; its operators are applied one synchronous call
at a time (tomosipo/Operator.py:176-233, tomosipo/astra.py:147-153).
"""
import collections
import dataclasses
import time

import numpy as np

from . import epsilon


@dataclasses.dataclass
class Request:
    """One job. ``op`` is "fp", "bp" or "sirt"; ``kind`` "cone" or "parallel"."""
    rid: int
    arrival: float          # seconds after the stream starts
    kind: str
    side: int               # volume side (voxels) of a unit-cube volume
    n_angles: int
    op: str
    iterations: int = 20    # SIRT only
    seed: int = 0


def make_stream(n_requests, rate=200.0, seed=0, sides=(64, 128, 192, 256), angles=(60, 720),
                ops=("fp", "bp", "sirt"), sirt_iterations=20):
    """Seeded Poisson stream: exponential inter-arrival times at ``rate`` per second."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
    out = []
    for k in range(n_requests):
        out.append(Request(rid=k, arrival=float(t[k]),
                           kind=("cone", "parallel")[int(rng.integers(2))],
                           side=int(sides[int(rng.integers(len(sides)))]),
                           n_angles=int(rng.integers(angles[0], angles[1] + 1)),
                           op=ops[int(rng.integers(len(ops)))], iterations=sirt_iterations,
                           seed=int(rng.integers(1 << 30))))
    return out


def request_geometry(req):
    """Unit-cube volume; detector (s, 1.5 s) of size (sqrt 2, 1.5 sqrt 2); cone
    source and detector distances from cone_angle 0.5 (SOD = SDD = 2 sqrt 2)."""
    from . import volume, cone, parallel

    s = req.side
    vg = volume(shape=s, size=1)
    size = (np.sqrt(2), 1.5 * np.sqrt(2))
    if req.kind == "cone":
        pg = cone(angles=req.n_angles, shape=(s, 3 * s // 2), size=size, cone_angle=0.5).to_vec()
    else:
        pg = parallel(angles=req.n_angles, shape=(s, 3 * s // 2), size=size).to_vec()
    return vg, pg


def request_bytes(req):
    """Device bytes a request holds while admitted (data, SIRT state, workspaces).
    The SIRT weights of its geometry (cached, shared) are counted too."""
    s = req.side
    vol = 4 * s ** 3
    sino = 4 * s * req.n_angles * (3 * s // 2)
    pad = 8 * (s + 4) ** 2 * s * 2       # row-paired padded volume copies of the forward projector
    ly = 4 * (s + 6) * req.n_angles * (3 * s // 2)  # length-weighted projections (padded rows)
    if req.op == "fp":
        return vol + sino + pad
    if req.op == "bp":
        return vol + sino + ly
    return 2 * vol + 2 * sino + ly + pad + (vol + sino)  # x, ones | y, r, l*r | C, R


def _input_data(req, torch, dev):
    g = torch.Generator(device=dev)
    g.manual_seed(req.seed)
    vg_shape = (req.side,) * 3
    sino_shape = (req.side, req.n_angles, 3 * req.side // 2)
    shape = sino_shape if req.op == "bp" else vg_shape
    x = torch.rand(shape, generator=g, device=dev)
    return x


def sirt_weights(A, torch, dev):
    """R = 1 / A(1) and C = 1 / A.T(1), clipped at 1 / epsilon (the SIRT weights
    of notebooks/sirt_benchmark.py:116-139)."""
    R = A(torch.ones(A.domain_shape, device=dev))
    R = torch.clamp(1.0 / R, max=1.0 / epsilon)
    C = A.T(torch.ones(A.range_shape, device=dev))
    C = torch.clamp(1.0 / C, max=1.0 / epsilon)
    return R, C


class _Job:
    """Device state of an admitted request and its remaining steps.

    ``weights``: a function returning the (R, C) of the request's geometry, or
    None to compute them for this request (the one-at-a-time reference)."""

    def __init__(self, req, A, torch, dev, weights=None):
        self.req = req
        self.A = A
        self.steps = collections.deque()
        data = _input_data(req, torch, dev)
        if req.op == "fp":
            self.inp, self.out = data, torch.empty(A.range_shape, device=dev)
            self.steps.append("fp")
        elif req.op == "bp":
            self.inp, self.out = data, torch.empty(A.domain_shape, device=dev)
            self.steps.append("bp")
        else:
            # SIRT on the projections of the request's phantom
            self.y = A(data)
            self.R, self.C = weights() if weights is not None else sirt_weights(A, torch, dev)
            self.x = torch.zeros(A.domain_shape, device=dev)
            self.tmp_y = torch.empty(A.range_shape, device=dev)
            self.tmp_x = None
            for _ in range(req.iterations):
                self.steps.append("fp")
                self.steps.append("bp")
        self.done_event = None

    # what the next fused launch of each kind reads and writes, and how
    def fp_io(self):
        from .fused import Epilogue

        if self.req.op == "fp":
            return self.inp, self.out, Epilogue("set")
        return self.x, self.tmp_y, Epilogue("residual", y=self.y, w=self.R)

    def bp_io(self):
        from .fused import Epilogue

        if self.req.op == "bp":
            return self.inp, self.out, Epilogue("set")
        return self.tmp_y, self.x, Epilogue("scaled_add", w=self.C)

    def result(self):
        return self.out if self.req.op in ("fp", "bp") else self.x

    def release(self):
        for name in ("inp", "out", "y", "R", "C", "x", "tmp_y", "tmp_x"):
            if hasattr(self, name):
                setattr(self, name, None)


def request_key(req):
    """Geometry key of a request: requests with equal keys share an operator."""
    return (req.kind, req.side, req.n_angles)


def request_operator(req, operators):
    """The operator of ``req`` from the cache ``operators`` (a dict keyed by
    request_key), created and cached on first use."""
    from .Operator import operator

    key = request_key(req)
    A = operators.get(key)
    if A is None:
        A = operator(*request_geometry(req))
        operators[key] = A
    return A


class Server:
    """Continuous batching with HBM-budget admission on one HIP device.

    ``operators``: a dict of operators by request_key shared with other
    servers (e.g. filled by warm_plans), so that a long-running server keeps
    the plans of the shapes it has seen; None gives the server its own.

    ``overlap`` (default): the server works in phases; a phase takes every
    active request's next step, the forward steps in one fused launch on the
    device's current stream and the backprojection steps in one fused launch on
    a side stream, running at the same time (the forward projector is bound by
    texture addressing, the backprojector by vector ALU issue), and the next
    phase starts when both have finished, so a SIRT request advances one step
    per phase. ``overlap=False``: a tick runs one fused forward launch and then
    one fused backprojection launch on the current stream that also takes the
    SIRT requests whose forward step just ran (a SIRT iteration per tick)."""

    def __init__(self, hbm_budget_bytes=16 << 30, device=0, max_active=64, keep_results=False, operators=None,
                 overlap=True):
        import torch

        self.torch = torch
        self.overlap = overlap
        self.dev = torch.device("cuda", device)
        self.budget = int(hbm_budget_bytes)
        self.max_active = max_active
        self.keep_results = keep_results
        self._ops = {} if operators is None else operators
        self._weights = {}

    def _operator(self, req):
        return request_operator(req, self._ops)

    def _sirt_weights(self, req, A):
        key = request_key(req)
        w = self._weights.get(key)
        if w is None:
            w = sirt_weights(A, self.torch, self.dev)
            self._weights[key] = w
        return w

    def _launch(self, jobs, forward, records, failed_jobs):
        """One fused launch over ``jobs``; on failure one launch per job, and the
        jobs whose own launch fails are moved to ``failed_jobs``."""
        from ._tsx import TsxError
        from .fused import fused_forward, fused_backward

        fn = fused_forward if forward else fused_backward

        def run(js):
            io = [j.fp_io() if forward else j.bp_io() for j in js]
            fn([j.A for j in js], [a for a, _, _ in io], out=[b for _, b, _ in io],
               epilogues=[e for _, _, e in io])

        try:
            run(jobs)
            return list(jobs)
        except (TsxError, RuntimeError, ValueError):
            pass
        ok = []
        for j in jobs:
            try:
                run([j])
                ok.append(j)
            except (TsxError, RuntimeError, ValueError) as e:
                records[j.req.rid]["error"] = str(e)
                failed_jobs.append(j)
        return ok

    def run(self, requests):
        """Serve ``requests`` in real time; returns per-request records and a summary."""
        torch = self.torch
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        for r in pending:
            if request_bytes(r) > self.budget:
                raise ValueError(f"request {r.rid} needs {request_bytes(r)} bytes, more than the "
                                 f"HBM budget of {self.budget} bytes")
        active, finishing, records, results = [], [], {}, {}
        ticks = collections.deque()  # events of ticks queued on the device
        used = 0
        deferred = 0
        with torch.cuda.device(self.dev):
            from .fused import _side_stream

            main = torch.cuda.current_stream(self.dev)
            side = _side_stream(self.dev) if self.overlap else None
            t0 = time.perf_counter()
            while pending or active or finishing:
                now = time.perf_counter() - t0
                # admit everything that has arrived and fits
                while pending and pending[0].arrival <= now and len(active) < self.max_active:
                    r = pending[0]
                    need = request_bytes(r)
                    if used + need > self.budget:
                        deferred += 1
                        break
                    pending.popleft()
                    records[r.rid] = {"arrival": r.arrival, "admitted": now}
                    try:
                        A = self._operator(r)
                        job = _Job(r, A, torch, self.dev, weights=lambda r=r, A=A: self._sirt_weights(r, A))
                    except Exception as e:  # a request that cannot be set up fails alone
                        records[r.rid]["error"] = f"{type(e).__name__}: {e}"
                        records[r.rid]["done"] = time.perf_counter() - t0
                        continue
                    used += need
                    active.append(job)
                # retire finished jobs whose work has completed on the device
                still = []
                for job in finishing:
                    if job.done_event.query():
                        t = time.perf_counter() - t0
                        records[job.req.rid]["done"] = t
                        used -= request_bytes(job.req)
                        if self.keep_results:
                            results[job.req.rid] = job.result()
                        job.release()
                    else:
                        still.append(job)
                finishing = still
                if not active:
                    if finishing:
                        time.sleep(0.0002)
                    elif pending:
                        time.sleep(max(0.0, min(0.002, pending[0].arrival - now)))
                    continue
                # keep at most two ticks queued ahead of the device so that new
                # arrivals join the next launch instead of queueing behind many
                while len(ticks) >= 2:
                    ticks.popleft().synchronize()
                failed = []
                fps = [j for j in active if j.steps[0] == "fp"]
                if self.overlap:
                    # one phase: the forward steps here, the backprojection steps
                    # on the side stream after everything queued before the phase
                    # (their inputs come from earlier phases or admission)
                    bps = [j for j in active if j.steps[0] == "bp"]
                    start = torch.cuda.Event()
                    start.record(main)
                    if fps:
                        for j in self._launch(fps, True, records, failed):
                            j.steps.popleft()
                    if bps:
                        with torch.cuda.stream(side):
                            side.wait_event(start)
                            for j in self._launch(bps, False, records, failed):
                                j.steps.popleft()
                            bp_done = torch.cuda.Event()
                            bp_done.record(side)
                        main.wait_event(bp_done)
                else:
                    # one tick: one fused forward launch, then one fused
                    # backprojection launch that also takes the SIRT requests
                    # whose forward step just ran (a SIRT iteration per tick)
                    if fps:
                        for j in self._launch(fps, True, records, failed):
                            j.steps.popleft()
                    bps = [j for j in active if j.steps and j.steps[0] == "bp"]
                    if bps:
                        for j in self._launch(bps, False, records, failed):
                            j.steps.popleft()
                if failed:
                    torch.cuda.synchronize()
                    t = time.perf_counter() - t0
                    for j in failed:
                        records[j.req.rid]["done"] = t
                        used -= request_bytes(j.req)
                        j.release()
                    gone = {id(j) for j in failed}
                    active = [j for j in active if id(j) not in gone]
                ev = torch.cuda.Event()
                ev.record()
                ticks.append(ev)
                keep = []
                for j in active:
                    if j.steps:
                        keep.append(j)
                    else:
                        j.done_event = ev
                        finishing.append(j)
                active = keep
            torch.cuda.synchronize()
        return _summary(records, results, deferred)


def warm_plans(requests, device=0, operators=None):
    """Create the operator and device plan of every distinct request shape in
    ``requests``, which autotunes the large ones once (the library caches the
    choices per shape), so that no served request pays for creating or tuning
    the plan of its shape. The operators go into ``operators`` (a dict keyed by
    request_key, to be passed to the servers) when given. Returns the number
    of distinct shapes."""
    import torch

    ops = {} if operators is None else operators
    seen = set()
    with torch.cuda.device(torch.device("cuda", device)):
        for r in requests:
            key = request_key(r)
            if key not in seen:
                seen.add(key)
                request_operator(r, ops).projector.plan(device)
        torch.cuda.synchronize()
    return len(seen)


def serve_one_at_a_time(requests, device=0, keep_results=False, operators=None):
    """Reference server: arrival order, one request at a time, one launch per
    operator call, SIRT as tomosipo's own loop (its own R and C; separate
    elementwise operations). ``operators`` as for Server."""
    import torch

    ops = {} if operators is None else operators
    records, results = {}, {}
    dev = torch.device("cuda", device)
    with torch.cuda.device(dev):
        t0 = time.perf_counter()
        for r in sorted(requests, key=lambda r: r.arrival):
            now = time.perf_counter() - t0
            if r.arrival > now:
                time.sleep(r.arrival - now)
            start = time.perf_counter() - t0
            job = _Job(r, request_operator(r, ops), torch, dev)
            A = job.A
            if r.op == "fp":
                A(job.inp, out=job.out)
            elif r.op == "bp":
                A.T(job.inp, out=job.out)
            else:
                tmp_x = torch.empty(A.domain_shape, device=dev)
                for _ in range(r.iterations):
                    A(job.x, out=job.tmp_y)
                    torch.sub(job.y, job.tmp_y, out=job.tmp_y)
                    job.tmp_y.mul_(job.R)
                    A.T(job.tmp_y, out=tmp_x)
                    job.x.addcmul_(job.C, tmp_x)
            torch.cuda.synchronize()
            records[r.rid] = {"arrival": r.arrival, "admitted": start,
                              "done": time.perf_counter() - t0}
            if keep_results:
                results[r.rid] = job.result()
    return _summary(records, results, 0)


def request_work(req):
    """Voxel-angle updates of a request (FP or BP: N^3 n_angles; SIRT: per
    iteration one FP and one BP, plus the three projections of its set-up)."""
    va = req.side ** 3 * req.n_angles
    if req.op in ("fp", "bp"):
        return va
    return va * (2 * req.iterations + 3)


def _summary(records, results, deferred):
    done = {k: v for k, v in records.items() if "error" not in v}
    lat = np.array([v["done"] - v["arrival"] for v in done.values()]) if done else np.zeros(1)
    first = min(v["arrival"] for v in records.values())
    last = max(v["done"] for v in records.values())
    return {
        "completed": len(done),
        "failed": len(records) - len(done),
        "req_per_s": len(done) / (last - first),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "makespan_s": last - first,
        "deferred_admissions": deferred,
        "records": records,
        "results": results,
    }
