"""Benchmark: GUPS of forward + back projection on 512^3 cone-beam instances.

Workload (one "step"): 16 independent cone_vec instances, each a 512^3
float32 unit-cube volume, 720 angles over 2*pi, a 512 x 768 detector of size
(sqrt 2, 1.5 sqrt 2) through the origin with source distance sqrt(2)/0.5,
the source shifted linearly in z over +-0.25 (helical-like) plus seeded
per-angle jitter sigma = 1e-3. One forward projection of every instance and
one backprojection of its result. The instances run in 8 groups of fused
launches: a group's backprojection overlaps the next group's forward
projection on a second stream (the forward projector is bound by texture
addressing, the backprojector by vector ALU issue). Inputs are resident in
device memory before timing starts.

Metric: GUPS = 2 * 512^3 * 720 * 16 / (time of one step) (voxel-angle
updates per second over FP and BP together).

The CPU baseline is the FP64 oracle (C, OpenMP) on the host cores of the same
machine, timed on a bounded subset of the angles of instance 0.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_VOX = 512
N_ANGLES = 720
DET = (512, 768)
SIZE = (np.sqrt(2), 1.5 * np.sqrt(2))
# Algorithmic per-sample work of the Joseph projector: 4 float32 corners
# (16 B) and 14 FP32 flops per bilinear sample; the exact-transpose
# backprojection has the same non-zeros. Every ray samples every slice of its
# stepping axis: samples per voxel-angle = rows * cols * N / N^3.
FLOP_PER_SAMPLE = 14.0
GATHER_B_PER_SAMPLE = 16.0
# MI355X vector FP32: 157.3 TFLOP/s published (packed FP32, two flops per lane
# per instruction slot); this build disables packed FP32, so the issue-rate
# ceiling of its kernels is half of that.
FP32_PEAK_TFLOPS = 157.3
FP32_NONPACKED_PEAK_TFLOPS = 78.6
HBM_PEAK_TBS = 8.0
# Aggregate LDS read bandwidth of the MI355X, measured with a 16-byte-per-lane
# LDS read kernel (profiles/lds_bandwidth.txt); None until measured.
LDS_MEASURED_PROFILE = os.path.join(ROOT, "profiles", "lds_bandwidth.json")
LDS_ASSUMED_TBS = 78.6  # 256 CU x 128 B/clk x 2.4 GHz (CDNA3-class LDS)


def instance_geometry(i, ts):
    base = ts.cone(angles=N_ANGLES, shape=DET, size=SIZE, cone_angle=0.5).to_vec()
    src = base.src_pos
    src[:, 0] += np.linspace(-0.25, 0.25, N_ANGLES)
    src += np.random.default_rng(1000 + i).normal(0.0, 1e-3, size=src.shape)
    pg = ts.cone_vec(shape=base.det_shape, src_pos=src, det_pos=base.det_pos, det_v=base.det_v,
                     det_u=base.det_u)
    vg = ts.volume(shape=N_VOX, size=1)
    return vg, pg


def phantom(i, torch, dev):
    """Box of ones on [128, 384)^3 plus seeded N(0, 0.01) noise, built on the device."""
    g = torch.Generator(device=dev)
    g.manual_seed(2000 + i)
    x = torch.randn((N_VOX,) * 3, generator=g, device=dev, dtype=torch.float32) * 0.01
    x[128:384, 128:384, 128:384] += 1.0
    return x


def cpu_baseline(ts, target_s=12.0):
    """FP64 oracle FP + BP on a subset of angles of instance 0; GUPS over both."""
    import oracle

    vg, pg = instance_geometry(0, ts)
    x = np.zeros((N_VOX,) * 3, dtype=np.float64)
    x[128:384, 128:384, 128:384] = 1.0
    # calibrate on one angle, then pick the subset (every k-th angle)
    g1 = oracle.oracle_geometry(vg, pg[0:1])
    t0 = time.perf_counter()
    y1 = oracle.fp(g1, x)
    oracle.bp(g1, y1)
    t1 = time.perf_counter() - t0
    n_sub = int(max(1, min(N_ANGLES, target_s / max(t1, 1e-3))))
    step = max(1, N_ANGLES // n_sub)
    sub = pg[::step]
    g = oracle.oracle_geometry(vg, sub)
    t0 = time.perf_counter()
    y = oracle.fp(g, x)
    t_fp = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.bp(g, y)
    t_bp = time.perf_counter() - t0
    na = sub.num_angles
    gups = 2 * N_VOX ** 3 * na / (t_fp + t_bp) / 1e9
    gups_fp = N_VOX ** 3 * na / t_fp / 1e9
    gups_bp = N_VOX ** 3 * na / t_bp / 1e9
    # the small parallel scan (64^3, 96 angles, 64x96) in full: one FP and one BP
    vg1, pg1 = small_parallel_geometry(ts)
    g1 = oracle.oracle_geometry(vg1, pg1)
    x1 = np.zeros(g1.vol_shape)
    x1[16:40, 16:40, 16:40] = 1.0
    t0 = time.perf_counter()
    oracle.bp(g1, oracle.fp(g1, x1))
    t_small = time.perf_counter() - t0
    return {
        "value": round(gups, 4),
        "unit": "GUPS",
        "cores": oracle.num_threads(),
        "kind": "port",
        "gups_fp": round(gups_fp, 4),
        "gups_bp": round(gups_bp, 4),
        "host": host_cpu(),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
        "sample": (f"instance 0 of the workload, every {step}th angle ({na} of {N_ANGLES}): "
                   f"one FP ({t_fp:.2f} s) and one BP ({t_bp:.2f} s) in FP64"),
        "small_parallel_full_ms": round(t_small * 1e3, 2),
    }


def host_cpu():
    """lscpu's model name and socket / core / thread counts of this host."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        out = ""
    for line in out.splitlines():
        key, _, val = line.partition(":")
        key, val = key.strip(), val.strip()
        if key in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
            info[key] = val
    info["os_cpu_count"] = os.cpu_count()
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return info


def lds_peak():
    """(TB/s, source) of the aggregate LDS bandwidth used as the gather peak."""
    if os.path.exists(LDS_MEASURED_PROFILE):
        with open(LDS_MEASURED_PROFILE) as f:
            m = json.load(f)
        return float(m["lds_read_TBps"]), "measured (profiles/lds_bandwidth.json)"
    return LDS_ASSUMED_TBS, "assumed 256 CU x 128 B/clk x 2.4 GHz"


def small_parallel_geometry(ts):
    """parallel3d_vec, 64^3 unit cube, 96 angles over pi, 64 x 96 detector of size (1, 1.5)."""
    return ts.volume(shape=64, size=1), ts.parallel(angles=96, shape=(64, 96), size=(1, 1.5)).to_vec()


def small_parallel_benchmark(torch, ts, reps=20):
    """GPU time of one FP and one BP of the small parallel scan (single operator
    calls, launch-bound)."""
    vg, pg = small_parallel_geometry(ts)
    A = ts.operator(vg, pg)
    x = torch.zeros(A.domain_shape, device="cuda")
    x[16:40, 16:40, 16:40] = 1.0
    y = torch.empty(A.range_shape, device="cuda")
    b = torch.empty(A.domain_shape, device="cuda")
    A(x, out=y)
    A.T(y, out=b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        A(x, out=y)
        A.T(y, out=b)
    torch.cuda.synchronize()
    return {"workload": "parallel3d_vec 64^3, 96 angles, 64x96: one FP + one BP",
            "ms": round((time.perf_counter() - t0) * 1e3 / reps, 3)}


def progress(msg):
    """One line on stderr per finished part of the run."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _serve_both(reqs, warm, gups_basis):
    """Both servers on the same stream, with a capacity ceiling: the stream's
    total work (voxel-angle updates, request_work) at ``gups_basis`` GUPS (the
    FP+BP rate of the fused 512^3 launches measured in this run, the projectors'
    best sustained rate; smaller requests run slower, so the ceiling is an upper
    bound), and no request served before it arrives."""
    from tomosipo_amd import serving

    # create and tune every request shape once (a long-running server keeps its
    # plans), shared by both servers; then warm both servers on a few requests
    ops = {}
    n_shapes = serving.warm_plans(list(reqs) + list(warm), operators=ops)
    progress(f"serving: {n_shapes} request shapes planned")
    serving.Server(operators=ops).run(warm)
    serving.serve_one_at_a_time(warm, operators=ops)
    # 64 GB of the 288 GB for resident requests, up to 128 at a time (on this
    # stream: 24 GB / 64 requests 49.7 req/s, 64 GB / 128 50.1, 128 GB / 256 48.0)
    batched = serving.Server(hbm_budget_bytes=64 << 30, max_active=128, operators=ops).run(reqs)
    progress(f"serving: continuous batching {batched['req_per_s']:.1f} req/s")
    single = serving.serve_one_at_a_time(reqs, operators=ops)
    progress(f"serving: one request per launch {single['req_per_s']:.1f} req/s")
    keys = ("req_per_s", "p50_latency_s", "p99_latency_s", "makespan_s")
    # least work that serves the stream: each request's own projections (a SIRT
    # request: its data term A x plus two per iteration), and the SIRT weights
    # R = 1 / A(1), C = 1 / A.T(1) once per distinct SIRT geometry
    work = 0.0
    sirt_keys = set()
    for r in reqs:
        va = float(r.side) ** 3 * r.n_angles
        if r.op == "sirt":
            work += va * (2 * r.iterations + 1)
            if serving.request_key(r) not in sirt_keys:
                sirt_keys.add(serving.request_key(r))
                work += 2 * va
        else:
            work += va
    span = max(r.arrival for r in reqs) - min(r.arrival for r in reqs)
    min_makespan = max(span, work / (gups_basis * 1e9))
    ceiling = len(reqs) / min_makespan
    return {
        "continuous_batching": {k: round(batched[k], 4) for k in keys} |
        {"deferred_admissions": batched["deferred_admissions"], "failed": batched["failed"]},
        "one_request_per_launch": {k: round(single[k], 4) for k in keys},
        "req_per_s_speedup": round(batched["req_per_s"] / single["req_per_s"], 3),
        "capacity_ceiling": {
            "work_voxel_angles": work, "gups_basis": round(gups_basis, 1),
            "min_makespan_s": round(min_makespan, 3), "req_per_s": round(ceiling, 3),
            "continuous_batching_frac": round(batched["req_per_s"] / ceiling, 3),
            "one_request_per_launch_frac": round(single["req_per_s"] / ceiling, 3),
            "two_x_one_request_per_launch_exceeds_ceiling": bool(2 * single["req_per_s"] > ceiling),
        },
    }


def serving_benchmark(n_requests, gups_basis):
    """Mixed request stream (Poisson 200 req/s; sides 64-256; 60-720 angles; cone or
    parallel; FP, BP or 20-iteration SIRT): continuous batching vs one request per launch.
    A second stream, a burst of small FP/BP requests (sides 64-128, 60-180 angles, all
    arriving within a few ms), measures capacity where requests do not fill the GPU on
    their own."""
    from tomosipo_amd import serving

    out = {"stream": f"{n_requests} requests, Poisson 200 req/s, seed 7"}
    out |= _serve_both(serving.make_stream(n_requests, rate=200.0, seed=7),
                       serving.make_stream(6, rate=1000.0, seed=8, sirt_iterations=1), gups_basis)
    small = dict(sides=(64, 96, 128), angles=(60, 180), ops=("fp", "bp"))
    out["small_requests"] = {"stream": f"{3 * n_requests} FP/BP requests, sides 64-128, 60-180 "
                                       "angles, Poisson 1e5 req/s (a burst), seed 9"}
    out["small_requests"] |= _serve_both(serving.make_stream(3 * n_requests, rate=1e5, seed=9, **small),
                                         serving.make_stream(6, rate=1000.0, seed=10, **small), gups_basis)
    return out


def fdk_benchmark(torch, ts, reps=3):
    """FDK reconstruction (ts.fdk) of one instance of the workload: 512^3, 720 angles, 512x768."""
    vg, pg = instance_geometry(0, ts)
    A = ts.operator(vg, pg)
    y = A(phantom(0, torch, "cuda"))
    out = torch.empty(A.domain_shape, device="cuda")
    ts.fdk(A, y, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ts.fdk(A, y, out=out)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"workload": "FDK of instance 0 (cone_vec 512^3, 720 angles, 512x768)",
            "ms": round(dt * 1e3, 2), "gups": round(N_VOX ** 3 * N_ANGLES / dt / 1e9, 1)}


def autograd_batch_benchmark(torch, ts, reps=10):
    """Autograd step on a batch of 8 volumes 128^3, parallel3d 180 angles, 128x192 detector:
    loss = |f(x) - y|^2 with f = to_autograd(A, num_extra_dims=1), forward + backward."""
    from tomosipo_amd.torch_support import to_autograd

    vg = ts.volume(shape=128, size=1)
    pg = ts.parallel(angles=180, shape=(128, 192), size=(1, 1.5))
    A = ts.operator(vg, pg)
    f = to_autograd(A, num_extra_dims=1)
    g = torch.Generator(device="cuda")
    g.manual_seed(4000)
    x = torch.rand((8,) + A.domain_shape, generator=g, device="cuda", requires_grad=True)
    y = torch.rand((8,) + A.range_shape, generator=g, device="cuda")

    def step():
        x.grad = None
        loss = torch.sum((f(x) - y) ** 2)
        loss.backward()

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"workload": "autograd |A x - y|^2, batch 8 x 128^3, parallel3d 180 angles, 128x192",
            "ms_per_step": round(dt * 1e3, 3),
            "gups": round(2 * 8 * 128 ** 3 * 180 / dt / 1e9, 1)}


def sirt_benchmark(torch, ts, iterations=50):
    """256^3 unit-cube volume, cone beam with 360 angles over 2 pi, 256 x 384 detector of size
    (sqrt 2, 1.5 sqrt 2), cone_angle 0.5: SIRT iterations with the fused residual / update
    kernels or as tomosipo's own loop (separate operator calls and elementwise kernels), each
    captured as a HIP graph or run eagerly."""
    from tomosipo_amd.sirt import SIRT

    vg = ts.volume(shape=256, size=1)
    pg = ts.cone(angles=360, shape=(256, 384), size=SIZE, cone_angle=0.5)
    A = ts.operator(vg, pg)
    g = torch.Generator(device="cuda")
    g.manual_seed(3000)
    x = torch.randn((256,) * 3, generator=g, device="cuda") * 0.01
    x[64:192, 64:192, 64:192] += 1.0
    y = A(x)
    res = {"workload": "SIRT, 256^3 cone, 360 angles, 256x384 detector", "iterations": iterations}
    for name, kw in (("graph_fused", dict(fused=True, graph=True)),
                     ("eager_fused", dict(fused=True, graph=False)),
                     ("graph_unfused", dict(fused=False, graph=True)),
                     ("eager_unfused", dict(fused=False, graph=False))):
        s = SIRT(A, y, fused=kw["fused"])
        s.run(1, graph=kw["graph"])          # warm-up (and graph capture)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run(iterations, graph=kw["graph"])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[name] = {"ms_per_iteration": round(dt * 1e3 / iterations, 3),
                     "gups": round(2 * 256 ** 3 * 360 * iterations / dt / 1e9, 1)}
    return res


def kernel_choices(A):
    """The autotuner's choices for an operator's device plan."""
    info = A.projector.plan(0).info()
    return {k: info[k] for k in ("fp_kernel", "fp_variant", "fp_tile_rows", "fp_slabs", "bp_kernel",
                                 "bp_brick_z", "bp_launch_angles", "tuned")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=16)
    ap.add_argument("--groups", type=int, default=8,
                    help="instance groups of the pipelined FP/BP schedule (1: FP launch, then BP)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--serving-requests", type=int, default=1000,
                    help="requests in the mixed serving stream (0 skips the serving run)")
    ap.add_argument("--no-sirt", action="store_true", help="skip the 256^3 SIRT timing")
    ap.add_argument("--no-fdk", action="store_true", help="skip the 512^3 FDK timing")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="finish with a 1 GiB device-to-device copy (reference for HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); each instance of this "
              "workload is independent, so N GPUs would run N/16 of the instances each with no "
              "communication. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import tomosipo_amd as ts

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    ops, xs, ys, bs = [], [], [], []
    for i in range(args.instances):
        vg, pg = instance_geometry(i, ts)
        A = ts.operator(vg, pg)
        ops.append(A)
        xs.append(phantom(i, torch, dev))
        ys.append(torch.empty(A.range_shape, device=dev))
        bs.append(torch.empty(A.domain_shape, device=dev))
    torch.cuda.synchronize()

    from tomosipo_amd.fused import fused_normal

    groups = max(1, min(args.groups, args.instances))

    def step(events=None, g=groups):
        # instances in `g` groups: group k's fused forward launch on this stream,
        # its fused backprojection launch on a side stream, overlapping the
        # forward launch of group k + 1 (g = 1: one FP launch, then one BP launch)
        fused_normal(ops, xs, ys, bs, groups=g, events=events)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # per-launch durations from events on the stream each launch runs on
    fp_launch = [e[2].elapsed_time(e[3]) for ev in evs for e in ev]
    bp_launch = [e[4].elapsed_time(e[5]) for ev in evs for e in ev]
    per_launch = args.instances / groups
    fp_ms = float(np.mean(fp_launch)) / per_launch
    bp_ms = float(np.mean(bp_launch)) / per_launch
    ms_per_step = wall * 1e3 / args.steps
    va = float(N_VOX) ** 3 * N_ANGLES * args.instances
    gups = 2 * va / (ms_per_step * 1e-3) / 1e9
    # the same work unpipelined (one FP launch, then one BP launch), for comparison
    seq_ms = None
    if groups > 1:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step(g=1)
        torch.cuda.synchronize()
        seq_ms = (time.perf_counter() - t0) * 1e3

    # roofline of the dominant kernel: its average launch over `per_launch` instances
    samples = float(DET[0]) * DET[1] * N_VOX * N_ANGLES * per_launch
    dom, dom_ms = ("backprojection", bp_ms * per_launch) if bp_ms >= fp_ms else \
        ("forward", fp_ms * per_launch)
    achieved_tf = samples * FLOP_PER_SAMPLE / (dom_ms * 1e-3) / 1e12
    gather_tbs = samples * GATHER_B_PER_SAMPLE / (dom_ms * 1e-3) / 1e12
    lds_tbs, lds_source = lds_peak()
    roofline = {
        # the projectors are gathers with FP32 vector arithmetic and no matrix
        # instructions: the compute bound is the vector ALU, not MFMA
        "bound": "valu",
        "achieved": round(achieved_tf, 3),
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tf / FP32_PEAK_TFLOPS, 4),
        "peak_nonpacked": FP32_NONPACKED_PEAK_TFLOPS,
        "frac_nonpacked": round(achieved_tf / FP32_NONPACKED_PEAK_TFLOPS, 4),
        "traffic": None,
        "kernel": f"fused {dom} launch over {per_launch:g} instances (incl. its pre-pass), "
                  f"average of {len(bp_launch)} launches, concurrent with the other projector",
        "kernel_ms": round(dom_ms, 3),
        "note": ("kernel_ms is the launch's duration while the other projector's launch runs "
                 "beside it on a second stream; step_achieved / step_frac count both projectors' "
                 "flops over the step time"),
        # both projectors' flops over the step time: the GPU's use while they overlap
        "step_achieved": round(2 * samples / per_launch * args.instances * FLOP_PER_SAMPLE /
                               (ms_per_step * 1e-3) / 1e12, 3),
        "basis": ("FP32 vector flops, 14 per bilinear sample, every detector ray x every slice of "
                  "its stepping axis (no MFMA: projection is a gather); peak = published packed-FP32 "
                  "rate, peak_nonpacked = the rate without packed FP32, which this build does not use"),
    }
    roofline["step_frac"] = round(roofline["step_achieved"] / FP32_PEAK_TFLOPS, 4)
    rec = {
        "metric": "GUPS (voxels x angles / s), FP+BP, 512^3 cone_vec, 720 angles, 512x768 detector",
        "value": round(gups, 2),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded box + N(0, 0.01) phantoms, helical cone_vec with seeded jitter)",
        "config": {
            "workload": f"{args.instances} x cone_vec 512^3, 720 angles, 512x768 detector, "
                        f"one FP and one BP of each, in {groups} groups of fused launches "
                        "(a group's BP overlaps the next group's FP on a second stream)",
            "instances": args.instances,
            "groups": groups,
            "sequential_ms_per_step": None if seq_ms is None else round(seq_ms, 2),
            "sequential_gups": None if seq_ms is None else round(2 * va / (seq_ms * 1e-3) / 1e9, 1),
            "fp_ms_per_instance": round(fp_ms, 3),
            "bp_ms_per_instance": round(bp_ms, 3),
            "launch_times_note": ("per-instance FP and BP times are launch durations while the two "
                                  "projectors share the GPU; their sum exceeds the step time"),
            "gups_fp": round(N_VOX ** 3 * N_ANGLES / (fp_ms * 1e-3) / 1e9, 2),
            "gups_bp": round(N_VOX ** 3 * N_ANGLES / (bp_ms * 1e-3) / 1e9, 2),
            "kernels": kernel_choices(ops[0]),
        },
        "roofline": roofline,
        "roofline_gather": {
            "achieved": round(gather_tbs, 3), "peak": round(lds_tbs, 2), "unit": "TB/s",
            "frac": round(gather_tbs / lds_tbs, 4),
            "basis": "16 B of bilinear corners per sample against aggregate LDS read bandwidth",
            "peak_source": lds_source,
        },
    }
    # HBM bytes per launch of the dominant kernel from the committed rocprofv3 counter
    # summary (FETCH_SIZE + WRITE_SIZE per instance, copy-calibrated), times instances
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        key = "bp_separable_bytes_per_instance" if dom == "backprojection" else "fp_padded_bytes_per_instance"
        if key in t:
            roofline["traffic"] = float(t[key]) * per_launch
            roofline["traffic_unit"] = "bytes per launch"
            # the kernel's HBM rate at this run's duration (bytes from the counters)
            roofline["hbm_gbps"] = round(roofline["traffic"] / (dom_ms * 1e-3) / 1e9, 1)
            roofline["hbm_frac"] = round(roofline["hbm_gbps"] / 8000.0, 4)
            roofline["traffic_source"] = ("read from profiles/hbm_traffic.json: rocprofv3 --pmc "
                                          "FETCH_SIZE + WRITE_SIZE of a one-instance run, copy-"
                                          f"calibrated, times {per_launch:g} instances; not "
                                          "measured in this run")
    progress(f"throughput workload: {gups:.1f} GUPS FP+BP")
    if args.serving_requests > 0:
        rec["serving"] = serving_benchmark(args.serving_requests, gups)
    rec["small_parallel"] = small_parallel_benchmark(torch, ts)
    rec["autograd_batch"] = autograd_batch_benchmark(torch, ts)
    progress("small parallel scan and autograd batch done")
    if not args.no_sirt:
        rec["sirt"] = sirt_benchmark(torch, ts)
        progress("SIRT done")
    if not args.no_fdk:
        rec["fdk"] = fdk_benchmark(torch, ts)
        progress("FDK done")
    if args.calibrate_copy:
        src = torch.empty(256 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    if not args.no_cpu:
        progress("CPU baseline (FP64 oracle) starting")
        rec["cpu_baseline"] = cpu_baseline(ts, args.cpu_seconds)
        rec["cpu_baseline"]["gpu_over_cpu"] = round(gups / rec["cpu_baseline"]["value"], 1)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
