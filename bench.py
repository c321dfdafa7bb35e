#!/usr/bin/env python3
"""Benchmark: photons/s of the MI355X transport on the 384k-tet box.

Workload (one "step" = one MonteCarlo call): 40x40x40 mm box, 1 mm grid,
6 tets per cube (384,000 tets, 68,921 nodes); background mua=0.01/mm,
mus=1/mm, g=0.9, n=1.4 (n=1.0 outside); a 5 mm sphere (mua=0.05, mus=2) at
a seeded position; pencil source at (0.3, 0.2, 20) on the top face; CW;
1e8 photons per step.  The timed path is FP32 transport with FP64 tallies;
an untimed FP64 run on the same photon streams gives the accuracy check
(relative L2 of element fluence) and the FP64 throughput.

Prints one JSON line.  `value` is photons/s over the timed steps with the
mesh already resident in HBM.  `roofline` uses the walk kernel's duration
from HIP events on its own stream; `cpu_baseline` times the reference
engine (compiled from its sources into oracle/_ref/) with OpenMP on the host
cores over a bounded photon sample of the same problem (at least 30 s).
The same line carries the metric's other two parts: `batch` (256 forward
solves on the same box, 64 pencil positions x 4 optical-parameter sets, one
autotuned persistent launch replayed from a HIP graph, FP32 against FP64)
and `serve` (a Poisson stream of mixed 2D/3D, CW/100 MHz requests through
the serving loop: photons/s, p50/p99 latency, alone vs in the mix).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photons/s at <=1% fluence L2 err, batched FP32 vs plain FP64; p99 latency"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per element step: face planes + neighbour codes +
# optical parameters of the element, plus the tally read-modify-write
B_STEP = {("fp32", False): 112, ("fp32", True): 128, ("fp64", False): 192, ("fp64", True): 208}
B_STEP_2D = 96   # 2D FP32 CW: 3 edge lines + neighbour codes + optics + tally
# vector arithmetic per element step when no counted profile is committed,
# an estimate: 3-4 plane tests ~36 FLOP, move 6, deposit and weight ~6,
# scatter amortised ~6, one or two transcendentals ~6
FLOP_STEP_ESTIMATE = 60
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=float, default=None,
                    help="photons per step (default: 1e8 box, 3e8 fd, 1e6 square2d)")
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="least duration of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--refill-min", type=int, default=0)
    ap.add_argument("--replicas", type=int, default=0)
    ap.add_argument("--tally", choices=("fp64", "fixed"), default="fp64")
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32",
                    help="transport arithmetic of the timed steps (fp64: the plain FP64 path, for profiling it; "
                         "the headline is fp32)")
    ap.add_argument("--requests", type=int, default=1000, help="serve workload: number of requests")
    ap.add_argument("--serve-requests", type=int, default=500,
                    help="default line's serve part: number of requests (0: leave the part out)")
    ap.add_argument("--serve-pool-scale", type=float, default=1.0,
                    help="serve: mesh sizes of the serving pool times this (1.0: 1k-200k triangles, 50k-2M tets)")
    ap.add_argument("--batch-scale", type=float, default=1.0,
                    help="default line's batch part: photon counts times this (0: leave the part out)")
    ap.add_argument("--serve-launch-photons", type=float, default=0,
                    help="serve: photons per server launch (0: the server's default)")
    ap.add_argument("--policy", choices=("fair", "shortest"), default="shortest",
                    help="serve workload: the server's scheduling policy")
    ap.add_argument("--utilization", type=float, default=0.8,
                    help="serve workload: arrival rate as a fraction of the measured service rate")
    ap.add_argument("--workload", choices=("box", "fd", "square2d", "batch", "serve"), default="box",
                    help="box: the 384k-tet box (the headline line); batch: 256 instances "
                         "(64 pencil positions x 4 parameter sets) of the same box in one launch")
    ap.add_argument("--photon-scale", type=float, default=1.0,
                    help="batch workload: photon counts are log-uniform in [1e6, 1e8] times this "
                         "(1.0: about 5.5e9 photons per batch)")
    ap.add_argument("--autotune", action="store_true",
                    help="box/fd/square2d: pick block size, refill threshold, photons per wave fetch and "
                         "tally strategy with vmc_problem_autotune before the timed steps (tuning time "
                         "is reported, not timed)")
    ap.add_argument("--line-probe", action="store_true",
                    help="only run valomc_amd/bin/line_probe: random 128-byte line reads and 64-byte atomic line "
                         "requests per second at footprints from 2 MiB to 1 GiB, on the walk's launch shape")
    ap.add_argument("--copy-bytes", type=float, default=0,
                    help="only run a streaming kernel (out = 2 * in, float32) over this many bytes "
                         "three times: the reference dispatch that scales FETCH_SIZE / WRITE_SIZE")
    a = ap.parse_args()
    a.photons_given = a.photons is not None
    if a.photons is None:
        a.photons = 1e8
    return a


def copy_reference(nbytes: int):
    """A streaming kernel of known size: out = 2 * in over float32, three
    launches.  Each launch reads and writes `nbytes` and makes one FP32
    multiply per element, which scales FETCH_SIZE / WRITE_SIZE and shows how
    SQ_INSTS_VALU_FLOPS_FP32 counts."""
    import torch
    n = nbytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(src, 2.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel": "torch.mul float32 by 2", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "fp32_flops": n, "launches": 3}), flush=True)


TRAFFIC_PROFILES = {"box": "walk3d_fp32_box384k_hbm_traffic.txt", "fd": "walk3d_fp32fd_box2m_hbm_traffic.txt"}
FLOP_PROFILES = {"box": "walk3d_fp32_box384k_flops.txt", "fd": "walk3d_fp32fd_box2m_flops.txt",
                 "square2d": "walk2d_fp32_square200_flops.txt"}


def _profile_value(name, key):
    """The number after `key:` in a committed profile summary (None if the
    file or the line is absent)."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as fh:
            for line in fh:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


def traffic_per_step(workload="box"):
    """HBM bytes per element step of the workload's FP32 walk, from its
    committed FETCH_SIZE / WRITE_SIZE profile (None if absent)."""
    if workload not in TRAFFIC_PROFILES:
        return None
    return _profile_value(TRAFFIC_PROFILES[workload], "hbm_bytes_per_element_step")


COUNTER_PROFILES = {("box", "fp32"): "walk3d_fp32_box384k_counters.txt", ("fd", "fp32"): "walk3d_fp32fd_box2m_counters.txt",
                    ("box", "fp64"): "walk3d_fp64_box384k_counters.txt"}
# random 128-byte line reads and 64-byte FP64 atomic line requests per
# second on the walk's launch shape (profiles/line_probe_mi355x.txt): reads
# at the records' footprint (64 MiB: the box's FP32 records; 1 GiB: larger
# ones), atomics at any footprint up to 192 MiB
READ_CEILING = {("box", "fp32"): 61.44e9, ("box", "fp64"): 53.49e9, ("fd", "fp32"): 53.49e9}
ATOMIC_CEILING = 23.5e9
# reads and atomics at once, in the walk's proportions (the probe's mixed
# mode): the box's 0.38 atomic lines per read over 64 MiB of records; the FD
# box's 0.55 per read over 256 MiB, between the probe's 0.5 and 0.625 rows
MIXED_CEILING = {("box", "fp32"): (48.36e9, 18.14e9), ("fd", "fp32"): (32.5e9, 17.8e9)}


def memory_side(workload, precision, steps_per_s):
    """The walk's memory-side request rates (TCC_EA0 read and atomic
    requests per element step from its committed counter profile, times the
    measured step rate) against the probe's ceilings; None without a
    profile."""
    key = (workload, precision)
    if key not in COUNTER_PROFILES:
        return None
    name = COUNTER_PROFILES[key]
    rd, at = None, None
    try:
        with open(os.path.join(ROOT, "profiles", name)) as fh:
            for ln in fh:
                f = ln.split()
                if len(f) >= 3 and f[0] == "TCC_EA0_RDREQ_sum":
                    rd = float(f[2])
                if len(f) >= 3 and f[0] == "TCC_EA0_ATOMIC_sum":
                    at = float(f[2])
    except (OSError, ValueError):
        return None
    if rd is None or at is None:
        return None
    out = {"read_requests_per_step": rd, "atomic_requests_per_step": at,
           "read_g_per_s": round(rd * steps_per_s / 1e9, 2), "atomic_g_per_s": round(at * steps_per_s / 1e9, 2),
           "read_frac_of_probe": round(rd * steps_per_s / READ_CEILING[key], 3),
           "atomic_frac_of_probe": round(at * steps_per_s / ATOMIC_CEILING, 3),
           "source": f"profiles/{name}, profiles/line_probe_mi355x.txt"}
    if key in MIXED_CEILING:
        mr, ma = MIXED_CEILING[key]
        out["read_frac_of_mixed_probe"] = round(rd * steps_per_s / mr, 3)
        out["atomic_frac_of_mixed_probe"] = round(at * steps_per_s / ma, 3)
    return out


def flops_per_step(workload="box"):
    """Counted FP32 vector FLOPs per element step of the workload's FP32
    walk (SQ_INSTS_VALU_FLOPS_FP32 over the steps of the same dispatch),
    from its committed profile; None if absent."""
    if workload not in FLOP_PROFILES:
        return None
    return _profile_value(FLOP_PROFILES[workload], "fp32_flop_per_element_step")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(p, target_s: float, what: str = "the same problem"):
    """Time the reference engine (or, without it, the oracle) on the host:
    a pilot of at least 1 s sizes a sample of at least `target_s`."""
    # the OpenMP runtime reads these when the reference library first loads
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    kind = "reference" if pyoracle.have_reference() else "port"
    run = pyoracle.run_reference if kind == "reference" else pyoracle.run
    p.Nphoton = 20000
    while True:
        r = run(p, threads)
        if r.seconds >= 1.0 or p.Nphoton >= 10_000_000:
            break
        p.Nphoton = int(p.Nphoton * min(20.0, 1.5 / max(r.seconds, 1e-3)))
    rate = p.Nphoton / max(r.seconds, 1e-6)
    n = int(min(max(rate * target_s * 1.1, 20000), 2e8))
    p.Nphoton = n
    r = run(p, threads)
    return {"value": n / r.seconds, "unit": "photons/s", "cores": threads, "kind": kind,
            "sample": f"{n} photons of {what}, MonteCarlo() only, "
                      f"{threads} OpenMP threads (OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                      f"OMP_PLACES={os.environ.get('OMP_PLACES')}) on {cpu_model()}, {r.seconds:.1f} s"}


def batch_leg(a, photon_scale: float, steps: int, full: bool):
    """256 forward solves on the 384k-tet box (64 pencil positions x 4
    optical-parameter sets, FP32 transport, separate FP64 tallies per
    instance): launch knobs picked by vmc_batch_autotune on the batch
    itself, the batch recorded once as a HIP graph (vmc_batch_plan) and
    replayed `steps` times -- the timed figure -- with FP32 against FP64
    runs of the same streams for the first four items.  `full` adds the
    same batch launched eagerly and the items run one vmc_problem_run at a
    time.  Returns the figures."""
    import valomc_amd as vmc
    from valomc_amd import meshgen
    base_p, sources, media, photons = meshgen.batched_box_instances(seed=1234, photon_scale=photon_scale)
    base = vmc.Problem(base_p)
    meds = [base.variant(mua=mua, mus=mus) for mua, mus in media]
    probs = [[m.variant(BCType=s["BCType"], BCLDirType=s["BCLDirType"], BCLNormal=s["BCLNormal"])
              for s in sources] for m in meds]
    items = [(probs[i][j], dict(nphoton=int(photons[i, j]), seed=1000 + 64 * i + j))
             for i in range(4) for j in range(64)]
    total = int(photons.sum())
    kw = dict(precision=vmc.FP32, block_size=a.block_size, refill_min=a.refill_min, replicas=a.replicas,
              tally=vmc.TALLY_FIXED64 if a.tally == "fixed" else vmc.TALLY_FP64)
    small = [(pr, dict(nphoton=max(1, k["nphoton"] // 100), seed=k["seed"])) for pr, k in items]
    vmc.batch_run(small, **kw)   # warm-up: code loading, arena allocation
    print(f"batch: {len(items)} items, {total:.3e} photons; autotune", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    tuned = vmc.batch_autotune(items, trial_fraction=0.02 if full else 0.005, **kw)
    tuned["tuning_s"] = round(time.perf_counter() - t0, 3)
    kw.update({k: tuned[k] for k in ("block_size", "refill_min", "chunk", "replicas", "lds_slots", "cache_epoch")})
    # the batch as a HIP graph: recorded once, replayed (with a full run, a
    # warm-up replay with other seeds first)
    t0 = time.perf_counter()
    plan = vmc.BatchPlan(items, **kw)
    plan_s = time.perf_counter() - t0
    print(f"batch: knobs {tuned}; graph replays", file=sys.stderr, flush=True)
    if full:
        plan.run([(pr, dict(k, seed=k["seed"] + 7)) for pr, k in items])
    walls = []
    for i in range(steps):
        t0 = time.perf_counter()
        res, st = plan.run()
        walls.append(time.perf_counter() - t0)
        print(f"batch: replay {i + 1}/{steps} {walls[-1]:.2f} s", file=sys.stderr, flush=True)
    wall = sum(walls) / len(walls)
    plan.close()
    fig = {"instances": len(items), "photons": total, "photon_scale": photon_scale,
           "photons_per_s": round(total / wall, 1), "ms": round(wall * 1e3, 3),
           "kernel_ms": round(st.kernel_ms, 3),
           "element_steps_per_s": round(st.steps / (st.kernel_ms / 1e3), 1),
           "autotune": tuned, "graph_plan_create_s": round(plan_s, 3)}
    if full:
        # the same batch launched eagerly
        t0 = time.perf_counter()
        eres, est = vmc.batch_run(items, **kw)
        eager_wall = time.perf_counter() - t0
        print(f"batch: eager launch {eager_wall:.2f} s; one solve at a time", file=sys.stderr, flush=True)
        fig["eager_batch_photons_per_s"] = round(total / eager_wall, 1)
        fig["graph_vs_eager_rel_l2_max"] = max(float(np.linalg.norm(r1.ER - r2.ER) / np.linalg.norm(r2.ER))
                                               for r1, r2 in zip(res, eres))
        # the same items one solve at a time
        t0 = time.perf_counter()
        seq_kernel = 0.0
        seq = []
        for pr, k in items:
            r = pr.run(k["nphoton"], seed=k["seed"], **kw)
            seq_kernel += r.kernel_ms
            seq.append(r)
        seq_wall = time.perf_counter() - t0
        fig["sequential_photons_per_s"] = round(total / seq_wall, 1)
        fig["sequential_kernel_ms"] = round(seq_kernel, 3)
        fig["batch_vs_sequential_rel_l2_max"] = max(float(np.linalg.norm(r1.ER - r2.ER) / np.linalg.norm(r2.ER))
                                                    for r1, r2 in zip(res, seq))
    print("batch: FP64 runs of the first 4 items", file=sys.stderr, flush=True)
    # accuracy: FP32 batch items against FP64 runs of the same streams (first 4)
    rel = []
    for (pr, k), r in list(zip(items, res))[:4]:
        r64 = pr.run(k["nphoton"], seed=k["seed"], precision=vmc.FP64)
        rel.append(float(np.linalg.norm(r.ER - r64.ER) / np.linalg.norm(r64.ER)))
    fig["fp32_vs_fp64_rel_l2_first4"] = rel
    for row in probs:
        for pr in row:
            pr.close()
    for m in meds:
        m.close()
    base.close()
    return fig


def run_batch(a):
    """The batch workload as its own line (`--workload batch`)."""
    fig = batch_leg(a, a.photon_scale, a.steps, full=True)
    line = {
        "metric": METRIC, "value": fig["photons_per_s"], "unit": "photons/s", "n_gpus": 1,
        "steps": a.steps, "warmup": 1, "ms_per_step": fig["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "256 instances on the 384k-tet box (64 pencil positions x 4 optical-parameter "
                               "sets), one persistent launch replayed from a HIP graph, autotuned, FP32 "
                               "transport, separate FP64 tallies",
                   "instances": fig["instances"], "photons_per_step": fig["photons"],
                   "photon_scale": a.photon_scale, "parallelism": "single GPU"},
        **{k: v for k, v in fig.items() if k not in ("instances", "photons", "photon_scale", "photons_per_s", "ms")},
    }
    print(json.dumps(line), flush=True)


def serve_leg(a, nreq: int, ncal: int):
    """Poisson stream of `nreq` forward solves through the serving loop: 40%
    2D meshes of 1k-200k triangles and 60% 3D boxes of 50k-2M tetrahedra
    from a seeded pool of structured-then-jittered meshes kept resident,
    photons lognormal (median 1e7, sigma 1), CW or 100 MHz half and half.
    The rate gives about `utilization` of the service rate measured per mesh
    first, corrected on a calibration stream of `ncal` requests.  Returns
    the figures."""
    import valomc_amd as vmc
    from valomc_amd import meshgen, serving
    pool = meshgen.serving_pool(seed=2024, scale=a.serve_pool_scale)
    reqs = meshgen.serving_requests(nreq, seed=2025)
    cache = serving.MeshCache(budget_bytes=16 << 30)
    kw = dict(precision=vmc.FP32, tally=vmc.TALLY_FIXED64 if a.tally == "fixed" else vmc.TALLY_FP64)
    skw = dict(kw, launch_photons=int(a.serve_launch_photons))
    policy = vmc.SERVE_SHORTEST if a.policy == "shortest" else vmc.SERVE_FAIR
    print(f"serve: service rates of the {len(pool)} meshes", file=sys.stderr, flush=True)
    # service rate per (mesh, mode) from short single runs
    rate = {}
    for k, p in enumerate(pool):
        prob = cache.get(p)
        for f in (0.0, 100e6):
            prob.run(200000, seed=1, f=f, **kw)
            r = prob.run(2000000, seed=2, f=f, **kw)
            rate[(p.notes["kind"], k % 4, f)] = 2000000 / (r.kernel_ms / 1e3)   # device time per photon
        cache.release(prob)
    service = [q["photons"] / rate[(q["kind"], q["index"], q["f"])] for q in reqs]
    lam0 = a.utilization / float(np.mean(service))
    # requests in flight together share launches, so the device is busy for
    # less than the sum of their alone times: one correction of the rate
    # from a calibration stream (other seeds) at the estimated rate
    ncal = min(len(reqs), ncal)
    cal = [dict(q, seed=q["seed"] + 100000) for q in reqs[:ncal]]
    print(f"serve: calibration stream of {ncal} requests at {lam0:.2f}/s", file=sys.stderr, flush=True)
    with vmc.Server(arena_bytes=16 << 30, policy=policy, **skw) as srv:
        _, wcal = serving.drive(srv, cache, pool, cal, lam0, seed=11)
        busy0 = srv.counters()["busy_ms"] / 1e3 / wcal
    # the calibration requests' device time in the mix against their alone
    # service times gives the sharing factor; the rate then targets the
    # utilisation for the whole stream's alone service times
    share = float(np.sum(service[:ncal])) / max(busy0 * wcal, 1e-9)
    lam = lam0 * min(2.0, max(0.5, share))
    keep = list(range(0, len(reqs), max(1, len(reqs) // 4)))[:4]
    print(f"serve: calibration busy {busy0:.3f}, sharing {share:.3f}; {len(reqs)} requests at {lam:.2f}/s",
          file=sys.stderr, flush=True)
    with vmc.Server(arena_bytes=16 << 30, policy=policy, **skw) as srv:
        out, wall = serving.drive(srv, cache, pool, reqs, lam, seed=7, keep=keep)
        ctr = srv.counters()
    lat = np.array([o["latency_ms"] for o in out])
    total = int(sum(o["photons"] for o in out))
    # each kept request's result alone against its result in the mix
    by_kind = {"2d": [p for p in pool if p.notes["kind"] == "2d"], "3d": [p for p in pool if p.notes["kind"] == "3d"]}
    diffs = []
    for o in out:
        if "ER" not in o:
            continue
        prob = cache.get(by_kind[o["kind"]][o["index"]])
        alone = prob.run(o["photons"], seed=o["seed"], f=o["f"], **kw)
        cache.release(prob)
        diffs.append(float(np.linalg.norm(alone.ER - o["ER"]) / np.linalg.norm(alone.ER)))
    cache.close()
    return {
        "requests": len(out), "photons": total, "photons_per_s": round(total / wall, 1),
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
        "latency_ms_max": round(float(lat.max()), 2),
        "arrival_rate_per_s": round(lam, 3), "target_utilization": a.utilization,
        "calibration": {"requests": ncal, "rate_per_s": round(lam0, 3), "busy_fraction": round(busy0, 3),
                        "sharing_factor": round(share, 3)},
        "policy": a.policy, "wall_s": round(wall, 3), "gpu_busy_fraction": round(ctr["busy_ms"] / 1e3 / wall, 3),
        "launches": ctr["launches"], "alone_vs_mix_rel_l2_max": max(diffs) if diffs else None,
    }


SERVE_DESC = ("serving: Poisson stream, 40% 2D (1k-200k triangles) / 60% 3D (50k-2M tets) resident jittered "
              "meshes, photons lognormal(median 1e7, sigma 1), CW/100 MHz 50/50, FP32 transport")


def run_serve(a):
    """The serving workload as its own line (`--workload serve`)."""
    fig = serve_leg(a, a.requests, 150)
    line = {
        "metric": METRIC, "value": fig["photons_per_s"], "unit": "photons/s", "n_gpus": 1,
        "steps": fig["requests"], "warmup": 0, "ms_per_step": round(fig["wall_s"] / fig["requests"] * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": SERVE_DESC, "requests": fig["requests"], "photons_total": fig["photons"],
                   "parallelism": "single GPU"},
        **{k: v for k, v in fig.items() if k not in ("requests", "photons", "photons_per_s")},
    }
    print(json.dumps(line), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs one GPU only (asked for --gpus {a.gpus}); the multi-GPU path is a "
              "design note, not a measured configuration", file=sys.stderr)
        sys.exit(2)
    if a.copy_bytes:
        copy_reference(int(a.copy_bytes))
        return
    if a.line_probe:
        import subprocess
        exe = os.path.join(ROOT, "valomc_amd", "bin", "line_probe")
        if not os.path.exists(exe):
            sys.exit(f"{exe} not built: run `python -m valomc_amd.build probe`")
        sys.exit(subprocess.call([exe]))
    if a.workload == "batch":
        run_batch(a)
        return
    if a.workload == "serve":
        run_serve(a)
        return
    import valomc_amd as vmc
    from valomc_amd import meshgen

    # single-problem workloads: the headline box, the FD box, the 2D square
    if a.workload == "fd":
        nphot = int(a.photons) if a.photons_given else 300_000_000
        make = lambda n: meshgen.fd_box_3d(nphoton=n, seed=7)
        desc = ("2,035,200-tet jittered 60x60x40 mm box, smooth-field medium, Gaussian-profile source, "
                "100 MHz, FP32 transport + FP64 tallies")
    elif a.workload == "square2d":
        nphot = int(a.photons) if a.photons_given else 1_000_000
        make = lambda n: meshgen.square_2d(nphoton=n, seed=1)
        desc = "200-triangle 10x10 mm square, cosinic source on edges 4-7, CW, FP32 transport + FP64 tallies"
    else:
        nphot = int(a.photons)
        make = lambda n: meshgen.sphere_box_3d(nphoton=n, seed=1234)
        desc = "384k-tet heterogeneous box, pencil source, CW, FP32 transport + FP64 tallies"
    if a.precision == "fp64":
        desc = desc.replace("FP32 transport + FP64 tallies", "FP64 transport + FP64 tallies")
    p = make(nphot)
    fd = float(getattr(p, "f", 0.0) or 0.0) > 0.0
    t_create = time.perf_counter()
    prob = vmc.Problem(p)   # ErrorChecks + Init on the host, records to the device
    t_create = time.perf_counter() - t_create
    kw = dict(block_size=a.block_size, blocks_per_cu=a.blocks_per_cu, refill_min=a.refill_min,
              replicas=a.replicas, tally=vmc.TALLY_FIXED64 if a.tally == "fixed" else vmc.TALLY_FP64,
              f=float(getattr(p, "f", 0.0) or 0.0))

    prec = vmc.FP64 if a.precision == "fp64" else vmc.FP32
    # the FP64 comparison run keeps the FP64 launch defaults (FP32 launch
    # knobs, such as 1024-lane workgroups, are not valid for it)
    kw64 = dict(tally=kw["tally"], f=kw["f"])
    tuned = None
    if a.autotune:
        t = time.perf_counter()
        tuned = prob.autotune(nphot, precision=prec, **kw)
        tuned["tuning_s"] = round(time.perf_counter() - t, 3)
        kw.update({k: tuned[k] for k in ("block_size", "refill_min", "chunk", "replicas", "lds_slots", "cache_epoch")})
    for i in range(a.warmup):
        prob.run(nphot, seed=100 + i, precision=prec, **kw)
    runs = []
    t0 = time.perf_counter()
    for i in range(a.steps):
        runs.append(prob.run(nphot, seed=1 + i, precision=prec, **kw))
    t1 = time.perf_counter()
    wall = t1 - t0
    value = a.steps * nphot / wall
    kernel_s = sum(r.kernel_ms for r in runs) / 1e3
    steps_total = sum(r.steps for r in runs)
    bstep = B_STEP[(a.precision, fd)] if p.dim == 3 else B_STEP_2D
    achieved = steps_total * bstep / kernel_s / 1e9
    tps = traffic_per_step(a.workload) if a.precision == "fp32" else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None if tps is None else round(tps * steps_total / a.steps),
                "kernel": f"{'walk_small_kernel' if runs[0].lds_mesh else 'walk_kernel'}"
                          f"<{p.dim},{'double' if a.precision == 'fp64' else 'float'},"
                          f"{'FD' if fd else 'CW'},{'fixed' if a.tally == 'fixed' else 'fp64'}-tally>",
                # a small mesh is walked from records held in LDS: the
                # algorithmic bytes are then served by LDS, not by HBM
                "records_in_lds": bool(runs[0].lds_mesh),
                "bytes_per_step": bstep,
                "steps_per_launch": steps_total // a.steps,
                "kernel_ms_per_launch": round(kernel_s * 1e3 / a.steps, 3)}

    fps = flops_per_step(a.workload) if a.precision == "fp32" else None
    flop_step = fps if fps is not None else FLOP_STEP_ESTIMATE
    vpeak = FP64_PEAK_TFLOPS if a.precision == "fp64" else FP32_PEAK_TFLOPS
    vector = {"achieved": round(steps_total * flop_step / kernel_s / 1e12, 3), "peak": vpeak,
              "unit": "TFLOP/s", "flop_per_step": round(flop_step, 2),
              "flop_source": "SQ_INSTS_VALU_FLOPS_FP32 per element step (committed profile)" if fps is not None
              else "estimate"}
    vector["frac"] = round(vector["achieved"] / vpeak, 4)

    extra = {}
    if not a.no_fp64 and a.precision == "fp32":
        f32 = runs[0]
        t = time.perf_counter()
        f64 = prob.run(nphot, seed=1, precision=vmc.FP64, **kw64)
        t64 = time.perf_counter() - t
        rel = float(np.linalg.norm(f32.ER - f64.ER) / np.linalg.norm(f64.ER))
        extra = {"fp64_photons_per_s": round(nphot / t64, 1),
                 "fp32_speedup_vs_fp64": round((nphot / (wall / a.steps)) / (nphot / t64), 3),
                 "fp32_vs_fp64_rel_l2": rel,
                 "fp64_kernel_ms": round(f64.kernel_ms, 3)}
        if fd:
            extra["fp32_vs_fp64_rel_l2_EI"] = float(np.linalg.norm(f32.EI - f64.EI) / np.linalg.norm(f64.EI))
    prob.close()

    led = runs[-1].ledger
    balance = abs(led[0] + led[1] + led[2] - led[3] + led[4] - nphot) / nphot
    cpu = None
    if not a.no_cpu:
        print("cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(make(1), a.cpu_seconds, desc.split(", ")[0])
    # the metric's batched and serving parts ride on the headline line
    parts = {}
    if a.workload == "box" and a.batch_scale > 0:
        parts["batch"] = batch_leg(a, a.batch_scale, 1, full=False)
    if a.workload == "box" and a.serve_requests > 0:
        parts["serve"] = {"workload": SERVE_DESC, **serve_leg(a, a.serve_requests, 100)}

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "photons/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32", "data": "synthetic",
        "config": {"workload": desc, "elements": p.Ne, "boundary_faces": p.Nb, "nodes": p.Nr,
                   "photons_per_step": nphot, "parallelism": "single GPU"},
        "roofline": roofline,
        "vector_roofline": vector,
        "cpu_baseline": cpu,
        "element_steps_per_s": round(steps_total / kernel_s, 1),
        "memory_side": memory_side(a.workload, a.precision, steps_total / kernel_s) if p.dim == 3 else None,
        "energy_balance_rel": balance,
        "loss_per_step": runs[-1].loss,
        "rechecks_per_step": runs[-1].rechecks,
        "grid": [runs[-1].grid_blocks, runs[-1].block_size],
        "problem_create_s": round(t_create, 3),
        **extra,
        **parts,
    }
    if tuned is not None:
        line["autotune"] = tuned
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
