"""LBVH build + closest-hit benchmark on one MI355X.

Default workload (the headline): the skewed single-scene case -- 1,048,576
FP32 triangles whose centres come from 64 Gaussian clusters (seed 2, sigma
0.02, clamped to [0,1], edge offsets U[-0.002, 0.002]^3), 63-bit Morton
codes, primitive-id leaves; 1920x1080 pinhole primary rays from
(0.5, 0.5, -1) looking +z (60 degree vertical FOV, tmin 1e-4, tmax 1e30)
plus one cosine-weighted diffuse bounce ray per primary hit.  Inputs are
generated once on the host and are resident in HBM before the timed region.

One step = build the BVH (bounds, encode, sort, tree, fit) + closest-hit of
the primary rays + closest-hit of the bounce rays, all on one stream.

Prints one JSON line.  `value` is the build throughput in Mprims/s; the line
also carries closest-hit Mrays/s, the per-phase split, the roofline of the
dominant kernel, the error of the hit distance against FP64, the reference
CPU path timed on this host, and the two throughput / latency comparisons
of the metric: 4,096 scenes built and traced as one segmented batch against
the plain per-scene path, and an open-loop request stream through the
request server against the plain path one request at a time.

Other workloads (--workload): config0 (65,536 triangles, 30-bit codes, the
reference leaf order, beside the reference CPU path), batched, serving,
large (16,777,216 triangles), f64 (builder<double>), hbm-copy (a copy of
known size, the calibration of the rocprofv3 byte counters).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_PRIMS = 1 << 20
WIDTH, HEIGHT = 1920, 1080
MISS = 0xFFFFFFFF
SORT_NAMES = {1: "onesweep", 2: "reduce-then-scan", 3: "hybrid"}
METRIC = "build Mprims/s + closest-hit Mrays/s @1M tris; p99 latency batched vs plain"
# Algorithmic bytes per primitive of the build with 63-bit codes (u64 key +
# u32 id), the canonical accounting whatever sort the tuner picks: bounds 36,
# encode 48, sort 8 + 8 digit passes x 24, tree 16, fit 92.
SORT_BYTES_PER_PRIM_63 = 8 + 8 * 24
BUILD_BYTES_PER_PRIM_63 = 36 + 48 + SORT_BYTES_PER_PRIM_63 + 16 + 92
# With 30-bit codes (u32 key + u32 id): bounds 36, encode 44, sort 4 + 4
# digit passes x 16, tree 12, fit 92.
BUILD_BYTES_PER_PRIM_30 = 36 + 44 + (4 + 4 * 16) + 12 + 92
# Bytes the closest-hit traversal touches on the benchmark workload (primary
# + bounce rays on the 1M clustered tree): popped nodes x 32 + box-tested
# children x 32 + tested triangles x 48, counted by the instrumented CPU
# traversal of the same tree (tests/test_workloads.py recomputes it).
TRACE_TOUCHED_BYTES_PER_STEP = 3857317376
RAY_COMPULSORY_BYTES = 48  # 32-byte ray in, 16-byte hit out
# HBM bytes per k_trace dispatch on this workload, averaged over the primary
# and bounce dispatches: rocprofv3 --pmc FETCH_SIZE (x 2.0 KiB) + WRITE_SIZE
# (x 1.0 KiB), scaled by a 512 MiB copy kernel
# (profiles/clustered1m_63bit_hbm_traffic.md: primary 719.1 + 47.6 MB, bounce
# 903.2 + 16.9 MB; mean 811.2 MB read + 32.3 MB written).
TRACE_HBM_TRAFFIC_PER_LAUNCH = 843.4e6
# Offered load of the serving comparison in the default line (requests/s):
# above the plain path's capacity, so both paths run saturated.
HEADLINE_SERVING_RATE = 4000.0
HEADLINE_SERVING_REQUESTS = 160


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="single: skip the batched and served comparisons; batched: skip the "
                         "plain per-scene path (profiling runs)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget of the CPU-baseline sample")
    ap.add_argument("--workload",
                    choices=["single", "config0", "batched", "serving", "large", "f64",
                             "hbm-copy"],
                    default="single",
                    help="single: the 1M-triangle scene (default); config0: 65,536 uniform "
                         "triangles, 30-bit codes, reference leaf order, 256x256 rays, beside the "
                         "reference CPU path; batched: 4,096 uneven scenes built in one "
                         "segmented pass against the plain per-scene path; serving: an open-loop "
                         "request stream through the request server against the plain path one "
                         "request at a time; large: 16,777,216 triangles, per-phase 30/63-bit "
                         "builds and a refit; f64: builder<double> over the 1M scene in FP64, and "
                         "16,777,216 boxes; hbm-copy: 5 copies of 512 MiB (counter calibration)")
    ap.add_argument("--scenes", type=int, default=4096, help="batched: number of scenes")
    ap.add_argument("--requests", type=int, default=160, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=1000.0,
                    help="serving: offered load, requests per second (Poisson)")
    return ap.parse_args()


def host_info():
    """CPU model, usable cores and SMT state of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    smt = None
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = f.read().strip() == "1"
    except OSError:
        pass
    tpc = None
    try:
        with open("/sys/devices/system/cpu/cpu0/topology/thread_siblings_list") as f:
            sib = f.read().strip()
        n = 0
        for part in sib.split(","):
            a, _, b = part.partition("-")
            n += (int(b) - int(a) + 1) if b else 1
        tpc = n
    except (OSError, ValueError):
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "usable_cores": usable, "logical_cpus": os.cpu_count(),
            "smt_active": smt, "threads_per_core": tpc}


def _ref_threads(fast):
    import oracle
    return oracle.ref(fast).ref_hardware_concurrency() or os.cpu_count() or 1


def cpu_baseline(tri, rays, bounce, budget_s):
    """The real lbvh.h (oracle/_ref, built from the reference sources) timed
    on this host: builds of the same scene, then the reference traverser on
    a sample of the same rays; falls back to the oracle restatement."""
    import numpy as np
    import oracle
    kind = "reference"
    fast = oracle.ref_available(fast=True)
    if not (fast or oracle.ref_available()):
        kind = "port"
    threads = os.cpu_count() or 1
    if kind == "reference":
        threads = _ref_threads(fast)
    t_end = time.time() + budget_s * 0.6
    times = []
    while time.time() < t_end or len(times) < 2:
        t0 = time.perf_counter()
        if kind == "reference":
            oracle.ref_build(tri, 63, 0, threads=threads, fast=fast)
        else:
            oracle.build(tri, 63, 0)
        times.append(time.perf_counter() - t0)
        if len(times) >= 6:
            break
    build_s = sorted(times)[len(times) // 2]
    # traversal sample on the primitive-id tree (the reference-order tree is
    # near brute force for this scene: hours of CPU)
    if kind == "reference":
        pid_nodes, _, _ = oracle.ref_build(tri, 63, 1, threads=threads, fast=fast)
    else:
        pid_nodes, _, _, _ = oracle.build(tri, 63, 1)
    sample = np.concatenate([rays[::16], bounce[::16]])
    t0 = time.perf_counter()
    if kind == "reference":
        oracle.ref_trace(pid_nodes, tri, sample, 0, threads=threads)
    else:
        oracle.trace(pid_nodes, tri, sample, 0, threads=threads)
    trace_s = time.perf_counter() - t0
    out = {
        "value": round(N_PRIMS / build_s / 1e6, 3),
        "unit": "Mprims/s",
        "cores": int(threads),
        "kind": kind,
        "flags": ("-O3 -ffast-math" if (kind == "reference" and fast) else "-O2 -ffp-contract=off"),
        "sample": "%d builds of the 1,048,576-triangle scene (63-bit codes, median %.1f ms); "
                  "closest hit of %d rays (every 16th primary and bounce ray) on the "
                  "primitive-id tree in %.2f s" % (len(times), build_s * 1e3, len(sample), trace_s),
        "build_ms": round(build_s * 1e3, 2),
        "closest_hit_mrays_s": round(len(sample) / trace_s / 1e6, 4),
    }
    out.update(host_info())
    return out


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def _p(xs, q, nd=4):
    v = percentile(xs, q)
    return None if v is None else round(v, nd)


def t_error_vs_fp64(tri, rays, prim, t):
    """Error of the FP32 hit distances against the plane distance in FP64:
    t64 = dot(n, p0 - o) / dot(n, d) with n = e0 x e1, all in FP64, on the
    stored triangle (p0, e0, e1) of the hit primitive.  Relative error and
    error in units of the FP32 spacing at t64, max and mean over the hits."""
    import numpy as np
    hit = prim != MISS
    if not hit.any():
        return {"hits": 0}
    tr = tri[prim[hit].astype(np.int64)].astype(np.float64)
    r = rays[hit].astype(np.float64)
    n = np.cross(tr[:, 3:6], tr[:, 6:9])
    c = tr[:, 0:3] - r[:, 0:3]
    t64 = (n * c).sum(axis=1) / (n * r[:, 3:6]).sum(axis=1)
    t32 = t[hit].astype(np.float64)
    err = np.abs(t32 - t64)
    rel = err / np.abs(t64)
    ulp = err / np.spacing(np.abs(t64).astype(np.float32)).astype(np.float64)
    return {"hits": int(hit.sum()), "max_rel": float("%.4g" % rel.max()),
            "mean_rel": float("%.4g" % rel.mean()), "max_ulp": float("%.4g" % ulp.max()),
            "mean_ulp": float("%.4g" % ulp.mean())}


# ---- batched vs plain ---------------------------------------------------------

BATCHED_DESC = ("%d scenes (seed 4), triangle counts log-uniform in [64, 262144], uniform scenes "
                "in their own unit cubes, 30-bit codes, primitive-id leaves; 1,024 box-surface "
                "rays per scene")


def batched_compare(nseg, steps, warmup, plain_path=True):
    """4,096 independent scenes (seed 4) with triangle counts log-uniform in
    [64, 262144], each a uniform scene in its own unit cube (offsets
    U[-0.01, 0.01]^3), 30-bit codes, primitive-id leaves; 1,024 rays per scene
    from the scene's bounding-box surface in directions uniform on the sphere.

    Batched step: one segmented rebuild of every scene (replayed from a HIP
    graph) + one closest-hit launch over every scene's rays.  Plain path: the
    same scenes built and traced one call at a time through the single-scene
    entry points (a fresh tree per scene, as the plain API returns), once
    queued back to back and once with a host synchronization per scene.

    Latency: every scene is requested at once.  A batched step completes all
    of them together, so each scene's latency is the step's host wall time
    (distribution over the timed steps); served one at a time, scene s
    completes after every scene before it (distribution over the scenes).
    Both produce identical hits (checked)."""
    import numpy as np

    import lbvh_amd
    from lbvh_amd import hip
    from lbvh_amd import workload as W

    log("bench: generating %d scenes (seed 4, log-uniform [64, 262144])" % nseg)
    sizes = W.batched_scene_sizes(4, nseg, 64, 262144)
    tri, off = W.batched_scenes(4, sizes=sizes)
    rays, roff = W.batched_box_rays(4, tri, off, 1024)
    n, m = int(off[-1]), rays.shape[0]
    log("bench: %d triangles, %d rays" % (n, m))
    ctx = lbvh_amd.Context(0)
    d_tri = hip.DeviceArray.from_host(tri)
    d_rays = hip.DeviceArray.from_host(rays)
    d_roff = hip.DeviceArray.from_host(roff)
    out = (hip.DeviceArray(m, np.uint32), hip.DeviceArray(m, np.float32),
           hip.DeviceArray((m, 2), np.float32))
    PID = lbvh_amd.LEAF_PRIMITIVE_ID
    mode = lbvh_amd.ISECT_REFERENCE
    del tri

    t0 = time.perf_counter()
    b = ctx.build_batch(d_tri, off, leaf_order=PID)
    ctx.synchronize()
    first_build_s = time.perf_counter() - t0
    info = b.info()
    tune_log = ctx.tune_log()
    log("bench: batch built (first call incl. allocation and tuning: %.1f ms); sort %d-bit "
        "digits, tile %d, %d passes" % (first_build_s * 1e3, info.sort_digit_bits, info.sort_tile,
                                        info.sort_passes))
    for _ in range(warmup):
        b.rebuild(graph=True)
        b.trace(d_rays, d_roff, mode, out=out)
    ctx.synchronize()
    # device time of back-to-back steps
    ctx.set_timing(2)
    hip.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.rebuild(graph=True)
        b.trace(d_rays, d_roff, mode, out=out)
    hip.synchronize()
    wall_b = (time.perf_counter() - t0) / steps
    ctx.synchronize()
    bt, _ = ctx.phase_stats(1)
    tt, _ = ctx.phase_stats(2)
    ctx.set_timing(0)
    build_ms = sum(bt) / steps
    trace_ms = sum(tt) / steps
    # latency: each step submitted alone and waited for
    lat_b = []
    for _ in range(steps):
        ts = time.perf_counter()
        b.rebuild(graph=True)
        b.trace(d_rays, d_roff, mode, out=out)
        ctx.synchronize()
        lat_b.append((time.perf_counter() - ts) * 1e3)
    batch_prim = out[0].numpy()
    batch_t = out[1].numpy()
    ctx.set_timing(1)
    b.rebuild()
    ctx.synchronize()
    split = ctx.phase_ms()
    ctx.set_timing(0)
    log("bench: batched step %.3f ms (build %.3f ms, trace %.3f ms)" % (wall_b * 1e3, build_ms,
                                                                      trace_ms))
    res = {"workload": BATCHED_DESC % nseg, "scenes": nseg, "prims": n, "rays": m,
           "steps": steps,
           "step_ms": round(wall_b * 1e3, 4), "build_ms": round(build_ms, 4),
           "trace_ms": round(trace_ms, 4),
           "build_mprims_s": round(n / (build_ms * 1e-3) / 1e6, 3),
           "closest_hit_mrays_s": round(m / (trace_ms * 1e-3) / 1e6, 3),
           "build_phases_ms": ({"bounds+encode": round(split[0], 4), "sort": round(split[1], 4),
                                "tree+fit": round(split[2] + split[3], 4)}
                               if len(split) == 4 else split),
           "sort": {"digit_bits": info.sort_digit_bits, "tile": info.sort_tile,
                    "strategy": SORT_NAMES.get(info.sort_strategy, "?"),
                    "digit_passes": info.sort_passes, "picked_by": "autotuner",
                    "candidates_ms": {"%s-d%d-t%d" % (r["strategy"], r["digit_bits"],
                                                      r["tile"]): r["ms"]
                                      for r in tune_log if r["batch"]}},
           "first_build_ms": round(first_build_s * 1e3, 2)}
    if not plain_path:
        b.destroy()
        ctx.close()
        return res

    def plain(sync_each):
        done, svc, trees = [], [], []
        hip.synchronize()
        t0 = time.perf_counter()
        for s in range(nseg):
            ts = time.perf_counter()
            bv = ctx.build_triangles(d_tri.rows(int(off[s]), int(off[s + 1])), code_bits=30,
                                     leaf_order=PID)
            r0, r1 = int(roff[s]), int(roff[s + 1])
            ctx.trace(bv, d_rays.rows(r0, r1), mode,
                      out=(out[0].rows(r0, r1), out[1].rows(r0, r1), out[2].rows(r0, r1)))
            if sync_each:
                ctx.synchronize()
                now = time.perf_counter()
                svc.append((now - ts) * 1e3)
                done.append((now - t0) * 1e3)
            trees.append(bv)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        for bv in trees:
            bv.destroy()
        return wall, done, svc

    plain(False)  # warm the workspace and the tuner
    wall_async, _, _ = plain(False)
    wall_sync, done, svc = plain(True)
    same = bool(np.array_equal(out[0].numpy(), batch_prim) and
                np.array_equal(out[1].numpy().view(np.uint32), batch_t.view(np.uint32)))
    log("bench: plain %.1f ms queued, %.1f ms one at a time; hits identical: %s"
        % (wall_async * 1e3, wall_sync * 1e3, same))
    b.destroy()
    ctx.close()
    step_ms = wall_b * 1e3
    res.update({
        "plain_queued_ms": round(wall_async * 1e3, 3),
        "plain_one_at_a_time_ms": round(wall_sync * 1e3, 3),
        "throughput_vs_plain_queued": round(wall_async * 1e3 / step_ms, 3),
        "throughput_vs_plain_one_at_a_time": round(wall_sync * 1e3 / step_ms, 3),
        "latency_ms": {"all scenes requested at once": True,
                       "batched": {"p50": _p(lat_b, 50), "p99": _p(lat_b, 99),
                                   "samples": len(lat_b)},
                       "plain": {"p50": _p(done, 50), "p99": _p(done, 99),
                                 "samples": len(done)},
                       "plain_per_scene_service": {"p50": _p(svc, 50), "p99": _p(svc, 99)}},
        "hits_identical": same,
    })
    return res


def bench_batched(args):
    r = batched_compare(args.scenes, args.steps, args.warmup, plain_path=not args.no_extras)
    res = {
        "metric": METRIC,
        "value": r["build_mprims_s"],
        "unit": "Mprims/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["step_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched: " + r["workload"], "n_prims": r["prims"],
                   "n_scenes": r["scenes"], "rays": r["rays"]},
        "batched": r,
    }
    print(json.dumps(res), flush=True)


# ---- served vs plain ------------------------------------------------------------

SERVING_DESC = ("%d requests, Poisson arrivals at %.0f req/s (seed 5); new scenes log-uniform "
                "[1024, 4194304] clustered triangles, 30%% reuse, 10%% refits; rays log-uniform "
                "[4096, 2097152], half primary, half bounce; 63-bit codes, primitive-id leaves; "
                "request data resident in HBM")


def served_compare(count, rate):
    """Open-loop Poisson request stream (seed 5) at a fixed offered load.
    New scenes: triangle counts log-uniform in [1,024, 4,194,304], clustered
    as the 1M-triangle scene (64 Gaussian clusters, sigma 0.02, offsets
    U[-0.002, 0.002]^3); 30% of requests trace a resident scene, 10% refit one
    (every vertex perturbed by U[-1e-4, 1e-4]) before tracing; every request
    carries rays log-uniform in [4,096, 2,097,152], half primary (pinhole from
    (0.5, 0.5, -1)), half diffuse bounces off primary hits.  63-bit codes,
    primitive-id leaves, reference intersector.

    Request data (triangles, rays) is resident in HBM before the stream
    starts, and results stay there.  Served: each request is submitted at its
    arrival time to the request server, which admits, overlaps and coalesces
    them.  Plain: the same requests one at a time through the device entry
    points (build or refit, then trace, then synchronize), queued FIFO on the
    same arrival times.  Latency = completion - arrival; throughput = rays
    over the makespan.  Outputs of both are compared bit for bit."""
    import queue
    import threading

    import numpy as np

    import lbvh_amd
    from lbvh_amd import hip
    from lbvh_amd import workload as W

    stream = W.serving_stream(seed=5, count=count, rate_hz=rate)
    log("bench: %d requests (%d new scenes, %d traces, %d refits), offered load %.0f req/s"
        % (len(stream), sum(r["kind"] == "new" for r in stream),
           sum(r["kind"] == "trace" for r in stream), sum(r["kind"] == "refit" for r in stream),
           rate))
    PID = lbvh_amd.LEAF_PRIMITIVE_ID
    setup = lbvh_amd.Context(0)
    setup.set_autotune(False)
    trees = {}

    def hits(i, r, tri, prim_rays):
        k = r["scene"]
        if r["kind"] == "new":
            trees[k] = setup.build_triangles(tri, code_bits=63, leaf_order=PID)
        elif r["kind"] == "refit":
            setup.refit_triangles(trees[k], tri)
        p, t, _ = setup.trace(trees[k], prim_rays, 0, with_uv=False)
        return p, t

    reqs = []  # (device triangles or None, device rays)
    for i, r, tri, rays in W.serving_requests(stream, hits):
        reqs.append((hip.DeviceArray.from_host(tri) if tri is not None else None,
                     hip.DeviceArray.from_host(rays)))
    for b in trees.values():
        b.destroy()
    setup.close()
    total_rays = sum(r[1].shape[0] for r in reqs)

    def outputs(m):
        return (hip.DeviceArray(m, np.uint32), hip.DeviceArray(m, np.float32),
                hip.DeviceArray((m, 2), np.float32))
    out_s = [outputs(r[1].shape[0]) for r in reqs]
    out_p = [outputs(r[1].shape[0]) for r in reqs]
    log("bench: %d rays in all; streams ready" % total_rays)

    # both paths start with workspaces sized for the largest scene of the
    # stream: the server reserves its lanes, the plain context builds it once
    largest = max(range(len(stream)), key=lambda i: stream[i]["n_tri"]
                  if stream[i]["kind"] == "new" else 0)
    # ---- served ----------------------------------------------------------
    srv = lbvh_amd.Server(0, code_bits=63, leaf_order=PID, isect_mode=0,
                          reserve_prims=stream[largest]["n_tri"])
    # warm the lanes and pools outside the measured stream
    warm = outputs(reqs[0][1].shape[0])
    wt, _ = srv.submit(reqs[0][1], 0, reqs[0][0], out=warm)
    srv.wait(wt)
    tickets = queue.Queue()
    ids = {}

    def submitter():
        t0 = time.perf_counter()
        for i, r in enumerate(stream):
            delay = r["arrival_s"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tri, rays = reqs[i]
            if r["kind"] == "new":
                tk, sid = srv.submit(rays, 0, tri, out=out_s[i])
                ids[r["scene"]] = sid
            elif r["kind"] == "refit":
                tk, _ = srv.submit(rays, ids[r["scene"]], tri, out=out_s[i])
            else:
                tk, _ = srv.submit(rays, ids[r["scene"]], out=out_s[i])
            tickets.put((i, tk))

    th = threading.Thread(target=submitter)
    th.start()
    lat_s, first_sub, last_done = [], None, 0.0
    coalesced = []
    ph_queue, ph_prep, ph_trace = [], [], []
    for _ in stream:
        i, tk = tickets.get()
        _, _, _, info = srv.wait(tk)
        lat_s.append((info["done_us"] - info["submit_us"]) * 1e-3)
        ph_queue.append((info["admit_us"] - info["submit_us"]) * 1e-3)
        ph_prep.append((info["ready_us"] - info["admit_us"]) * 1e-3)
        ph_trace.append((info["done_us"] - info["ready_us"]) * 1e-3)
        coalesced.append(info["coalesced_with"])
        first_sub = info["submit_us"] if first_sub is None else min(first_sub, info["submit_us"])
        last_done = max(last_done, info["done_us"])
    th.join()
    sst = srv.stats()
    srv.close()
    served_span_s = (last_done - first_sub) * 1e-6
    log("bench: served makespan %.1f ms, p50 %.2f ms, p99 %.2f ms, %d launches"
        % (served_span_s * 1e3, percentile(lat_s, 50), percentile(lat_s, 99), sst["launches"]))

    # ---- plain, one request at a time ---------------------------------------
    ctx = lbvh_amd.Context(0)
    ctx.set_autotune(False)
    w0 = ctx.build_triangles(reqs[largest][0], code_bits=63, leaf_order=PID)
    ctx.trace(w0, reqs[0][1], 0, out=warm)
    ctx.synchronize()
    w0.destroy()
    trees = {}
    end = 0.0
    busy = 0.0
    lat_p = []
    ctx.set_timing(2)
    for i, r in enumerate(stream):
        tri, rays = reqs[i]
        arrival = r["arrival_s"]
        start = max(arrival, end)
        s0 = time.perf_counter()
        k = r["scene"]
        if r["kind"] == "new":
            trees[k] = ctx.build_triangles(tri, code_bits=63, leaf_order=PID)
        elif r["kind"] == "refit":
            ctx.refit_triangles(trees[k], tri)
        ctx.trace(trees[k], rays, 0, out=out_p[i])
        ctx.synchronize()
        busy += time.perf_counter() - s0
        end = start + (time.perf_counter() - s0)
        lat_p.append((end - arrival) * 1e3)
    dev_ms = sum(sum(ctx.phase_stats(k)[0]) for k in (1, 2, 3))
    ctx.set_timing(0)
    same = True
    for i in range(len(stream)):
        for a, b in zip(out_s[i], out_p[i]):
            same = same and np.array_equal(a.numpy().view(np.uint32), b.numpy().view(np.uint32))
    plain_span_s = end - stream[0]["arrival_s"]
    for b in trees.values():
        b.destroy()
    ctx.close()
    log("bench: plain makespan %.1f ms, p50 %.2f ms, p99 %.2f ms; identical: %s"
        % (plain_span_s * 1e3, percentile(lat_p, 50), percentile(lat_p, 99), same))
    served_mrays = total_rays / served_span_s / 1e6
    plain_mrays = total_rays / plain_span_s / 1e6
    return {
        "workload": SERVING_DESC % (len(stream), rate),
        "requests": len(stream), "offered_load_rps": rate, "rays": total_rays,
        "served": {"makespan_ms": round(served_span_s * 1e3, 3),
                   "mrays_s": round(served_mrays, 3),
                   "latency_ms_p50": round(percentile(lat_s, 50), 3),
                   "latency_ms_p99": round(percentile(lat_s, 99), 3),
                   "phases_ms_p50": {"admission wait": _p(ph_queue, 50, 3),
                                     "build / refit": _p(ph_prep, 50, 3),
                                     "traversal": _p(ph_trace, 50, 3)},
                   "traversal_launches": sst["launches"],
                   "mean_coalesced": round(float(np.mean(coalesced)), 2),
                   "max_coalesced": sst["max_coalesced"],
                   "memory_peak_gb": round(sst["memory_peak_bytes"] / 2**30, 3),
                   "evictions": sst["evictions"]},
        "plain": {"makespan_ms": round(plain_span_s * 1e3, 3), "mrays_s": round(plain_mrays, 3),
                  "busy_ms": round(busy * 1e3, 3),
                  "device_ms": round(dev_ms, 3),
                  "capacity_rps": round(len(stream) / busy, 1),
                  "latency_ms_p50": round(percentile(lat_p, 50), 3),
                  "latency_ms_p99": round(percentile(lat_p, 99), 3)},
        "throughput_vs_plain": round(served_mrays / plain_mrays, 3),
        "results_identical": bool(same),
    }


def bench_serving(args):
    r = served_compare(args.requests, args.rate)
    res = {
        "metric": METRIC,
        "value": r["served"]["mrays_s"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": r["requests"],
        "warmup": 1,
        "ms_per_step": round(r["served"]["makespan_ms"] / r["requests"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "serving: " + r["workload"], "requests": r["requests"],
                   "offered_load_rps": r["offered_load_rps"], "rays": r["rays"]},
        "serving": r,
    }
    print(json.dumps(res), flush=True)


def bench_large(args):
    """16,777,216 triangles (seed 3): centres U[0,1]^3, offsets
    U[-5e-4, 5e-4]^3.  Per-phase device times of builds with 30-bit and
    63-bit codes (reference leaf order, as builder<float>), then a refit
    after perturbing every vertex by U[-1e-4, 1e-4] with the topology kept.
    Inputs resident in HBM; steps = builds per code width."""
    import numpy as np

    import lbvh_amd
    from lbvh_amd import hip
    from lbvh_amd import workload as W

    n = 1 << 24
    log("bench: generating the 16,777,216-triangle uniform scene (seed 3)")
    v, tri = W.scene_uniform(3, n, 5e-4)
    tri2 = W.make_triangles(W.perturb(33, v, 1e-4))
    del v
    ctx = lbvh_amd.Context(0)
    d_tri = hip.DeviceArray.from_host(tri)
    d_tri2 = hip.DeviceArray.from_host(tri2)
    out = {}
    for bits in (30, 63):
        b = ctx.build_triangles(d_tri, code_bits=bits)  # tunes the sort for this size
        ctx.synchronize()
        b.destroy()
        for _ in range(args.warmup):
            ctx.build_triangles(d_tri, code_bits=bits).destroy()
        ctx.synchronize()
        ctx.set_timing(2)
        trees = [ctx.build_triangles(d_tri, code_bits=bits) for _ in range(args.steps)]
        ctx.synchronize()
        tot, calls = ctx.phase_stats(1)
        ctx.set_timing(1)
        split = [ctx.build_triangles(d_tri, code_bits=bits) for _ in range(3)]
        ctx.synchronize()
        ph, _ = ctx.phase_stats(1)
        ctx.set_timing(0)
        build_ms = sum(tot) / calls
        scale = build_ms / sum(ph)
        inf = trees[0].info()
        out[bits] = {"build_ms": round(build_ms, 4),
                     "mprims_s": round(n / (build_ms * 1e-3) / 1e6, 2),
                     "phases_ms": {"encode (bounds+codes)": round(ph[0] * scale, 4),
                                   "sort": round(ph[1] * scale, 4),
                                   "tree+fit": round((ph[2] + ph[3]) * scale, 4)},
                     "sort": {"digit_bits": inf.sort_digit_bits, "tile": inf.sort_tile,
                              "strategy": SORT_NAMES.get(inf.sort_strategy, "?"),
                              "global_digits": inf.sort_global_digits}}
        if bits == 30:
            r = trees[0]
            for _ in range(args.warmup):
                ctx.refit_triangles(r, d_tri2)
                ctx.refit_triangles(r, d_tri)
            ctx.synchronize()
            ctx.set_timing(2)
            for k in range(args.steps):
                ctx.refit_triangles(r, d_tri2 if k % 2 == 0 else d_tri)
            ctx.synchronize()
            rt, rc = ctx.phase_stats(3)
            ctx.set_timing(0)
            out["refit_ms"] = round(sum(rt) / rc, 4)
        for b in trees + split:
            b.destroy()
        log("bench: %d-bit build %.3f ms" % (bits, build_ms))
    ctx.close()
    res = {
        "metric": METRIC,
        "value": out[30]["mprims_s"],
        "unit": "Mprims/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out[30]["build_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "large build: 16,777,216 uniform FP32 triangles (seed 3, offsets "
                               "U[-5e-4, 5e-4]^3), reference leaf order; 30- and 63-bit codes; "
                               "refit after a U[-1e-4, 1e-4] vertex perturbation",
                   "n_prims": n},
        "build_30bit": out[30],
        "build_63bit": out[63],
        "refit_ms": out["refit_ms"],
        "refit_mprims_s": round(n / (out["refit_ms"] * 1e-3) / 1e6, 2),
    }
    print(json.dumps(res), flush=True)


# Algorithmic bytes per primitive of the FP64 build (builder<double>, 60-bit
# keys, reference leaves): bounds 48, encode 48 + 8, sort 8 + 8 digit passes
# x 16 (key only), tree 8 + 16 + 4 + 8 (codes, left/right, range end,
# parent links), fit 48 + 48 + 8 (leaf box, node box, links).
BUILD_BYTES_PER_PRIM_F64 = 48 + 56 + (8 + 8 * 16) + 36 + 104


def _f64_trace_timing(ctx, d_boxes, tri96, args):
    """traverser<double> on the 1M FP64 scene: the 1920x1080 primary rays
    widened to FP64, reference intersector, device-resident inputs; steps =
    timed traversals.  The tree is built with primitive-id leaves: with
    reference leaves on unsorted input the fit and the traverser read
    primitives[sorted position] (lbvh.h:1855, 1904-1905), so every box spans
    random triangles of the whole scene and a ray visits most of the tree."""
    import numpy as np

    import lbvh_amd
    from lbvh_amd import hip
    from lbvh_amd import workload as W
    rays = W.primary_rays(WIDTH, HEIGHT).astype(np.float64)
    m = rays.shape[0]
    n = tri96.shape[0]
    nodes = hip.DeviceArray((n - 1,), lbvh_amd.NODE64_DTYPE)
    ctx.build_aabbs_f64(d_boxes, lbvh_amd.LEAF_PRIMITIVE_ID, out=nodes)
    d_tri = hip.DeviceArray.from_host(tri96)
    d_rays = hip.DeviceArray.from_host(rays)
    outs = (hip.DeviceArray(m, np.uint64), hip.DeviceArray(m, np.float64),
            hip.DeviceArray((m, 2), np.float64))
    for _ in range(max(args.warmup, 1)):
        ctx.trace_f64(nodes, d_tri, d_rays, 0, out=outs)
        ctx.synchronize()
        log("bench: FP64 traversal warmup step done")
    ctx.set_timing(2)
    for _ in range(args.steps):
        ctx.trace_f64(nodes, d_tri, d_rays, 0, out=outs)
    ctx.synchronize()
    tot, calls = ctx.phase_stats(2)
    ctx.set_timing(0)
    hits = int((outs[0].numpy() != np.uint64(2**64 - 1)).sum())
    for a in outs + (d_tri, d_rays, nodes):
        a.close()
    ms = sum(tot) / calls
    log("bench: FP64 traversal of %d rays %.3f ms" % (m, ms))
    return {"rays": m, "hits": hits, "trace_ms": round(ms, 4),
            "mrays_s": round(m / (ms * 1e-3) / 1e6, 2)}


def bench_f64(args):
    """builder<double>: the 1,048,576-triangle clustered scene of the default
    workload as triangle<double> (vertices widened to FP64), converter boxes
    built on the device with u64 codes into node<double>; then 16,777,216
    uniform FP64 boxes.  Reference leaf order (builder<double>'s bytes).
    Inputs resident in HBM; steps = timed builds per size."""
    import numpy as np

    import lbvh_amd
    from lbvh_amd import hip
    from lbvh_amd import workload as W

    log("bench: FP64 scenes (1,048,576 clustered triangles, seed 2; 16,777,216 uniform boxes)")
    v, _ = W.scene_clustered(2, N_PRIMS)
    tri96 = W.triangles_f64(v)
    del v
    rng = np.random.default_rng(1)
    lo = rng.random((1 << 24, 3))
    big = np.concatenate([lo, lo + rng.random((1 << 24, 3)) * 5e-4], axis=1)
    del lo
    ctx = lbvh_amd.Context(0)
    out = {}
    for name, boxes in (("1m", W.triangle_boxes_f64(tri96)), ("16m", big)):
        n = boxes.shape[0]
        d = hip.DeviceArray.from_host(boxes)
        nodes = hip.DeviceArray((n - 1,), lbvh_amd.NODE64_DTYPE)
        ctx.build_aabbs_f64(d, 0, out=nodes)  # tunes the sort for this size
        for _ in range(args.warmup):
            ctx.build_aabbs_f64(d, 0, out=nodes)
        ctx.synchronize()
        ctx.set_timing(2)
        for _ in range(args.steps):
            ctx.build_aabbs_f64(d, 0, out=nodes)
        ctx.synchronize()
        tot, calls = ctx.phase_stats(1)
        ctx.set_timing(1)
        for _ in range(3):
            ctx.build_aabbs_f64(d, 0, out=nodes)
        ctx.synchronize()
        ph, _ = ctx.phase_stats(1)
        ctx.set_timing(0)
        build_ms = sum(tot) / calls
        scale = build_ms / sum(ph)
        out[name] = {"n_prims": n, "build_ms": round(build_ms, 4),
                     "mprims_s": round(n / (build_ms * 1e-3) / 1e6, 2),
                     "phases_ms": {"encode (bounds+codes)": round(ph[0] * scale, 4),
                                   "sort": round(ph[1] * scale, 4),
                                   "tree+fit": round((ph[2] + ph[3]) * scale, 4)}}
        log("bench: FP64 build of %d boxes %.3f ms" % (n, build_ms))
        if name == "1m":
            out["trace_1m"] = _f64_trace_timing(ctx, d, tri96, args)
        d.close()
        nodes.close()
    ctx.close()
    del big
    one = out["1m"]
    achieved = BUILD_BYTES_PER_PRIM_F64 * N_PRIMS / (one["build_ms"] * 1e-3) / 1e9
    res = {
        "metric": METRIC,
        "value": one["mprims_s"],
        "unit": "Mprims/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": one["build_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "builder<double>: 1,048,576 clustered triangles (seed 2) as "
                               "triangle<double>, converter boxes in FP64, reference leaf order; "
                               "also 16,777,216 uniform FP64 boxes (seed 1)",
                   "n_prims": N_PRIMS},
        "roofline": {"bound": "hbm", "scope": "whole build (all kernels of one build)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_prim": BUILD_BYTES_PER_PRIM_F64},
        "build_1m": one,
        "build_16m": out["16m"],
        "trace_1m": out["trace_1m"],
    }
    if not args.no_cpu_baseline:
        import oracle
        fast = oracle.ref_available(fast=True)
        if fast or oracle.ref_available():
            threads = oracle.ref(fast).ref_hardware_concurrency() or 1
            times = []
            t_end = time.time() + args.cpu_seconds * 0.6
            while time.time() < t_end or len(times) < 2:
                t0 = time.perf_counter()
                oracle.ref_build_f64(tri96, "triangles", 0, threads=threads, fast=fast)
                times.append(time.perf_counter() - t0)
                if len(times) >= 6:
                    break
            build_s = sorted(times)[len(times) // 2]
            res["cpu_baseline"] = {
                "value": round(N_PRIMS / build_s / 1e6, 3), "unit": "Mprims/s",
                "cores": int(threads), "kind": "reference",
                "flags": "-O3 -ffast-math" if fast else "-O2 -ffp-contract=off",
                "sample": "%d builds of the 1,048,576-triangle scene with lbvh::builder<double> "
                          "(median %.1f ms)" % (len(times), build_s * 1e3)}
    print(json.dumps(res), flush=True)



# ---- config 0: the reference's own CPU-runnable case ----------------------------

def bench_config0(args):
    """65,536 uniform triangles (seed 1, offsets U[-0.01, 0.01]^3), 30-bit
    codes, the reference leaf order (lbvh::builder<float>'s exact tree: a leaf
    index is the primitive's Morton-order position); 256x256 pinhole primary
    rays from (0.5, 0.5, -1), closest hit with tmin 1e-4, tmax 1e30.  The GPU
    build and trace (inputs resident in HBM) beside the real lbvh.h on this
    host's cores: its builder and its traverser over the same tree and rays.
    On this tree a ray visits most leaves (the leaf index is the sorted
    position, not the primitive), so the CPU traverser is near brute force."""
    import numpy as np

    import lbvh_amd
    from lbvh_amd import hip
    from lbvh_amd import workload as W

    n = 65536
    _, tri = W.scene_uniform(1, n, 0.01)
    rays = W.primary_rays(256, 256)
    m = rays.shape[0]
    ctx = lbvh_amd.Context(0)
    d_tri = hip.DeviceArray.from_host(tri)
    d_rays = hip.DeviceArray.from_host(rays)
    out = (hip.DeviceArray(m, np.uint32), hip.DeviceArray(m, np.float32),
           hip.DeviceArray((m, 2), np.float32))
    REF = lbvh_amd.LEAF_REFERENCE
    for _ in range(args.warmup):
        b = ctx.build_triangles(d_tri, code_bits=30, leaf_order=REF)
        ctx.trace(b, d_rays, 0, out=out)
        ctx.synchronize()
        b.destroy()
    trees = []
    ctx.set_timing(2)
    hip.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b = ctx.build_triangles(d_tri, code_bits=30, leaf_order=REF)
        ctx.trace(b, d_rays, 0, out=out)
        trees.append(b)
    hip.synchronize()
    wall = time.perf_counter() - t0
    ctx.synchronize()
    bt, nb = ctx.phase_stats(1)
    tt, nt = ctx.phase_stats(2)
    ctx.set_timing(0)
    nodes = trees[-1].nodes()
    for b in trees:
        b.destroy()
    gp, gt = out[0].numpy(), out[1].numpy()
    ctx.close()
    build_ms = sum(bt) / args.steps
    trace_ms = sum(tt) / args.steps
    res = {
        "metric": METRIC,
        "value": round(n / (build_ms * 1e-3) / 1e6, 3),
        "unit": "Mprims/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "config 0: 65,536 uniform FP32 triangles (seed 1), 30-bit codes, "
                               "reference leaf order; 256x256 primary rays", "n_prims": n,
                   "code_bits": 30, "rays": m},
        "build_ms": round(build_ms, 4),
        "build_roofline_frac": round(BUILD_BYTES_PER_PRIM_30 * n / (build_ms * 1e-3) / 1e9
                                     / HBM_PEAK_GBS, 5),
        "closest_hit_mrays_s": round(m / (trace_ms * 1e-3) / 1e6, 3),
        "trace_ms": round(trace_ms, 4),
        "hits": int((gp != MISS).sum()),
    }
    if not args.no_cpu_baseline:
        import oracle
        fast = oracle.ref_available(fast=True)
        if fast or oracle.ref_available():
            threads = _ref_threads(fast)
            times = []
            for _ in range(6):
                t0 = time.perf_counter()
                rn, _, _ = oracle.ref_build(tri, 30, 0, threads=threads, fast=fast)
                times.append(time.perf_counter() - t0)
            build_s = sorted(times)[len(times) // 2]
            # the traverser on a first sample sizes the bounded run
            probe = rays[::64]
            t0 = time.perf_counter()
            oracle.ref_trace(rn, tri, probe, 0, threads=threads, fast=fast)
            per_ray = (time.perf_counter() - t0) / probe.shape[0]
            step = 1
            while per_ray * m / step > args.cpu_seconds * 0.8 and step < 64:
                step *= 2
            sample = rays[::step]
            t0 = time.perf_counter()
            oracle.ref_trace(rn, tri, sample, 0, threads=threads, fast=fast)
            trace_s = time.perf_counter() - t0
            # hits are compared with the reference built with its strict flags
            # (-O2 -ffp-contract=off): -ffast-math changes the traversal's
            # arithmetic, not the tree
            rp, rt, _ = oracle.ref_trace(rn, tri, sample, 0, threads=threads, fast=False)
            res["cpu_baseline"] = {
                "value": round(n / build_s / 1e6, 3), "unit": "Mprims/s", "cores": int(threads),
                "kind": "reference",
                "flags": "-O3 -ffast-math" if fast else "-O2 -ffp-contract=off",
                "sample": "6 builds of the 65,536-triangle scene (30-bit codes, reference leaf "
                          "order, median %.2f ms); closest hit of %s on that tree in %.2f s"
                          % (build_s * 1e3, "all 65,536 rays" if step == 1 else
                             "%d rays (every %dth of the 256x256)" % (sample.shape[0], step),
                             trace_s),
                "build_ms": round(build_s * 1e3, 3),
                "closest_hit_mrays_s": round(sample.shape[0] / trace_s / 1e6, 5),
                "same_nodes_as_gpu": bool(rn.tobytes() == nodes.tobytes()),
                "same_hits_as_gpu_strict_reference": bool(
                    np.array_equal(rp, gp[::step]) and
                    np.array_equal(rt.view(np.uint32), gt[::step].view(np.uint32))),
            }
            res["cpu_baseline"].update(host_info())
    print(json.dumps(res), flush=True)


# ---- counter calibration -------------------------------------------------------

def bench_hbm_copy(args):
    """Five device-to-device copies of 512 MiB (the runtime's copy kernel
    reads 512 MiB and writes 512 MiB per dispatch): the known-size kernel
    that scales rocprofv3's FETCH_SIZE and WRITE_SIZE counters to bytes."""
    import numpy as np

    from lbvh_amd import hip
    nbytes = 512 << 20
    a = hip.DeviceArray(nbytes // 4, np.float32)
    b = hip.DeviceArray(nbytes // 4, np.float32)
    hip.check(hip.hip().hipMemset(a.ptr, 1, nbytes), "hipMemset")
    for i in range(5):
        t0 = time.perf_counter()
        b.copy_from(a)
        hip.synchronize()
        log("bench: copy %d: %d bytes read, %d written in %.3f ms"
            % (i, nbytes, nbytes, (time.perf_counter() - t0) * 1e3))
    a.close()
    b.close()


# ---- the headline --------------------------------------------------------------

def bench_single(args):
    import numpy as np

    import lbvh_amd
    from lbvh_amd import hip
    from lbvh_amd import workload as W

    log("bench: generating the 1,048,576-triangle clustered scene (seed 2) and %dx%d rays"
        % (WIDTH, HEIGHT))
    _, tri = W.scene_clustered(2, N_PRIMS)
    rays = W.primary_rays(WIDTH, HEIGHT)

    ctx = lbvh_amd.Context(0)
    d_tri = hip.DeviceArray.from_host(tri)
    d_rays = hip.DeviceArray.from_host(rays)

    # bounce rays from the GPU's own primary hits (identical to the oracle's)
    # (the first build of this size also runs the sort autotuner)
    bvh = ctx.build_triangles(d_tri, code_bits=63, leaf_order=lbvh_amd.LEAF_PRIMITIVE_ID)
    p, t, _ = ctx.trace(bvh, d_rays, lbvh_amd.ISECT_REFERENCE)
    ctx.synchronize()
    sort_info = bvh.info()
    hp = p.numpy()
    ht = t.numpy()
    bounce, _ = W.bounce_rays(5, tri, rays, hp, ht)
    d_bounce = hip.DeviceArray.from_host(bounce)
    n_primary, n_bounce = rays.shape[0], bounce.shape[0]
    log("bench: %d primary rays, %d hits -> %d bounce rays" % (n_primary, int((hp != MISS).sum()),
                                                             n_bounce))
    # hit-distance error against FP64, both ray sets, reference and
    # watertight intersectors (outside the timed region)
    t_err = {}
    for name, mode in (("reference", lbvh_amd.ISECT_REFERENCE),
                       ("watertight", lbvh_amd.ISECT_WATERTIGHT)):
        e = {}
        for rname, rs, drs in (("primary", rays, d_rays), ("bounce", bounce, d_bounce)):
            pp, tt, _ = ctx.trace(bvh, drs, mode, with_uv=False)
            ctx.synchronize()
            e[rname] = t_error_vs_fp64(tri, rs, pp.numpy(), tt.numpy())
        t_err[name] = e
    # the first traversal of each size bucket tunes the traversal
    ctx.trace(bvh, d_bounce, lbvh_amd.ISECT_REFERENCE, with_uv=False)
    ctx.synchronize()
    tune_log = ctx.tune_log()
    trace_log = [r for r in tune_log if r["kind"] == "traversal" and r["log2_n"] in
                 (int(math.ceil(math.log2(n_primary))), int(math.ceil(math.log2(n_bounce))))
                 and r["mode"] == lbvh_amd.ISECT_REFERENCE]
    bvh.destroy()
    out_p = (hip.DeviceArray(n_primary, np.uint32), hip.DeviceArray(n_primary, np.float32),
             hip.DeviceArray((n_primary, 2), np.float32))
    out_b = (hip.DeviceArray(n_bounce, np.uint32), hip.DeviceArray(n_bounce, np.float32),
             hip.DeviceArray((n_bounce, 2), np.float32))

    def step():
        b = ctx.build_triangles(d_tri, code_bits=63, leaf_order=lbvh_amd.LEAF_PRIMITIVE_ID)
        ctx.trace(b, d_rays, lbvh_amd.ISECT_REFERENCE, out=out_p)
        ctx.trace(b, d_bounce, lbvh_amd.ISECT_REFERENCE, out=out_b)
        return b

    prev = None
    for i in range(args.warmup):
        b = step()
        if prev is not None:
            prev.destroy()
        prev = b
    ctx.synchronize()
    log("bench: warmup done (%d steps)" % args.warmup)

    # timed region: K steps back to back; the engine records one HIP event
    # pair per call on the stream its kernels run on (the context's stream)
    trees = []
    ctx.set_timing(2)
    hip.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        trees.append(step())
    hip.synchronize()
    wall = time.perf_counter() - t0
    ctx.synchronize()
    build_tot, nb = ctx.phase_stats(1)
    trace_tot, nt = ctx.phase_stats(2)
    ctx.set_timing(0)
    for b in trees:
        b.destroy()
    if prev is not None:
        prev.destroy()
    log("bench: timed %d steps in %.3f s" % (args.steps, wall))
    # phase split of the build, measured separately (events between phases
    # add short idle gaps, so they stay out of the timed steps)
    split = []
    ctx.set_timing(1)
    for i in range(5):
        split.append(ctx.build_triangles(d_tri, code_bits=63,
                                         leaf_order=lbvh_amd.LEAF_PRIMITIVE_ID))
    ctx.synchronize()
    split_ph, ns = ctx.phase_stats(1)
    ctx.set_timing(0)
    for b in split:
        b.destroy()
    ctx.close()
    d_tri.close()
    d_rays.close()
    d_bounce.close()
    for a in out_p + out_b:
        a.close()

    K = args.steps
    build_ms = sum(build_tot) / K
    split_sum = sum(split_ph)
    phases = [build_ms * x / split_sum for x in split_ph]  # split scaled to the timed total
    trace_ms = sum(trace_tot) / K  # primary + bounce per step
    ms_per_step = wall * 1e3 / K
    mprims = N_PRIMS / (build_ms * 1e-3) / 1e6
    mrays = (n_primary + n_bounce) / (trace_ms * 1e-3) / 1e6

    digit = sort_info.sort_digit_bits
    kernels = {
        "trace (primary+bounce)": trace_ms,
        "sort": phases[1],
        "bounds+encode": phases[0],
        "tree+fit": phases[2] + phases[3],
    }
    dominant = max(kernels, key=kernels.get)
    build_bytes = N_PRIMS * BUILD_BYTES_PER_PRIM_63
    build_frac = build_bytes / (build_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    if dominant == "sort":
        sb = N_PRIMS * SORT_BYTES_PER_PRIM_63
        ach = sb / (phases[1] * 1e-3) / 1e9
        roof = {"kernel": "radix sort (canonical 8 digit passes, u64 key + u32 id)",
                "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": sb, "avg_launch_ms": round(phases[1], 5)}
    else:
        # traversal: the bytes it touches (an upper-bound proxy for its HBM
        # traffic); the compulsory ray and hit bytes beside them
        ach = TRACE_TOUCHED_BYTES_PER_STEP / (trace_ms * 1e-3) / 1e9
        comp = (n_primary + n_bounce) * RAY_COMPULSORY_BYTES / (trace_ms * 1e-3) / 1e9
        roof = {"kernel": "k_trace (closest hit, reference intersector)", "bound": "hbm",
                "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": TRACE_HBM_TRAFFIC_PER_LAUNCH,
                "bytes": "touched: popped nodes x 32 + tested children x 32 + tested triangles x 48",
                "bytes_per_launch": TRACE_TOUCHED_BYTES_PER_STEP / 2,
                "avg_launch_ms": round(trace_ms / 2, 5),
                "compulsory_gbs": round(comp, 2)}

    res = {
        "metric": METRIC,
        "value": round(mprims, 3),
        "unit": "Mprims/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "clustered scene, 1,048,576 FP32 triangles (64 Gaussian clusters, "
                               "seed 2), 63-bit Morton codes, primitive-id leaves; 1920x1080 primary "
                               "rays + one diffuse bounce per hit",
                   "n_prims": N_PRIMS, "code_bits": 63, "primary_rays": n_primary,
                   "bounce_rays": n_bounce},
        "build_ms": round(build_ms, 4),
        "build_phases_ms": {"bounds+encode": round(phases[0], 4), "sort": round(phases[1], 4),
                            "tree+fit": round(phases[2] + phases[3], 4)},
        "sort": {"digit_bits": digit, "tile": sort_info.sort_tile,
                 "strategy": SORT_NAMES.get(sort_info.sort_strategy, "?"),
                 "global_digits": sort_info.sort_global_digits, "picked_by": "autotuner",
                 "candidates_ms": {"%s-d%d-t%d" % (r["strategy"], r["digit_bits"], r["tile"]): r["ms"]
                                   for r in tune_log if not r["batch"] and r["kind"] == "sort"}},
        "traversal_tuning": {
            "2^%d rays" % lg: {"picked": next(("%s, %d waves/SIMD"
                                               % ("key order" if r["reorder"] else "as given",
                                                  r["waves_per_simd"])
                                               for r in trace_log if r["log2_n"] == lg
                                               and r["picked"]), None),
                               "candidates_ms": {"%s-w%d" % ("sorted" if r["reorder"] else "given",
                                                             r["waves_per_simd"]): r["ms"]
                                                 for r in trace_log if r["log2_n"] == lg}}
            for lg in sorted(set(r["log2_n"] for r in trace_log))},
        "build_roofline_frac": round(build_frac, 4),
        "build_algorithmic_bytes_per_prim": BUILD_BYTES_PER_PRIM_63,
        "closest_hit_mrays_s": round(mrays, 3),
        "trace_ms": round(trace_ms, 4),
        "t_error_vs_fp64": t_err,
        "dominant_kernel": dominant,
        "roofline": roof,
    }
    if not args.no_extras:
        log("bench: batched vs plain (4,096 scenes)")
        res["batched"] = batched_compare(4096, 10, 3)
        log("bench: served vs plain (%d requests at %.0f req/s)"
            % (HEADLINE_SERVING_REQUESTS, HEADLINE_SERVING_RATE))
        res["serving"] = served_compare(HEADLINE_SERVING_REQUESTS, HEADLINE_SERVING_RATE)
    if not args.no_cpu_baseline:
        log("bench: timing the reference CPU path (bounded sample)")
        res["cpu_baseline"] = cpu_baseline(tri, rays, bounce, args.cpu_seconds)
    print(json.dumps(res), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported "
              "(scenes shard across GPUs only by assigning whole scenes)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    return {"single": bench_single, "config0": bench_config0, "batched": bench_batched,
            "serving": bench_serving, "large": bench_large, "f64": bench_f64,
            "hbm-copy": bench_hbm_copy}[args.workload](args)


if __name__ == "__main__":
    main()
