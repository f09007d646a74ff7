"""LBVH build + closest-hit benchmark on one MI355X.

Workload: the skewed single-scene case — 1,048,576 FP32 triangles whose
centres come from 64 Gaussian clusters (seed 2, sigma 0.02, clamped to [0,1],
edge offsets U[-0.002, 0.002]^3), 63-bit Morton codes, primitive-id leaves;
1920x1080 pinhole primary rays from (0.5, 0.5, -1) looking +z (60 degree
vertical FOV, tmin 1e-4, tmax 1e30) plus one cosine-weighted diffuse bounce
ray per primary hit.  Inputs are generated once on the host and are resident
in HBM before the timed region.

One step = build the BVH (bounds, encode, sort, tree, fit) + closest-hit of
the primary rays + closest-hit of the bounce rays, all on one stream.

Prints one JSON line.  `value` is the build throughput in Mprims/s; the line
also carries closest-hit Mrays/s, the per-phase split, the roofline of the
dominant kernel, and the reference CPU path timed on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_PRIMS = 1 << 20
WIDTH, HEIGHT = 1920, 1080
SORT_NAMES = {1: "onesweep", 2: "reduce-then-scan", 3: "hybrid"}
# Algorithmic bytes per primitive of the build with 63-bit codes (u64 key +
# u32 id), the canonical accounting whatever sort the tuner picks: bounds 36,
# encode 48, sort 8 + 8 digit passes x 24, tree 16, fit 92.
SORT_BYTES_PER_PRIM_63 = 8 + 8 * 24
BUILD_BYTES_PER_PRIM_63 = 36 + 48 + SORT_BYTES_PER_PRIM_63 + 16 + 92
# Bytes the closest-hit traversal touches on the benchmark workload (primary
# + bounce rays on the 1M clustered tree): popped nodes x 32 + box-tested
# children x 32 + tested triangles x 48, counted by the instrumented CPU
# traversal of the same tree (tests/test_workloads.py recomputes it).
TRACE_TOUCHED_BYTES_PER_STEP = 3857317376
RAY_COMPULSORY_BYTES = 48  # 32-byte ray in, 16-byte hit out
# HBM bytes per k_trace dispatch on this workload, averaged over the primary
# and bounce dispatches: rocprofv3 --pmc FETCH_SIZE (x 2.0 KiB) + WRITE_SIZE
# (x 1.0 KiB), scaled by a 512 MiB copy kernel
# (profiles/clustered1m_63bit_hbm_traffic.md: primary 633.4 + 35.5 MB, bounce
# 910.7 + 4.2 MB; mean 772.1 MB read + 19.9 MB written).
TRACE_HBM_TRAFFIC_PER_LAUNCH = 791.9e6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget of the CPU-baseline sample")
    ap.add_argument("--workload", choices=["single", "batched", "serving", "large", "f64"],
                    default="single",
                    help="single: the 1M-triangle scene (default); batched: 4,096 uneven "
                         "scenes built in one segmented pass against the plain per-scene path; "
                         "serving: an open-loop request stream through the request server "
                         "against the plain path one request at a time; large: 16,777,216 "
                         "triangles, per-phase 30/63-bit builds and a refit; f64: "
                         "builder<double> over the 1M scene in FP64, and 16,777,216 boxes")
    ap.add_argument("--scenes", type=int, default=4096, help="batched: number of scenes")
    ap.add_argument("--requests", type=int, default=160, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=1000.0,
                    help="serving: offered load, requests per second (Poisson)")
    return ap.parse_args()


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
        threads = oracle.ref(fast).ref_hardware_concurrency() or threads
    t_end = time.time() + budget_s * 0.6
    times = []
    nodes = None
    while time.time() < t_end or len(times) < 2:
        t0 = time.perf_counter()
        if kind == "reference":
            nodes, _, _ = oracle.ref_build(tri, 63, 0, threads=threads, fast=fast)
        else:
            nodes, _, _, _ = oracle.build(tri, 63, 0)
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
    return {
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


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def bench_batched(args):
    """4,096 independent scenes (seed 4) with triangle counts log-uniform in
    [64, 262144], each a uniform scene in its own unit cube (offsets
    U[-0.01, 0.01]^3), 30-bit codes, primitive-id leaves; 1,024 rays per scene
    from the scene's bounding-box surface in directions uniform on the sphere.

    Batched step: one segmented rebuild of every scene (replayed from a HIP
    graph) + one closest-hit launch over every scene's rays.  Plain path: the
    same scenes built and traced one call at a time through the single-scene
    entry points (a fresh tree per scene, as the plain API returns), once with
    a host synchronization per scene (per-request latency) and once queued
    back to back.  Both produce identical hits (checked)."""
    import numpy as np

    import lbvh_amd
    from lbvh_amd import hip
    from lbvh_amd import workload as W

    nseg = args.scenes
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

    # ---- batched ----------------------------------------------------------
    t0 = time.perf_counter()
    b = ctx.build_batch(d_tri, off, leaf_order=PID)
    ctx.synchronize()
    first_build_s = time.perf_counter() - t0
    info = b.info()
    tune_log = ctx.tune_log()
    log("bench: batch built (first call incl. allocation and tuning: %.1f ms); sort %d-bit "
        "digits, tile %d, %d passes" % (first_build_s * 1e3, info.sort_digit_bits, info.sort_tile,
                                        info.sort_passes))
    for _ in range(args.warmup):
        b.rebuild(graph=True)
        b.trace(d_rays, d_roff, mode, out=out)
    ctx.synchronize()
    ctx.set_timing(2)
    hip.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.rebuild(graph=True)
        b.trace(d_rays, d_roff, mode, out=out)
    hip.synchronize()
    wall_b = (time.perf_counter() - t0) / args.steps
    ctx.synchronize()
    bt, _ = ctx.phase_stats(1)
    tt, _ = ctx.phase_stats(2)
    ctx.set_timing(0)
    build_ms = sum(bt) / args.steps
    trace_ms = sum(tt) / args.steps
    batch_prim = out[0].numpy()
    batch_t = out[1].numpy()
    # phase split of one eager rebuild
    ctx.set_timing(1)
    b.rebuild()
    ctx.synchronize()
    split = ctx.phase_ms()
    ctx.set_timing(0)
    log("bench: batched step %.3f ms (build %.3f ms, trace %.3f ms)" % (wall_b * 1e3, build_ms,
                                                                      trace_ms))

    # ---- plain, one request at a time ---------------------------------------
    def plain(sync_each):
        lat, trees = [], []
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
                lat.append((time.perf_counter() - ts) * 1e3)
            trees.append(bv)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        for bv in trees:
            bv.destroy()
        return wall, lat

    plain(False)  # warm the workspace
    wall_async, _ = plain(False)
    wall_sync, lat = plain(True)
    same = bool(np.array_equal(out[0].numpy(), batch_prim) and
                np.array_equal(out[1].numpy().view(np.uint32), batch_t.view(np.uint32)))
    log("bench: plain %.1f ms queued, %.1f ms one at a time; hits identical: %s"
        % (wall_async * 1e3, wall_sync * 1e3, same))
    b.destroy()
    ctx.close()

    step_ms = build_ms + trace_ms
    res = {
        "metric": "build Mprims/s + closest-hit Mrays/s @1M tris; p99 latency batched vs plain",
        "value": round(n / (build_ms * 1e-3) / 1e6, 3),
        "unit": "Mprims/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_b * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched: %d scenes (seed 4), triangle counts log-uniform in "
                               "[64, 262144], uniform scenes in their own unit cubes, 30-bit "
                               "codes, primitive-id leaves; 1,024 box-surface rays per scene"
                               % nseg,
                   "n_prims": n, "n_scenes": nseg, "rays": m},
        "batched": {"build_ms": round(build_ms, 4), "trace_ms": round(trace_ms, 4),
                    "build_phases_ms": {"bounds+encode": round(split[0], 4),
                                        "sort": round(split[1], 4),
                                        "tree+fit": round(split[2] + split[3], 4)}
                    if len(split) == 4 else split,
                    "closest_hit_mrays_s": round(m / (trace_ms * 1e-3) / 1e6, 3),
                    "sort": {"digit_bits": info.sort_digit_bits, "tile": info.sort_tile,
                             "strategy": SORT_NAMES.get(info.sort_strategy, "?"),
                             "digit_passes": info.sort_passes, "picked_by": "autotuner",
                             "candidates_ms": {"%s-d%d-t%d" % (r["strategy"], r["digit_bits"],
                                                               r["tile"]): r["ms"]
                                               for r in tune_log if r["batch"]}},
                    "first_build_ms": round(first_build_s * 1e3, 2),
                    "latency_ms_p50": round(wall_b * 1e3, 4),
                    "latency_ms_p99": round(wall_b * 1e3, 4)},
        "plain": {"queued_ms": round(wall_async * 1e3, 3),
                  "one_at_a_time_ms": round(wall_sync * 1e3, 3),
                  "latency_ms_p50": round(percentile(lat, 50), 4),
                  "latency_ms_p99": round(percentile(lat, 99), 4)},
        "speedup_vs_plain_queued": round(wall_async * 1e3 / step_ms, 3),
        "speedup_vs_plain_one_at_a_time": round(wall_sync * 1e3 / step_ms, 3),
        "hits_identical": same,
    }
    print(json.dumps(res), flush=True)


def bench_serving(args):
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

    stream = W.serving_stream(seed=5, count=args.requests, rate_hz=args.rate)
    log("bench: %d requests (%d new scenes, %d traces, %d refits), offered load %.0f req/s"
        % (len(stream), sum(r["kind"] == "new" for r in stream),
           sum(r["kind"] == "trace" for r in stream), sum(r["kind"] == "refit" for r in stream),
           args.rate))
    PID = lbvh_amd.LEAF_PRIMITIVE_ID
    setup = lbvh_amd.Context(0)
    setup.set_autotune(False)
    verts = {}
    reqs = []  # (tri or None, rays)
    trees = {}
    for i, r in enumerate(stream):
        k = r["scene"]
        tri = None
        if r["kind"] == "new":
            verts[k] = W.cluster_vertices((5 << 32) | k, r["n_tri"])
            tri = W.make_triangles(verts[k])
            trees[k] = setup.build_triangles(tri, code_bits=63, leaf_order=PID)
        elif r["kind"] == "refit":
            verts[k] = W.perturb((5 << 32) | (1 << 20) | i, verts[k], 1e-4)
            tri = W.make_triangles(verts[k])
            setup.refit_triangles(trees[k], tri)
        m = r["n_rays"]
        half = m // 2
        w = int(np.ceil(np.sqrt(max(half, 1) * 16.0 / 9.0)))
        h = max(1, (half + w - 1) // w)
        prim_rays = W.primary_rays(w, h)[:half]
        hp, ht, _ = setup.trace(trees[k], prim_rays, 0, with_uv=False)
        cur_tri = W.make_triangles(verts[k])
        bounce, _ = W.bounce_rays((5 << 32) | (2 << 20) | i, cur_tri, prim_rays, hp, ht)
        need = m - half
        if bounce.shape[0] == 0:
            extra = W.primary_rays(w, h + 1)
            bounce = extra[np.arange(need) % extra.shape[0]]
        else:
            bounce = bounce[np.arange(need) % bounce.shape[0]]
        reqs.append((hip.DeviceArray.from_host(tri) if tri is not None else None,
                     hip.DeviceArray.from_host(np.concatenate([prim_rays, bounce]))))
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

    # ---- served ------------------------------------------------------------
    srv = lbvh_amd.Server(0, code_bits=63, leaf_order=PID, isect_mode=0)
    # warm the lanes and pools outside the measured stream
    warm = outputs(reqs[0][1].shape[0])
    wt, ws = srv.submit(reqs[0][1], 0, reqs[0][0], out=warm)
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
            tickets.put((i, tk, t0))

    th = threading.Thread(target=submitter)
    th.start()
    served = [None] * len(stream)
    lat_s, first_sub, last_done, t0_sub = [], None, 0.0, None
    coalesced = []
    for _ in stream:
        i, tk, t0 = tickets.get()
        p, t, uv, info = srv.wait(tk)
        served[i] = True
        lat_s.append((info["done_us"] - info["submit_us"]) * 1e-3)
        coalesced.append(info["coalesced_with"])
        first_sub = info["submit_us"] if first_sub is None else min(first_sub, info["submit_us"])
        last_done = max(last_done, info["done_us"])
    th.join()
    sst = srv.stats()
    srv.close()
    served_span_s = (last_done - first_sub) * 1e-6
    log("bench: served makespan %.1f ms, p50 %.2f ms, p99 %.2f ms, %d launches"
        % (served_span_s * 1e3, percentile(lat_s, 50), percentile(lat_s, 99), sst["launches"]))

    # ---- plain, one request at a time ----------------------------------------
    ctx = lbvh_amd.Context(0)
    ctx.set_autotune(False)
    w0 = ctx.build_triangles(reqs[0][0], code_bits=63, leaf_order=PID)
    ctx.trace(w0, reqs[0][1], 0, out=warm)
    ctx.synchronize()
    w0.destroy()
    trees = {}
    end = 0.0
    busy = 0.0
    lat_p = []
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
    res = {
        "metric": "build Mprims/s + closest-hit Mrays/s @1M tris; p99 latency batched vs plain",
        "value": round(served_mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": len(stream),
        "warmup": 1,
        "ms_per_step": round(served_span_s * 1e3 / len(stream), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "serving: %d requests, Poisson arrivals at %.0f req/s (seed 5); "
                               "new scenes log-uniform [1024, 4194304] clustered triangles, 30%% "
                               "reuse, 10%% refits; rays log-uniform [4096, 2097152], half "
                               "primary, half bounce; 63-bit codes, primitive-id leaves; "
                               "request data resident in HBM" % (len(stream), args.rate),
                   "requests": len(stream), "offered_load_rps": args.rate, "rays": total_rays},
        "served": {"makespan_ms": round(served_span_s * 1e3, 3),
                   "mrays_s": round(served_mrays, 3),
                   "latency_ms_p50": round(percentile(lat_s, 50), 3),
                   "latency_ms_p99": round(percentile(lat_s, 99), 3),
                   "traversal_launches": sst["launches"],
                   "mean_coalesced": round(float(np.mean(coalesced)), 2),
                   "max_coalesced": sst["max_coalesced"],
                   "memory_peak_gb": round(sst["memory_peak_bytes"] / 2**30, 3),
                   "evictions": sst["evictions"]},
        "plain": {"makespan_ms": round(plain_span_s * 1e3, 3), "mrays_s": round(plain_mrays, 3),
                  "busy_ms": round(busy * 1e3, 3),
                  "capacity_rps": round(len(stream) / busy, 1),
                  "latency_ms_p50": round(percentile(lat_p, 50), 3),
                  "latency_ms_p99": round(percentile(lat_p, 99), 3)},
        "throughput_vs_plain": round(served_mrays / plain_mrays, 3),
        "results_identical": bool(same),
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
        "metric": "build Mprims/s + closest-hit Mrays/s @1M tris; p99 latency batched vs plain",
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


def _tri96(v):
    """triangle<double> records {p0, e0 = a - b, e1 = c - a, normal} of FP64
    vertex triples."""
    import numpy as np
    v = np.asarray(v, dtype=np.float64).reshape(-1, 9)
    a, b, c = v[:, 0:3], v[:, 3:6], v[:, 6:9]
    e0 = a - b
    e1 = c - a
    return np.ascontiguousarray(np.concatenate([a, e0, e1, np.cross(e0, e1)], axis=1))


def _boxes48(tri96):
    """triangle_aabb_converter<double>: min / max of p0, p0 - e0, p0 + e1."""
    import numpy as np
    p0 = tri96[:, 0:3]
    p1 = p0 - tri96[:, 3:6]
    p2 = p0 + tri96[:, 6:9]
    return np.ascontiguousarray(np.concatenate(
        [np.minimum(np.minimum(p0, p1), p2), np.maximum(np.maximum(p0, p1), p2)], axis=1))


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
    tri96 = _tri96(v)
    del v
    rng = np.random.default_rng(1)
    lo = rng.random((1 << 24, 3))
    big = np.concatenate([lo, lo + rng.random((1 << 24, 3)) * 5e-4], axis=1)
    del lo
    ctx = lbvh_amd.Context(0)
    out = {}
    for name, boxes in (("1m", _boxes48(tri96)), ("16m", big)):
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
        d.close()
        nodes.close()
        log("bench: FP64 build of %d boxes %.3f ms" % (n, build_ms))
    ctx.close()
    del big
    one = out["1m"]
    achieved = BUILD_BYTES_PER_PRIM_F64 * N_PRIMS / (one["build_ms"] * 1e-3) / 1e9
    res = {
        "metric": "build Mprims/s + closest-hit Mrays/s @1M tris; p99 latency batched vs plain",
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


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported "
              "(scenes shard across GPUs only by assigning whole scenes)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "large":
        return bench_large(args)
    if args.workload == "f64":
        return bench_f64(args)
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
    tune_log = ctx.tune_log()
    bvh.destroy()
    hp = p.numpy()
    ht = t.numpy()
    bounce, _ = W.bounce_rays(5, tri, rays, hp, ht)
    d_bounce = hip.DeviceArray.from_host(bounce)
    n_primary, n_bounce = rays.shape[0], bounce.shape[0]
    log("bench: %d primary rays, %d hits -> %d bounce rays" % (n_primary, int((hp != 0xFFFFFFFF).sum()),
                                                             n_bounce))
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
        split.append(ctx.build_triangles(d_tri, code_bits=63, leaf_order=lbvh_amd.LEAF_PRIMITIVE_ID))
    ctx.synchronize()
    split_ph, ns = ctx.phase_stats(1)
    ctx.set_timing(0)
    for b in split:
        b.destroy()

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
        "metric": "build Mprims/s + closest-hit Mrays/s @1M tris; p99 latency batched vs plain",
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
                                   for r in tune_log if not r["batch"]}},
        "build_roofline_frac": round(build_frac, 4),
        "build_algorithmic_bytes_per_prim": BUILD_BYTES_PER_PRIM_63,
        "closest_hit_mrays_s": round(mrays, 3),
        "trace_ms": round(trace_ms, 4),
        "dominant_kernel": dominant,
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        log("bench: timing the reference CPU path (bounded sample)")
        res["cpu_baseline"] = cpu_baseline(tri, rays, bounce, args.cpu_seconds)
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
