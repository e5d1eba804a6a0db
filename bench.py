"""Benchmark of the MI355X SVO builder on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batch|single|soup]

Default workload (the metric's configuration): 256 mixed meshes -- icospheres
of subdivision 2-6 and tori, 320-81,920 triangles log-uniform, resolution
drawn from {64, 128, 256} (seeded) -- with the meshes already resident in
device memory.  A step is one full build of all 256 octrees (voxelize ->
sorted unique voxels -> levels -> node records -> neighbors -> vertex/edge
records, reference semantics) in one batched pass; the same meshes built one
per pass (the plain path) are timed beside it.  Output: one JSON line.

`roofline` is computed for the kernel stage that takes the most device time,
from HIP events recorded on the build's stream, and its algorithmic bytes.
`cpu_baseline` times the CPU restatement of the reference (the oracle) on a
bounded sample of the same meshes on this host, on every core the process may
use (the single-threaded figure, the host's CPU count and model beside it).
A short served stream (Poisson arrivals, res 64-512) follows the batch and
its p50 / p99 latency go into the same line.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "gpu-svo_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "meshes/sec + voxels/sec at res 256, batched vs plain; p99 latency; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stage_bytes(stats, name):
    """Algorithmic HBM bytes of one pipeline stage (read + write, each once)."""
    kb = stats["key_bytes"]
    H, U, N = stats["n_hits"], stats["n_voxels"], stats["n_nodes"]
    B = stats["bitmap_bytes"]
    if name == "sort":
        return 2 * kb * H * stats["sort_passes"]
    if name == "bitmap_clear":
        return stats["bitmap_fill_bytes"]
    if name == "voxelize":
        # triangles in; raw hits out (keys, or one bitmap word update each)
        return 36 * stats["n_tris"] + (4 if B else kb) * H
    if name == "unique_level1":
        # sorted hits (or the bitmap) in; unique keys, parent keys and first
        # children out; the bitmap's nonzero 16-byte pieces zeroed
        return ((B if B else kb * H) + kb * U + (kb + 4) * stats["n_groups1"] +
                stats["bitmap_reset_bytes"])
    if name == "levels":
        return (kb + 4) * stats["level_elems"]
    if name == "neighbor_fill":
        return 108 * N
    if name == "nodes":
        # records written (code, is_leaf, origin, width, parent, childs), the
        # level keys read, and the voxel codes copied out
        return (2 * kb + 53) * N + 2 * kb * U
    if name == "attr_write":
        # vertex / edge / face records written, node origin and width read
        return 16 * stats["n_vertices"] + 28 * stats["n_edges"] + 28 * stats["n_faces"] + 16 * N
    if name == "attr_count":
        return 32 * N
    return 0


# Measured HBM traffic per algorithmic byte of each stage's kernels on the
# default workload (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch,
# scaled by the same counters on a device-to-device copy of 1 GiB: FETCH_SIZE
# x 2, WRITE_SIZE x 1; profiles/batch256_hbm_traffic.md).  Used to fill
# roofline.traffic for this run's launch.
TRAFFIC_PER_ALG_BYTE = {"attr_write": 1.000, "neighbor_fill": 1.000, "nodes": 1.00,
                        "unique_level1": 1.00, "voxelize": 4.2}


def dedup_gbs(stats, stages):
    """HBM rate of the de-duplication stages (bitmap clear or sort, and the
    compaction into level 1): algorithmic bytes over their device time."""
    names = [k for k in ("bitmap_clear", "sort", "unique_level1") if k in stages]
    t = sum(stages[k] for k in names)
    return round(sum(stage_bytes(stats, k) for k in names) / (t * 1e-3) / 1e9, 1) if t else None


def compulsory_bytes(stats):
    """End-to-end compulsory bytes of a step: the triangles read once and
    every output written once (the unique voxel codes, the node arrays code,
    is_leaf, origin, width, parent, childs and neighbors, and the vertex,
    edge and face records), with nothing re-read."""
    kb, N = stats["key_bytes"], stats["n_nodes"]
    return (36 * stats["n_tris"] + kb * stats["n_voxels"] + (kb + 161) * N +
            16 * stats["n_vertices"] + 28 * stats["n_edges"] + 28 * stats["n_faces"])


def compulsory_gbs(stats, ms_per_step):
    """compulsory_bytes over the whole step's wall time (GB/s): extra passes
    inside the step can only lower it."""
    return round(compulsory_bytes(stats) / (ms_per_step * 1e-3) / 1e9, 1)


def pass_stats(ctx):
    import ctypes as C
    import svo_amd
    L = svo_amd.lib()

    class PassStats(C.Structure):
        _fields_ = [(k, C.c_uint64) for k in (
            "n_meshes", "n_tris", "box_cells", "n_hits", "n_voxels", "level_elems", "n_nodes",
            "n_vertices", "n_edges", "n_faces", "out_bytes")] + \
            [(k, C.c_uint32) for k in ("key_bytes", "sort_bits", "sort_passes", "n_levels")] + \
            [(k, C.c_uint64) for k in ("bitmap_bytes", "n_groups1", "bitmap_fill_bytes",
                                       "bitmap_reset_bytes")]
    fn = L.svo_ctx_last_pass_stats
    fn.argtypes = [C.c_void_p, C.POINTER(PassStats)]
    ps = PassStats()
    fn(ctx.handle, C.byref(ps))
    return {k: getattr(ps, k) for k, _ in PassStats._fields_}


def cpu_threads():
    """Host threads for the multi-threaded CPU baseline: every core this
    process may run on."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(meshes, budget_s=12.0):
    """Oracle (CPU restatement of the reference) on a bounded sample of the meshes:
    single-threaded until `budget_s` of build time, then the same sample with
    OpenMP on the host threads."""
    import pyoracle as po
    done = 0
    t1 = 0.0
    for P, T, res in meshes:
        r = po.build(P, T, res, threads=1, fields=["voxels"])
        t1 += r.info.build_seconds
        done += 1
        if t1 > budget_s:
            break
    # all host threads: meshes are independent, so each thread builds whole
    # meshes (the oracle releases the interpreter lock while it runs)
    from concurrent.futures import ThreadPoolExecutor
    nt = cpu_threads()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(nt) as ex:
        list(ex.map(lambda m: po.build(m[0], m[1], m[2], threads=1, fields=["voxels"]),
                    meshes[:done]))
    tn = time.perf_counter() - t0
    return {"value": round(done / tn, 3), "unit": "meshes/s", "cores": nt, "kind": "port",
            "host_cpus": os.cpu_count(), "cpu_model": cpu_model(),
            "single_thread_value": round(done / t1, 3),
            "sample": f"first {done} meshes of the same 256-mesh set, build only "
                      f"(voxelize -> vertex/edge): {t1:.1f} s on 1 thread, {tn:.1f} s wall "
                      f"on {nt} threads (one mesh per thread)"}


TUNING_FILE = os.path.join(ROOT, "gpu-svo_amd", "tuning", "mi355x.tuning")


def load_tuning(ctx, args):
    """Loads the tuning file (the library's tuning search, run on this kind of
    device by gpu-svo_amd/tuning/make_tuning.py) onto the context; returns
    what it set, for the JSON line."""
    path = args.tuning
    if path == "none" or not os.path.exists(path):
        return None
    defaults = {}
    names = []
    with open(path) as f:
        for line in f:
            if line.strip():
                names.append(line.split()[0])
    for nm in names:
        defaults[nm] = ctx.get_option(nm)
    ctx.load_tuning(path)
    changed = {nm: ctx.get_option(nm) for nm in names if ctx.get_option(nm) != defaults[nm]}
    return {"file": os.path.relpath(path, ROOT), "options": changed}


def run_batch_workload(args):
    import numpy as np
    import svo_amd
    from svo_amd import meshgen

    log("bench: generating 256 mixed meshes")
    meshes = meshgen.mixed_batch(256, seed=3)
    n_tris = sum(int(T.shape[0]) for _, T, _ in meshes)
    ctx = svo_amd.Context(0)
    tuning = load_tuning(ctx, args)  # explicit options below take precedence
    if args.batch_lanes:
        ctx.set_option("batch_lanes", args.batch_lanes)
    for opt in args.option:
        name, value = opt.split("=", 1)
        ctx.set_option(name, int(value))
    dms = [ctx.upload_mesh(P, T) for P, T, _ in meshes]
    descs = [dm.desc(res) for dm, (_, _, res) in zip(dms, meshes)]
    if args.autotune:
        log("bench: tuning search on the batch (untimed)")
        tuning = {"search": ctx.autotune(descs, on_device=True, reps=5)}

    sizes = {}

    def batch_step():
        trees, st = ctx.build_batch(descs, on_device=True)
        bad = [s for s in st if s != 0]
        if bad:
            raise RuntimeError(f"batched build failed: statuses {set(bad)}")
        nv = sum(t.info.n_voxels for t in trees)
        nn = sum(t.info.n_nodes for t in trees)
        sizes["vertices"] = sum(t.info.n_vertices for t in trees)
        sizes["edges"] = sum(t.info.n_edges for t in trees)
        for t in trees:
            t.free()
        return nv, nn

    def plain_step():
        nv = nn = 0
        for d in descs:
            t = ctx.build_device(*d)
            nv += t.info.n_voxels
            nn += t.info.n_nodes
            t.free()
        return nv, nn

    for i in range(args.warmup):
        log(f"bench: warmup batched {i}")
        batch_step()
    # timed steps without stage events; the same number of steps with HIP
    # events on every stage follows, for the stage times and the roofline
    log(f"bench: timing {args.steps} batched steps")
    step_ms = []
    mallocs0 = ctx.device_mallocs()
    gc.collect()
    gc.disable()  # no collector pause inside the timed steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        nv, nn = batch_step()
        step_ms.append((time.perf_counter() - ts) * 1e3)
    t1 = time.perf_counter()
    gc.enable()
    mallocs_timed = ctx.device_mallocs() - mallocs0
    ms_batch = (t1 - t0) * 1e3 / args.steps
    log(f"bench: {args.steps} batched steps with stage events")
    ctx.set_timing(True)
    stage_acc: dict = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch_step()
        for k, v in ctx.stage_ms().items():
            stage_acc[k] = stage_acc.get(k, 0.0) + v
    ms_events = (time.perf_counter() - t0) * 1e3 / args.steps
    ctx.set_timing(False)
    stats = pass_stats(ctx)
    stages = {k: v / args.steps for k, v in stage_acc.items()}

    graph = None
    if not args.no_graph:
        # the same batch recorded once and replayed as one HIP graph per step
        # (svo_plan_*): identical outputs, one host synchronization per step
        log("bench: batch plan (recording, capture)")
        plan = ctx.plan(descs, on_device=True)
        replays = []

        def graph_step():
            trees, st, replayed = plan.run()
            if any(s != 0 for s in st):
                raise RuntimeError(f"plan replay failed: statuses {set(st)}")
            replays.append(replayed)
            for t in trees:
                t.free()

        for _ in range(args.warmup):
            graph_step()
        log(f"bench: timing {args.steps} graph replays")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            graph_step()
        ms_graph = (time.perf_counter() - t0) * 1e3 / args.steps
        plan.close()
        graph = {"ms_per_step": round(ms_graph, 3),
                 "meshes_per_s": round(256 / (ms_graph * 1e-3), 2),
                 "all_replayed": all(replays),
                 "vs_eager_batched": round(ms_batch / ms_graph, 3)}

    log("bench: plain path (one mesh per pass)")
    plain_step()  # warmup
    ps = max(1, args.plain_steps)
    t0 = time.perf_counter()
    for _ in range(ps):
        plain_step()
    t1 = time.perf_counter()
    ms_plain = (t1 - t0) * 1e3 / ps

    out = {
        "metric": METRIC,
        "value": round(256 / (ms_batch * 1e-3), 2),
        "unit": "meshes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_batch, 3),
        "ms_per_step_with_stage_events": round(ms_events, 3),
        "step_ms": [round(x, 3) for x in step_ms],
        "device_mallocs_in_timed_steps": mallocs_timed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32" if stats["key_bytes"] == 4 else "u64",
        "data": "synthetic",
        "config": {
            "workload": "256 mixed meshes (icospheres subdiv 2-6 + tori, 320-81,920 triangles "
                        "log-uniform, res from {64,128,256}, seed 3), one batched pass per step, "
                        "reference semantics, inputs resident in HBM",
            "n_meshes": 256, "n_tris": n_tris, "parallelism": "single GPU",
        },
        "voxels_per_s": round(nv / (ms_batch * 1e-3)),
        "nodes_per_s": round(nn / (ms_batch * 1e-3)),
        "plain_meshes_per_s": round(256 / (ms_plain * 1e-3), 2),
        "plain_ms_per_pass_of_256": round(ms_plain, 3),
        "batched_vs_plain": round(ms_plain / ms_batch, 3),
        "graph_replay": graph,
        "sizes": dict({"voxels": nv, "nodes": nn, "hits": stats["n_hits"]}, **sizes),
        "passes": ctx.get_option("batch_lanes"),
        "tuning": tuning,
        "options": {o.split("=", 1)[0]: int(o.split("=", 1)[1]) for o in args.option} or None,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "dedup_gbs": dedup_gbs(stats, stages),
        "compulsory_gbs": compulsory_gbs(stats, ms_batch),
        "dedup_path": "bitmap" if stats["bitmap_bytes"] else "sort",
        "roofline": _roofline(stats, stages),
    }
    for dm in dms:
        dm.free()
    ctx.close()
    if args.served_requests > 0:
        # the metric's latency figures: a short served stream of the same
        # kind of meshes (Poisson arrivals, continuous batching)
        out["served"] = served_stream(args.served_requests, args.served_warmup_batch,
                                      args.max_batch, args.lanes, args.max_wait_us)
        out["p50_ms"] = out["served"]["p50_ms"]
        out["p99_ms"] = out["served"]["p99_ms"]
    if not args.no_cpu:
        log("bench: CPU baseline (oracle, bounded sample)")
        out["cpu_baseline"] = cpu_baseline(meshes)
    else:
        out["cpu_baseline"] = None
    return out


def _roofline(stats, stages):
    """Roofline of the stage with the most device time in the last pass:
    algorithmic bytes of its launch over its HIP-event time on the pass's
    stream; traffic = measured bytes per algorithmic byte x this launch."""
    dev = {k: v for k, v in stages.items() if not k.startswith("setup") and k != "start"}
    dom = max(dev, key=dev.get)
    b = stage_bytes(stats, dom)
    if dom == "nodes" and "neighbor_fill" not in stages:
        b += stage_bytes(stats, "neighbor_fill")  # concurrent passes: the node kernels fill
    ach = b / (stages[dom] * 1e-3) / 1e9 if stages[dom] > 0 else 0.0
    ratio = TRAFFIC_PER_ALG_BYTE.get(dom)
    return {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": round(ratio * b) if ratio else None, "algorithmic_bytes": b,
            "stage_ms": round(stages[dom], 4)}


def run_single_workload(args, name, P, T, res, key_bits, flags=0, desc=""):
    """One large mesh per step (plain path), inputs resident in HBM."""
    import svo_amd
    ctx = svo_amd.Context(0)
    tuning = load_tuning(ctx, args)  # explicit options below take precedence
    for opt in args.option:
        oname, value = opt.split("=", 1)
        ctx.set_option(oname, int(value))
    ctx.set_timing(True)
    dm = ctx.upload_mesh(P, T)
    stage_acc: dict = {}

    def step():
        t = ctx.build_device(*dm.desc(res), key_bits=key_bits, flags=flags)
        info = (t.info.n_voxels, t.info.n_nodes, t.info.depth)
        t.free()
        return info

    for i in range(args.warmup):
        log(f"bench: warmup {name} {i}")
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nv, nn, depth = step()
        for k, v in ctx.stage_ms().items():
            stage_acc[k] = stage_acc.get(k, 0.0) + v
    t1 = time.perf_counter()
    stats = pass_stats(ctx)
    ms = (t1 - t0) * 1e3 / args.steps
    stages = {k: v / args.steps for k, v in stage_acc.items()}
    out = {
        "metric": METRIC, "value": round(1e3 / ms, 3), "unit": "meshes/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32" if stats["key_bytes"] == 4 else "u64", "data": "synthetic",
        "config": {"workload": desc, "n_tris": int(T.shape[0]), "res": res, "key_bits": key_bits,
                   "parallelism": "single GPU"},
        "voxels_per_s": round(nv / (ms * 1e-3)), "nodes_per_s": round(nn / (ms * 1e-3)),
        "sizes": {"voxels": nv, "nodes": nn, "depth": depth, "hits": stats["n_hits"],
                  "box_cells": stats["box_cells"], "out_gb": round(stats["out_bytes"] / 1e9, 3)},
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "dedup_gbs": dedup_gbs(stats, stages),
        "compulsory_gbs": compulsory_gbs(stats, ms),
        "options": {o.split("=", 1)[0]: int(o.split("=", 1)[1]) for o in args.option} or None,
        "tuning": tuning,
        "dedup_path": "bitmap" if stats["bitmap_bytes"] else "sort",
        "roofline": _roofline(stats, stages),
    }
    dm.free()
    ctx.close()
    if not args.no_cpu:
        import pyoracle as po
        nt = cpu_threads()
        log(f"bench: CPU baseline for {name} (oracle, 1 and {nt} threads)")
        if T.shape[0] > 400_000:
            # bounded sample: the oracle holds every output array in host memory
            # (~650 B per node); time a prefix of the triangles and scale by the
            # voxel count
            k = 100_000
            r1 = po.build(P, T[:k], res, key_bits=key_bits, threads=1, fields=["voxels"])
            rn = po.build(P, T[:k], res, key_bits=key_bits, threads=nt, fields=["voxels"])
            scale = nv / max(1, rn.info.n_voxels)
            est1 = r1.info.build_seconds * scale
            estn = rn.info.build_seconds * scale
            out["cpu_baseline"] = {
                "value": round(1.0 / estn, 5), "unit": "meshes/s", "cores": nt, "kind": "port",
                "single_thread_value": round(1.0 / est1, 5),
                "sample": f"first {k} of {T.shape[0]} triangles, build only: "
                          f"{r1.info.build_seconds:.1f} s on 1 thread, "
                          f"{rn.info.build_seconds:.1f} s on {nt} threads for "
                          f"{rn.info.n_voxels} voxels, scaled to {nv} voxels"}
        else:
            r1 = po.build(P, T, res, key_bits=key_bits, threads=1, fields=["voxels"])
            rn = po.build(P, T, res, key_bits=key_bits, threads=nt, fields=["voxels"])
            out["cpu_baseline"] = {"value": round(1.0 / rn.info.build_seconds, 4),
                                   "unit": "meshes/s", "cores": nt, "kind": "port",
                                   "single_thread_value": round(1.0 / r1.info.build_seconds, 4),
                                   "sample": f"the same mesh, build only: "
                                             f"{r1.info.build_seconds:.1f} s on 1 thread, "
                                             f"{rn.info.build_seconds:.1f} s on {nt} threads"}
    else:
        out["cpu_baseline"] = None
    return out


def served_stream(n, n_warm, max_batch, lanes, max_wait_us):
    """Poisson request stream at 2000 req/s (95% mixed meshes, 5% icosphere-4,
    res 64-512) admitted under a 64 GB budget (continuous batching), 1000
    neighbor queries per resident tree.  An untimed warm-up stream of the same
    kind (another seed) runs first, as the other workloads run warm-up steps:
    the first requests of a process pay for first-time device allocations
    (reported as cold_p99_ms).  Latency = completion - submission."""
    import numpy as np
    import svo_amd
    from svo_amd import meshgen
    cache = {}

    def stream(k, seed):
        out = []
        for t_arr, s in meshgen.request_stream(k, rate=2000.0, seed=seed):
            key = (s.kind, s.params)
            if key not in cache:
                cache[key] = s.build()
            out.append((t_arr, cache[key][0], cache[key][1], s.res))
        return out

    meshes = stream(n, 4)
    warm = stream(n_warm, 14) if n_warm > 0 else []
    ctx = svo_amd.Context(0)
    # warm the lanes (kernel code objects, workspaces)
    P0, T0 = meshgen.icosphere(3)
    ctx.build(P0, T0, 64).free()
    srv = svo_amd.Server(ctx, mem_budget=64 << 30, max_batch=max_batch, n_lanes=lanes,
                         max_wait_us=max_wait_us)
    rng = np.random.default_rng(5)

    def run(reqs):
        lat, bsz, wait = [], [], []
        done = 0
        queries = 0
        t_start = time.perf_counter()
        t_last = t_start
        gap = [0.0, 0.0]  # longest time without a completion while requests were pending, and when
        i = 0
        while done < len(reqs):
            now = time.perf_counter() - t_start
            while i < len(reqs) and reqs[i][0] <= now:
                _, P, T, res = reqs[i]
                srv.submit(P, T, res)
                i += 1
            got = srv.poll(cap=256, timeout_us=200 if i >= len(reqs) else 0)
            t_now = time.perf_counter()
            if got or i == done:
                if got and t_now - t_last > gap[0]:
                    gap[0], gap[1] = t_now - t_last, t_last - t_start
                t_last = t_now
            for d in got:
                if d.status != 0:
                    raise RuntimeError(f"request {d.request_id} failed: {d.status}")
                lat.append(d.t_done - d.t_submit)
                wait.append(d.t_start - d.t_submit)
                bsz.append(d.batch_size)
                nn = d.tree.info.n_nodes
                q = d.tree.query_neighbors(rng.integers(0, nn, 1000).astype(np.uint32),
                                           rng.integers(0, 27, 1000).astype(np.uint8))
                queries += q.size
                d.tree.free()
                done += 1
        return lat, bsz, wait, queries, time.perf_counter() - t_start, gap

    cold_p99 = None
    if warm:
        log(f"bench: served warm-up stream of {len(warm)} requests (untimed)")
        wl, _, _, _, _, _ = run(warm)
        cold_p99 = round(float(np.percentile(np.array(wl) * 1e3, 99)), 3)
    log(f"bench: served stream of {len(meshes)} requests")
    lat, bsz, wait, queries, wall, gap = run(meshes)
    srv.close()
    ctx.close()
    lat = np.array(lat) * 1e3
    return {
        "requests": n, "warmup_requests": len(warm), "wall_ms": round(wall * 1e3, 3),
        "meshes_per_s": round(n / wall, 2),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "cold_p99_ms": cold_p99,
        "queue_wait_p50_ms": round(float(np.percentile(np.array(wait) * 1e3, 50)), 3),
        "queue_wait_p99_ms": round(float(np.percentile(np.array(wait) * 1e3, 99)), 3),
        "mean_batch": round(float(np.mean(bsz)), 2), "queries": queries,
        "longest_completion_gap_ms": round(gap[0] * 1e3, 3),
        "longest_gap_at_ms": round(gap[1] * 1e3, 3),
        "server": {"lanes": lanes, "max_batch": max_batch, "max_wait_us": max_wait_us,
                   "budget_gb": 64},
    }


def run_served_workload(args):
    """The served stream alone: one JSON line whose value is its meshes/s."""
    r = served_stream(args.requests, args.served_warmup, args.max_batch, args.lanes,
                      args.max_wait_us)
    n = r["requests"]
    return {
        "metric": METRIC, "value": r["meshes_per_s"], "unit": "meshes/s", "n_gpus": 1,
        "steps": 1, "warmup": r["warmup_requests"], "ms_per_step": r["wall_ms"],
        "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"served: {n} requests, Poisson 2000 req/s, 95% mixed meshes + 5% "
                               f"icosphere-4, res 64-512, 64 GB budget, 1000 neighbor queries "
                               f"per resident tree, after an untimed warm-up stream of "
                               f"{r['warmup_requests']} requests", "parallelism": "single GPU"},
        "cold_p99_ms": r["cold_p99_ms"], "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
        "queue_wait_p50_ms": r["queue_wait_p50_ms"], "queue_wait_p99_ms": r["queue_wait_p99_ms"],
        "server": r["server"], "mean_batch": r["mean_batch"], "queries": r["queries"],
        "roofline": None, "cpu_baseline": None,
    }


def run_io_workload(args):
    """OBJ in and out at speed: parse a large mesh file with the library's
    parallel reader (BaseModel::loadOBJ rules) and write a tree with its
    parallel byte-exact writer, each beside the oracle's sequential one."""
    import tempfile
    import numpy as np
    import pyoracle as po
    import svo_amd
    from svo_amd import meshgen
    d = tempfile.mkdtemp(prefix="svo_io_")
    P, T = meshgen.noisy_icosphere(7, seed=1)
    src = os.path.join(d, "ico7.obj")
    with open(src, "w") as f:
        f.write("".join("v %.9g %.9g %.9g\n" % tuple(p) for p in P.astype(np.float64)))
        f.write("".join("f %d %d %d\n" % tuple(t + 1) for t in T))
    in_bytes = os.path.getsize(src)
    log(f"bench: parsing {in_bytes / 1e6:.1f} MB")
    def best_of(fn, k):
        ts = []
        for _ in range(k):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)
    t_parse = best_of(lambda: svo_amd.load_obj(src), max(1, args.steps))
    t_parse_ref = best_of(lambda: po.load_obj(src), 1)
    P2, T2 = svo_amd.load_obj(src)
    assert np.array_equal(P2, po.load_obj(src)[1]) and np.array_equal(T2, po.load_obj(src)[2])
    ctx = svo_amd.Context(0)
    tree = ctx.build(P, T, 256)
    out = os.path.join(d, "gpu")
    os.makedirs(out)
    t_write = best_of(lambda: tree.write_obj(out, "ico7"), max(1, args.steps))
    path = tree.write_obj(out, "ico7")
    out_bytes = os.path.getsize(path)
    ref_path = os.path.join(d, "ref_tree.obj")
    t0 = time.perf_counter()
    st, _ = po.write_obj(P, T, 256, ref_path)
    t_write_ref = time.perf_counter() - t0  # includes the oracle's build
    same = open(path, "rb").read() == open(ref_path, "rb").read()
    n_nodes = tree.info.n_nodes
    tree.free()
    ctx.close()
    for f in os.listdir(out):
        os.remove(os.path.join(out, f))
    os.rmdir(out)
    for f in (src, ref_path):
        os.remove(f)
    os.rmdir(d)
    return {
        "metric": "OBJ parse and tree write MB/s", "value": round(out_bytes / t_write / 1e6, 1),
        "unit": "MB/s written", "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": round(t_write * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "text", "data": "synthetic",
        "config": {"workload": "parse icosphere-7 + noise as OBJ text (327,680 faces); write its "
                               "res-256 tree as <stem>_<depth>_tree.obj",
                   "parallelism": "host threads"},
        "parse_mb_s": round(in_bytes / t_parse / 1e6, 1),
        "parse_mb_s_oracle": round(in_bytes / t_parse_ref / 1e6, 1),
        "write_mb_s": round(out_bytes / t_write / 1e6, 1),
        "write_s_oracle_incl_build": round(t_write_ref, 2),
        "write_mb_s_oracle_incl_build": round(out_bytes / t_write_ref / 1e6, 1),
        "sizes": {"obj_in_mb": round(in_bytes / 1e6, 1), "tree_obj_mb": round(out_bytes / 1e6, 1),
                  "nodes": n_nodes},
        "byte_identical_to_oracle": same, "roofline": None, "cpu_baseline": None,
    }


def copy_probe(nbytes):
    """One device-to-device copy of nbytes (reads nbytes, writes nbytes)."""
    import ctypes as C
    import svo_amd
    L = svo_amd.lib()
    ctx = svo_amd.Context(0)
    a, b = C.c_void_p(), C.c_void_p()
    svo_amd._check(L.svo_device_alloc(ctx.handle, nbytes, C.byref(a)), "probe alloc")
    svo_amd._check(L.svo_device_alloc(ctx.handle, nbytes, C.byref(b)), "probe alloc")
    svo_amd._check(L.svo_device_copy(ctx.handle, b, a, nbytes, 2), "probe copy")
    L.svo_device_free(ctx.handle, a, nbytes)
    L.svo_device_free(ctx.handle, b, nbytes)
    ctx.close()
    log(f"bench: copy probe of {nbytes} bytes done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--workload", default="batch",
                    choices=["batch", "single", "soup", "soup2048", "served", "io"])
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--served-warmup", type=int, default=500,
                    help="served: requests of an untimed warm-up stream run first")
    ap.add_argument("--served-requests", type=int, default=800,
                    help="batch workload: requests of the served stream whose p50 / p99 "
                         "latency go into the line (0: none)")
    ap.add_argument("--served-warmup-batch", type=int, default=300,
                    help="batch workload: untimed warm-up requests of that stream")
    ap.add_argument("--max-batch", type=int, default=32)
    ap.add_argument("--max-wait-us", type=int, default=1000)
    ap.add_argument("--lanes", type=int, default=2, help="served: pipeline lanes")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="batch workload: skip the batch-plan (HIP graph) leg")
    ap.add_argument("--tuning", default=TUNING_FILE,
                    help="tuning file loaded on the build context before the --option "
                         "values ('none': library defaults)")
    ap.add_argument("--autotune", action="store_true",
                    help="batch workload: run the library's tuning search on the batch first")
    ap.add_argument("--batch-lanes", type=int, default=0,
                    help="concurrent passes a batch is split into (0: library default)")
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="batch and single-mesh workloads: a library option "
                         "(svo_ctx_set_option) to set first")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first copy this many MiB device-to-device once: a kernel of known "
                         "traffic for scaling HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(meshes shard across GPUs without exchange, see the design notes)",
              file=sys.stderr)
        sys.exit(2)
    from svo_amd import meshgen
    if args.copy_probe_mb > 0:
        copy_probe(args.copy_probe_mb << 20)
    if args.workload == "batch":
        out = run_batch_workload(args)
    elif args.workload == "single":
        log("bench: generating icosphere-7 with value noise")
        P, T = meshgen.noisy_icosphere(7, seed=1)
        out = run_single_workload(args, "ico7", P, T, 512, 32,
                                  desc="icosphere subdiv 7 (327,680 triangles) + 3-octave value "
                                       "noise 0.1, res 512, one mesh per pass")
    elif args.workload in ("soup", "soup2048"):
        log("bench: generating 2M-triangle soup")
        P, T = meshgen.soup(2_000_000, seed=2)
        if args.workload == "soup":
            out = run_single_workload(args, "soup1024", P, T, 1024, 32,
                                      desc="2,000,000-triangle soup (lognormal edges, median "
                                           "1/256), res 1024, uint32 keys")
        else:
            out = run_single_workload(args, "soup2048", P, T, 2048, 64, flags=1,
                                      desc="2,000,000-triangle soup, res 2048, uint64 keys, "
                                           "nodes and neighbors without vertex/edge records")
    elif args.workload == "io":
        out = run_io_workload(args)
    else:
        out = run_served_workload(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
