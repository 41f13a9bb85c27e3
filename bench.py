#!/usr/bin/env python3
"""Benchmark: traced ray segments per second of the Monte Carlo flux trace on
one MI355X.

Workload (default, ``--workload hole``): a triangle-mesh cylindrical hole,
radius 0.5, depth 5 (aspect 10), in a 2 x 2 periodic cell, 199,360
triangles; sticking s in {0.01, 0.1, 1} x source exponent n in {1, 10}; each
of the six (s, n) cases traces ``--rays`` primary rays (1e8 by default). One
step traces all six cases as one persistent launch replayed from a HIP graph;
inputs are resident in HBM before the timed region. The plain path (one
thread per ray, one launch per case) is timed once beside it, and the CPU
baseline (the oracle restatement of the reference tracer, OpenMP on the host
cores) on a bounded sample of the same cases.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
S_VALUES = (0.01, 0.1, 1.0)
N_VALUES = (1.0, 10.0)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def hole_cases(rays):
    from rti_amd import geometry
    sc = geometry.hole_triangles()
    cases = []
    for s in S_VALUES:
        for n in N_VALUES:
            cases.append(dict(scene=sc, n_rays=int(rays), sticking=s, cos_exponent=n, bc_x=1, bc_y=1, seed=2))
    return sc, cases


def bytes_per_segment(counters, prim_bytes):
    """Algorithmic bytes per segment from an instrumented run.

    64 B per BVH node visit (one node record), ``prim_bytes`` per primitive
    test (triangle 48 B, disc 32 B), 36 B per neighbour test (disc record +
    column index), 48 B per deposit (read + write of three 8-byte
    accumulators), 8 B per front hit (neighbour row pointers).
    """
    seg, hits = float(counters[0]), float(counters[1])
    nodes, prims, nbt, deps = (float(x) for x in counters[4:8])
    total = 64.0 * nodes + prim_bytes * prims + 36.0 * nbt + 48.0 * deps + 8.0 * hits
    return total / seg, {"V_node": nodes / seg, "V_prim": prims / seg, "K_nb": nbt / seg, "D": deps / seg}


def cpu_baseline(cases, budget_s=20.0):
    """Oracle (restated reference tracer, OpenMP) on a bounded ray sample."""
    import oracle
    from rti_amd import geometry
    threads = os.cpu_count() or 1
    segs = 0
    secs = 0.0
    sample = []
    per_case = max(1000, int(2e4))
    # calibrate: grow the sample until the budget is used (about budget_s seconds)
    for c in cases:
        box, sz = geometry.scene_box(c["scene"])
        r = oracle.trace(c["scene"], n_rays=per_case, box=box, src_z=sz, sticking=c["sticking"],
                         cos_exponent=c["cos_exponent"], bc_x=c["bc_x"], bc_y=c["bc_y"], seed=c["seed"],
                         threads=threads)
        segs += r["segments"]
        secs += r["seconds"]
        sample.append(per_case)
    if secs < budget_s * 0.5:
        scale = min(50.0, budget_s / max(secs, 1e-3))
        segs = 0
        secs = 0.0
        sample = []
        for c in cases:
            nr = int(per_case * scale)
            box, sz = geometry.scene_box(c["scene"])
            r = oracle.trace(c["scene"], n_rays=nr, box=box, src_z=sz, sticking=c["sticking"],
                             cos_exponent=c["cos_exponent"], bc_x=c["bc_x"], bc_y=c["bc_y"], seed=c["seed"],
                             threads=threads)
            segs += r["segments"]
            secs += r["seconds"]
            sample.append(nr)
    return {"value": segs / secs, "unit": "segments/s", "cores": threads, "kind": "port",
            "sample": f"oracle on {sum(sample)} rays ({sample[0]} per (s, n) case, same geometry), "
                      f"{segs} segments in {secs:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rays", type=float, default=1e8, help="primary rays per (s, n) case")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--lds", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the trace shards by ray-id range with one int64 all-reduce "
              "(design only); run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from rti_amd import Context
    sc, cases = hole_cases(args.rays)
    log(f"scene: {sc['tris'].shape[0]} triangles, {len(cases)} cases x {int(args.rays)} rays")
    ctx = Context(0)
    ctx.batch_set(cases)
    knobs = dict(block_size=args.block, lds_stack=args.lds)
    ctx.set_params(**knobs)
    t0 = time.time()
    ctx.prepare()
    log(f"build {time.time() - t0:.2f} s (first call includes HIP init)")

    # warmup + timed replays of the captured graph
    if args.warmup:
        ctx.batch_replay(args.warmup)
    t0 = time.perf_counter()
    ctx.batch_replay(args.steps)
    wall = time.perf_counter() - t0
    st = ctx.batch_stats()
    ms_per_step = st["trace_ms"]
    counters = ctx.counters()
    segments = int(counters[0])
    log(f"persistent: {ms_per_step:.1f} ms/step (wall {1e3 * wall / args.steps:.1f}), {segments} segments/step")
    value = segments / (ms_per_step / 1e3)

    # plain path, once
    plain = None
    if not args.no_plain:
        ctx.set_params(kernel=1, **knobs)
        ctx.batch_replay(1)
        pst = ctx.batch_stats()
        pc = ctx.counters()
        plain = {"segments_per_s": int(pc[0]) / (pst["trace_ms"] / 1e3), "ms_per_step": pst["trace_ms"]}
        log(f"plain: {pst['trace_ms']:.1f} ms/step")

    # instrumented run for the algorithmic bytes, on 1/20 of the rays
    cnt_ctx = Context(0)
    small = [dict(c, n_rays=max(1, int(c["n_rays"] // 20))) for c in cases]
    cnt_ctx.batch_set(small)
    cnt_ctx.set_params(kernel=2, **knobs)
    cnt_ctx.batch_run()
    B, parts = bytes_per_segment(cnt_ctx.counters(), 48.0)
    cnt_ctx.close()
    achieved = value * B / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_segment": round(B, 1), **{k: round(v, 3) for k, v in parts.items()}}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cases)
        log(f"cpu baseline {cpu['value']:.3e} segments/s on {cpu['cores']} threads")

    out = {
        "metric": "ray segments/sec @1 MI355X, batched vs plain GPU vs CPU; flux err vs radiosity",
        "value": value,
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "triangle-mesh cylindrical hole r=0.5 depth 5 (AR 10), 2x2 periodic cell, "
                               f"{sc['tris'].shape[0]} triangles; s in {list(S_VALUES)} x n in {list(N_VALUES)}, "
                               f"{int(args.rays)} rays each; batched persistent launch, HIP graph",
                   "rays_per_step": int(args.rays) * len(cases), "segments_per_step": segments},
        "plain": plain,
        "batched_vs_plain": (value / plain["segments_per_s"]) if plain else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
