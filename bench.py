#!/usr/bin/env python3
"""Benchmark: traced ray segments per second of the Monte Carlo flux trace on
one MI355X. Prints one JSON line.

Workloads (``--workload``):
  hole   (default) triangle-mesh cylindrical hole, radius 0.5, depth 5
         (aspect 10), 2 x 2 periodic cell, 199,360 triangles; sticking
         s in {0.01, 0.1, 1} x source exponent n in {1, 10}; ``--rays``
         primary rays per (s, n) case (1e8).
  trench disc-cloud trench, width 1, depth 4, 0.02 grid (51,051 discs),
         reflective x, periodic y, source at z = 4.1, cosine source,
         s = 0.1, 1e6 rays (the reference's CPU-runnable case).
  batch  64 disc-cloud trenches and holes (aspect 2-20, 1e4-5e5 discs,
         s log-uniform 1e-3..1, diffuse or specular), ray budgets
         proportional to area, 1e6-5e7 per instance (6.4e8 in total);
         ``--rays`` rescales the batch total.
  rough  periodic rough surface over 16 holes of aspect 6 (about 4M discs),
         70/30 diffuse/specular, s = 0.02, ``--rays`` (1e9); also reports
         64-bit fixed point vs FP64 atomic accumulation on the same paths.
  adjoint disc-cloud trench of aspect 8 (192,001 discs at 0.01 spacing),
         s = 0.05, source exponent n = 5, ``--rays`` (1e8), with d/ds and d/dn
         accumulated in the same pass; also reports the summed derivatives
         against central finite differences with common random numbers
         (ds = 1e-3, dn = 1e-2) over 8 independent seeds, with standard errors.

A step traces the whole workload as one persistent launch replayed from a
HIP graph, inputs resident in HBM, launch knobs from the autotuner
(rti_amd.tune). The plain path (one thread per ray, one launch per
geometry) is timed once beside it, and the CPU baseline (the oracle
restatement of the reference tracer, OpenMP on the host cores) on a bounded
sample of the same cases.
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
METRIC = "ray segments/sec @1 MI355X, batched vs plain GPU vs CPU; flux err vs radiosity"


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


def workload(name, rays):
    from rti_amd import geometry
    if name == "hole":
        r = 1e8 if rays is None else rays
        sc, cases = hole_cases(r)
        desc = (f"triangle-mesh cylindrical hole r=0.5 depth 5 (AR 10), 2x2 periodic cell, {sc['tris'].shape[0]} "
                f"triangles; s in {list(S_VALUES)} x n in {list(N_VALUES)}, {int(r)} rays each")
        return cases, desc, 48.0
    if name == "trench":
        r = 1e6 if rays is None else rays
        sc = geometry.trench_discs()
        cases = [dict(scene=sc, n_rays=int(r), sticking=0.1, bc_x=0, bc_y=1, src_z=4.1, seed=1)]
        desc = (f"disc-cloud trench w=1 d=4 (AR 4), 0.02 grid, {sc['xyzr'].shape[0]} discs, reflective x, "
                f"periodic y, source z=4.1, cosine, s=0.1, {int(r)} rays")
        return cases, desc, 32.0
    if name == "batch":
        inst = geometry.batch_instances(seed=3, count=64, total_rays=rays)
        r = sum(i["n_rays"] for i in inst)
        cases = [dict(scene=i["scene"], n_rays=i["n_rays"], sticking=i["sticking"], diffuse_prob=i["diffuse_prob"],
                      bc_x=1, bc_y=1, seed=3) for i in inst]
        nd = sum(c["scene"]["xyzr"].shape[0] for c in cases)
        desc = (f"64 disc-cloud trenches/holes (AR 2-20, {nd} discs total, s log-uniform 1e-3..1, diffuse or "
                f"specular), {int(r)} rays total proportional to area (1e6-5e7 per instance)")
        return cases, desc, 32.0
    if name == "adjoint":
        r = 1e8 if rays is None else rays
        sc = geometry.trench_discs(width=1.0, depth=8.0, flat=1.0, ylen=1.0, spacing=0.01)
        cases = [dict(scene=sc, n_rays=int(r), sticking=0.05, cos_exponent=5.0, bc_x=0, bc_y=1, src_z=8.1, seed=4)]
        desc = (f"disc-cloud trench w=1 d=8 (AR 8), 0.01 grid, {sc['xyzr'].shape[0]} discs, reflective x, periodic y, "
                f"source z=8.1, n=5, s=0.05, {int(r)} rays, adjoint d/ds and d/dn in the same pass")
        return cases, desc, 32.0
    if name == "rough":
        r = 1e9 if rays is None else rays
        sc = geometry.rough_surface_discs()
        cases = [dict(scene=sc, n_rays=int(r), sticking=0.02, diffuse_prob=0.7, bc_x=1, bc_y=1, seed=5)]
        desc = (f"rough periodic surface (4096^2 height field, 256 bumps) over 16 holes of AR 6, "
                f"{sc['xyzr'].shape[0]} discs, 70/30 diffuse/specular, s=0.02, {int(r)} rays")
        return cases, desc, 32.0
    raise SystemExit(f"unknown workload {name}")


def bytes_per_segment(counters):
    """Algorithmic bytes per segment from an instrumented run, with the
    per-unit weights of the measurement plan:
        B_seg = 32 V_node + 32 V_prim + 24 D + 32 K_nb + 4 K_nb + 64
    V_node: tree nodes visited, V_prim: primitives tested, D: deposits,
    K_nb: neighbour discs tested (disc record + column index), 64: the ray
    state read and written once per segment. The traversal here visits
    four-wide nodes; each visit counts as one node at 32 B (one visit tests
    four children's boxes, stored quantised in a 64-byte record)."""
    seg = float(counters[0])
    nodes, prims, nbt, deps = (float(x) for x in counters[4:8])
    parts = {"V_node": nodes / seg, "V_prim": prims / seg, "K_nb": nbt / seg, "D": deps / seg}
    b = 32.0 * parts["V_node"] + 32.0 * parts["V_prim"] + 24.0 * parts["D"] + 36.0 * parts["K_nb"] + 64.0
    return b, parts


def record_bytes_per_segment(parts, prim_bytes, deposit_bytes, node_bytes):
    """The same visits priced at the sizes of this layout's records:
    node_bytes per four-wide node visit (the 48 bytes a visit loads; a disc
    cloud's marked nodes load their 16-byte leaf word too, priced here at 64),
    prim_bytes per primitive test (48 B triangle, 32 B disc), 36 B per
    neighbour test, deposit_bytes per deposit (read + write of the 8-byte
    sums: three, five with the adjoint), no ray-state traffic (rays stay in
    registers)."""
    return (node_bytes * parts["V_node"] + prim_bytes * parts["V_prim"] + 36.0 * parts["K_nb"]
            + deposit_bytes * parts["D"])


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cases, budget_s=7.0, runs=3, adjoint=False):
    """Oracle (restated reference tracer, OpenMP on every host core) on a
    bounded ray sample of the same cases, sized by a 2e4-ray pilot to about
    budget_s seconds per run; the trace region only (the reference times
    tracer.hpp:122-286, without the BVH build); the median of ``runs`` runs
    after the pilot warmed up."""
    import oracle
    from rti_amd import geometry
    threads = os.cpu_count() or 1
    if len(cases) > 8:
        cases = cases[:: max(1, len(cases) // 8)]
    boxes = [geometry.scene_box(c["scene"], c.get("src_z")) for c in cases]

    def run(per_case):
        segs, secs = 0, 0.0
        for c, (box, sz) in zip(cases, boxes):
            r = oracle.trace(c["scene"], n_rays=per_case, box=box, src_z=sz, sticking=c["sticking"],
                             cos_exponent=c.get("cos_exponent", 1.0), diffuse_prob=c.get("diffuse_prob", 1.0),
                             bc_x=c["bc_x"], bc_y=c["bc_y"], seed=c["seed"], threads=threads, adjoint=adjoint)
            segs += r["segments"]
            secs += r["seconds"]
        return segs, secs
    pilot = 20000
    segs, secs = run(pilot)
    per_case = int(pilot * min(2000.0, max(1.0, budget_s / max(secs, 1e-3))))
    rates = []
    for _ in range(runs):
        segs, secs = run(per_case)
        rates.append(segs / secs)
    value = float(np.median(rates))
    return {"value": value, "unit": "segments/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "runs": [round(r, 1) for r in rates],
            "sample": f"oracle (restated reference tracer, own BVH, OpenMP, {threads} threads) on {per_case} rays "
                      f"for each of {len(cases)} cases of the same geometry, {segs} segments per run; median of "
                      f"{runs} runs"}


def radiosity_report(cases, knobs, shards=20):
    """Hole workload: the mean incident flux on 20 wall bands (depth 0.25
    each) of every (s, n) case, against the FP64 radiosity solution of the
    same hole integrated over the same wall triangles (committed in
    tests/golden/radiosity_hole_bands.json). Incident flux of a triangle =
    raw / (s * area) * source area / rays. The cases' rays are traced as
    ``shards`` consecutive ray ranges of equal size (the same rays as one
    run; integer sums add exactly), and a band's standard error is the spread
    of its shard means: a history that deposits several times in one band
    (most at small s) makes the per-deposit variance too small. Bound per
    band: 3 standard errors + 0.2 % (the 320-gon facets against the
    cylinder move the flux by about 1e-4; the solution's discretisation
    by less)."""
    from rti_amd import Context
    with open(os.path.join(ROOT, "tests", "golden", "radiosity_hole_bands.json")) as fh:
        ref = json.load(fh)
    sc = cases[0]["scene"]
    v = sc["verts"][sc["tris"]].astype(np.float64)
    c = v.mean(1)
    area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    top = float(sc["verts"][:, 2].max())
    depth = top - c[:, 2]
    wall = (depth > 1e-6) & (depth < sc["depth"] - 1e-6) & (np.abs(np.hypot(c[:, 0], c[:, 1]) - sc["radius"]) < 0.01)
    lo, hi = sc["verts"].min(0), sc["verts"].max(0)
    a_src = float((hi[0] - lo[0]) * (hi[1] - lo[1]))
    edges = ref["band_edges_depth"]
    nb = len(edges) - 1
    masks = [wall & (depth >= edges[b]) & (depth < edges[b + 1]) for b in range(nb)]
    n = cases[0]["n_rays"]
    assert all(cs["n_rays"] == n for cs in cases)
    per = n // shards
    est = np.zeros((shards, len(cases), nb))
    with Context(0) as ctx:
        for k in range(shards):
            ctx.batch_set([dict(cs, n_rays=per, ray_begin=k * per, rays_total=per * shards) for cs in cases])
            ctx.set_params(**knobs)
            ctx.batch_run()
            for i, cs in enumerate(cases):
                g = ctx.batch_flux(i)["raw"] / (cs["sticking"] * area) * a_src / per
                for b in range(nb):
                    w = area[masks[b]]
                    est[k, i, b] = np.sum(g[masks[b]] * w) / w.sum()
    out, ok = [], True
    for i, cs in enumerate(cases):
        s, n_exp = cs["sticking"], cs.get("cos_exponent", 1.0)
        want = np.array(next(r["wall_band_flux"] for r in ref["cases"] if r["sticking"] == s
                             and r["cos_exponent"] == n_exp))
        m = est[:, i, :].mean(0)
        se = est[:, i, :].std(0, ddof=1) / math.sqrt(shards)
        err = np.abs(m - want)
        within = bool(np.all(err < 3 * se + 0.002 * want))
        ok = ok and within
        out.append({"s": s, "n": n_exp, "max_rel_err": round(float((err / want).max()), 5),
                    "mean_rel_err": round(float((err / want).mean()), 5),
                    "max_err_over_sigma": round(float((err / se).max()), 2),
                    "within_3sigma_plus_0.2pct": within})
    return {"reference": "FP64 radiosity of the same hole over the same wall triangles "
                         "(tests/golden/radiosity_hole_bands.json)",
            "bands": f"20 wall bands, depth 0.25; standard errors from {shards} ray-range shards of {per} rays",
            "cases": out, "all_within": ok}


def adjoint_fd_check(case, knobs, ds=1e-3, dn=1e-2, seeds=8):
    """Summed d(raw)/ds and d(raw)/dn of the adjoint pass against central
    differences of the summed raw flux with common random numbers, repeated
    over independent seeds: means, standard errors, and the difference of the
    means in units of its standard error."""
    from rti_amd import Context
    base = {k: v for k, v in case.items() if k not in ("scene", "seed")}
    s0, n0 = base["sticking"], base["cos_exponent"]
    rows = []
    with Context(0) as ctx:
        ctx.set_scene(case["scene"])

        def total(seed, **over):
            ctx.set_params(**dict(base, **over), seed=seed, **knobs)
            ctx.run()
            return float(ctx.flux()["raw"].sum())
        for k in range(seeds):
            seed = case["seed"] + 100 + k
            ctx.set_params(**base, seed=seed, adjoint=1, **knobs)
            ctx.run()
            a_s, a_n = (float(x.sum()) for x in ctx.adjoint())
            f_s = (total(seed, sticking=s0 + ds) - total(seed, sticking=s0 - ds)) / (2 * ds)
            f_n = (total(seed, cos_exponent=n0 + dn) - total(seed, cos_exponent=n0 - dn)) / (2 * dn)
            rows.append((a_s, f_s, a_n, f_n))
    r = np.array(rows)
    out = {"seeds": seeds, "rays_per_run": base["n_rays"], "steps": {"ds": ds, "dn": dn}}
    for name, ia, ifd in (("d_ds", 0, 1), ("d_dn", 2, 3)):
        ma, mf = r[:, ia].mean(), r[:, ifd].mean()
        sa, sf = r[:, ia].std(ddof=1) / np.sqrt(seeds), r[:, ifd].std(ddof=1) / np.sqrt(seeds)
        out[name] = {"adjoint": ma, "adjoint_se": sa, "fd": mf, "fd_se": sf, "rel_diff": (ma - mf) / abs(mf),
                     "z": (ma - mf) / float(np.hypot(sa, sf))}
    return out


def profile_figure(name, key):
    """A figure of the committed counter summary of the trace kernel on this
    workload: the number after "key:" at the start of a line, or None without
    one. Keys: "bytes_per_segment" (HBM-side traffic from FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for a copy kernel of
    known size), "l2_hit_rate" (TCC_HIT / (TCC_HIT + TCC_MISS)), "valu_busy"
    (rocprof's VALUBusy: SQ_ACTIVE_INST_VALU / CU_NUM / GRBM_GUI_ACTIVE of
    one XCC; above 1 when two VALU instructions issue in one quad-cycle),
    "valu_dual_issue" (SQ_ACTIVE_INST_VALU2 / SQ_ACTIVE_INST_VALU),
    "active_lanes" (SQ_THREAD_CYCLES_VALU / SQ_ACTIVE_INST_VALU, of 64),
    "td_busy" / "ta_busy" (TD_TD_BUSY / TA_TA_BUSY summed over the CUs,
    over CU_NUM x GRBM_GUI_ACTIVE of one XCC: the texture data and address
    units of the vector-memory path)."""
    path = os.path.join(ROOT, "profiles", f"trace_persistent_{name}_pmc.md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith(key + ":"):
                return float(line.split(":")[1].split()[0])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="hole", choices=["hole", "trench", "batch", "rough", "adjoint"])
    ap.add_argument("--rays", type=float, default=None, help="primary rays (per case for hole, total otherwise)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the trace shards by ray-id range with one int64 all-reduce "
              "(design only); run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from rti_amd import Context, tune
    cases, desc, prim_bytes = workload(args.workload, args.rays)
    log(f"{args.workload}: {desc}")
    knobs = {}
    if not args.no_tune:
        t = tune.tune(cases, log=log)
        knobs = tune.knobs(t)
        log(f"tuned {knobs}")
    adj = {"adjoint": 1} if args.workload == "adjoint" else {}
    ctx = Context(0)
    ctx.batch_set(cases)
    ctx.set_params(**adj, **knobs)
    t0 = time.time()
    ctx.prepare()
    log(f"build {time.time() - t0:.2f} s")

    # warmup + timed replays of the captured graph (HIP events on the trace stream)
    if args.warmup:
        ctx.batch_replay(args.warmup)
    t0 = time.perf_counter()
    ctx.batch_replay(args.steps)
    wall = time.perf_counter() - t0
    st = ctx.batch_stats()
    ms_per_step = st["trace_ms"]
    counters = ctx.counters()
    segments = int(counters[0])
    rays_per_step = int(sum(c["n_rays"] for c in cases))
    log(f"persistent: {ms_per_step:.1f} ms/step (wall {1e3 * wall / args.steps:.1f}), {segments} segments/step")
    value = segments / (ms_per_step / 1e3)

    extra = {}
    if args.workload == "hole":
        # the same rays, in shards, against the FP64 radiosity solution
        extra["flux_vs_radiosity"] = radiosity_report(cases, knobs)
        log(f"flux vs radiosity: {extra['flux_vs_radiosity']}")
    if args.workload == "rough":
        # same paths, FP64 atomic accumulation: the fixed point's error
        ctx.batch_run()
        fixed = ctx.batch_flux(0)["raw"]
        shift = ctx.batch_stats()["fixed_shift"]
        ctx.set_params(accum=1, **knobs)
        ctx.batch_run()
        fp64 = ctx.batch_flux(0)["raw"]
        nz = fp64 > 0
        rel = np.abs(fixed[nz] - fp64[nz]) / fp64[nz]
        extra["fixed_vs_fp64"] = {"max_rel": float(rel.max()), "mean_rel": float(rel.mean()),
                                  "max_abs": float(np.abs(fixed - fp64).max()), "fixed_shift": shift}
        log(f"fixed vs fp64: max rel {rel.max():.3e}, mean rel {rel.mean():.3e}")
        ctx.set_params(**knobs)

    if args.workload == "adjoint":
        extra["adjoint_vs_fd"] = adjoint_fd_check(cases[0], knobs)
        log(f"adjoint vs FD: {extra['adjoint_vs_fd']}")

    plain = None
    if not args.no_plain:
        ctx.set_params(kernel=1, **adj, **knobs)
        ctx.batch_replay(1)
        pst = ctx.batch_stats()
        pc = ctx.counters()
        plain = {"segments_per_s": int(pc[0]) / (pst["trace_ms"] / 1e3), "ms_per_step": pst["trace_ms"]}
        log(f"plain: {pst['trace_ms']:.1f} ms/step")
    ctx.close()

    # instrumented run for the algorithmic bytes, on 1/20 of the rays
    cnt_ctx = Context(0)
    small = [dict(c, n_rays=max(1, int(c["n_rays"] // 20))) for c in cases]
    cnt_ctx.batch_set(small)
    cnt_ctx.set_params(kernel=2, **knobs)
    cnt_ctx.batch_run()
    B, parts = bytes_per_segment(cnt_ctx.counters())
    cnt_ctx.close()
    B_rec = record_bytes_per_segment(parts, prim_bytes, 80.0 if adj else 48.0, 48.0 if prim_bytes == 48.0 else 64.0)
    achieved = value * B / 1e9
    b_meas = profile_figure(args.workload, "bytes_per_segment")
    traffic = None if b_meas is None else b_meas * segments
    traffic_gbs = None if traffic is None else traffic / (ms_per_step / 1e3) / 1e9
    valu_busy = profile_figure(args.workload, "valu_busy")
    dual = profile_figure(args.workload, "valu_dual_issue")
    lanes = profile_figure(args.workload, "active_lanes")
    td_busy = profile_figure(args.workload, "td_busy")
    ta_busy = profile_figure(args.workload, "ta_busy")
    cu_busy = profile_figure(args.workload, "cu_busy")
    # the resource that binds, from the committed counters of this workload:
    # HBM when the traffic reaches 0.7 of the peak; below that the vector-
    # memory path (texture data unit busy >= 0.85 of the time the CUs are
    # busy: divergent loads, about one cycle per lane per load instruction);
    # else VALU issue when the SIMDs are nearly always issuing VALU work
    # (VALUBusy >= 0.85); otherwise neither is saturated: latency and the
    # launch's tail. cu_busy (SQ_BUSY_CU_CYCLES / CU_NUM / GRBM_GUI_ACTIVE of
    # one XCC) is the part of the launch before its drain; 1 where the
    # summary does not give it.
    binding = None
    if valu_busy is not None and traffic_gbs is not None:
        if traffic_gbs >= 0.7 * HBM_PEAK_GBS:
            binding = "hbm"
        elif td_busy is not None and td_busy >= 0.85 * (cu_busy if cu_busy is not None else 1.0):
            binding = "vmem"
        elif valu_busy >= 0.85:
            binding = "valu"
        else:
            binding = "latency"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_gbs": None if traffic_gbs is None else round(traffic_gbs, 1),
                "traffic_frac": None if traffic_gbs is None else round(traffic_gbs / HBM_PEAK_GBS, 4),
                "l2_hit_rate": profile_figure(args.workload, "l2_hit_rate"),
                "bytes_per_segment": round(B, 1), **{k: round(v, 3) for k, v in parts.items()},
                "bytes_per_segment_records": round(B_rec, 1),
                "binding": binding, "td_busy": td_busy, "ta_busy": ta_busy, "cu_busy": cu_busy,
                "valu_busy": valu_busy,
                "valu_dual_issue": dual, "active_lanes": lanes,
                "lane_frac": None if lanes is None else round(lanes / 64, 3),
                "note": "achieved and frac: B_seg = 32 V_node + 32 V_prim + 24 D + 36 K_nb + 64 bytes per segment "
                        "(V_node four-wide node visits) x segments/s over the HBM peak; bytes_per_segment_records "
                        "prices the same visits at this layout's record sizes. traffic, traffic_gbs, l2_hit_rate, "
                        "valu_busy (rocprof VALUBusy; above 1 with dual VALU issue), valu_dual_issue and "
                        "active_lanes (of 64, per issued VALU instruction), td_busy and ta_busy (texture data and "
                        "address units), cu_busy (CUs busy over the launch; the rest is its drain) come from the "
                        "committed counter summary of the trace kernel on this workload (null without one); "
                        "binding names the resource those counters show binding while the CUs are busy (hbm, "
                        "vmem: the vector-memory path's data unit, valu, or latency when none is saturated)"}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cases, adjoint=bool(adj))
        log(f"cpu baseline {cpu['value']:.3e} segments/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
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
        "config": {"workload": desc + "; one persistent launch per step, HIP graph",
                   "rays_per_step": rays_per_step, "segments_per_step": segments,
                   "longest_path_segments": int(st["max_segments"]), "knobs": knobs},
        "plain": plain,
        "batched_vs_plain": (value / plain["segments_per_s"]) if plain else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
