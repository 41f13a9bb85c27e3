"""Autotuner for the persistent trace kernel's launch knobs.

For a batch (the same item dicts ``Context.batch_set`` takes) it traces a
reduced ray budget with each candidate and keeps the fastest, by coordinate
descent (block size and leaf threshold visited twice): ``compact_lanes`` (lanes of a wave that must have finished their
closest-hit query before the wave shades and refills), then the resident
waves per CU (0 = as many as registers and LDS allow; fewer can suit
workloads whose working set outgrows the caches), then the LDS stack depth,
then the block size, then ``leaf_batch`` (-1: leaf tests at each node visit;
32..256: postponed leaf tests on the wave's shared leaf list, a leaf phase
once the entries it takes reach that many 64ths of the lanes with a node
left), then ``refill_lanes``
(lanes of a wave that must wait for a new ray before the wave refills); for
workloads of disc clouds only, the binary-tree topology (``tree_builder``,
LBVH or PLOC) first. When
the candidates were timed on a sample, the few that the sample ranks within
1 % of the best, and for disc clouds the other tree topology, are timed again
on the full workload, which decides. Results
are cached in a JSON file keyed by the batch signature, so a repeated
workload is not re-tuned.
"""
import hashlib
import json
import os

import numpy as np

from ._lib import Context, device_info

COMPACT = (40, 48, 52, 56, 60, 64)
LDS = (4, 6, 8, 16)
BLOCK = (128, 256)
LEAF = (-1, 64, 128, 256)
REFILL = (1, 8, 16)
WAVES = (0, 20, 24)
# binary-tree topology (RTI_TREE_LBVH, RTI_TREE_PLOC): tuned for workloads of
# disc clouds only, whose results do not depend on it (a triangle mesh's
# stored order follows its tree, so the tree decides exact ties at shared
# edges, and a mesh keeps the default)
TREE = (1, 2)
MIN_GAIN = 0.005  # relative gain a coordinate step must show to change a knob
FINAL_BAND = 0.01  # sample-rate band of the candidates re-timed on the full workload
FINALISTS = 4


def signature(items, device=None):
    """Cache key of a workload on a device: every item's geometry (a digest
    of its arrays), ray budget, physics, boundaries and source, and the
    device's architecture and compute-unit count."""
    h = hashlib.sha1()
    if device is None:
        device = device_info(0)
    h.update(f"{device[0]}:{device[1]}".encode())
    digests = {}
    for it in items:
        sc = it["scene"]
        if id(sc) not in digests:
            g = hashlib.sha1()
            for key in (("xyzr", "normals") if sc["kind"] == "discs" else ("verts", "tris")):
                g.update(np.ascontiguousarray(sc[key]).tobytes())
            digests[id(sc)] = g.hexdigest()
        h.update(f"{sc['kind']}:{digests[id(sc)]}:{it.get('n_rays')}:{it.get('sticking')}:"
                 f"{it.get('cos_exponent', 1.0)}:{it.get('diffuse_prob', 1.0)}:{it.get('bc_x', 0)}:"
                 f"{it.get('bc_y', 0)}:{it.get('source', 0)}:{it.get('src_z')}".encode())
    return h.hexdigest()[:16]


def _rate(ctx, knobs, timed=2):
    ctx.set_params(**knobs)
    ctx.batch_replay(1)          # capture + warm
    ctx.batch_replay(timed)      # mean of the timed replays (run-to-run spread ~2 %)
    st = ctx.batch_stats()
    return float(ctx.counters()[0]) / (st["trace_ms"] / 1e3)


def _finalists(passes, best):
    """The descent's result, the other tree topologies (their effect on the
    launch's longest paths does not show on a sample: on the 64-instance disc
    batch PLOC was 2 % slower on the sample and 3-5 % faster at full size),
    and, fastest on the sample first, single-knob variants: each knob set to
    a value that its last coordinate pass ranked within FINAL_BAND of that
    pass's best; at most FINALISTS in all."""
    out = [dict(best)]
    for v in passes.get("tree_builder", {}):
        if v != best["tree_builder"]:
            out.append(dict(best, tree_builder=v))
    cands = []
    for name, scores in passes.items():
        if name == "tree_builder":
            continue
        top = max(scores.values())
        for v, r in scores.items():
            if v != best[name] and r >= top * (1.0 - FINAL_BAND):
                cands.append((r, name, v))
    cands.sort(key=lambda c: -c[0])
    for r, name, v in cands:
        if len(out) >= FINALISTS:
            break
        out.append(dict(best, **{name: v}))
    return out


def default_fraction(items, sample_rays=2e8, full_up_to=2e8):
    """Fraction of the rays the candidates are timed on: all of them up to
    ``full_up_to`` rays (a small sample of a batch is dominated by its single
    longest path, which does not shrink with the sample), else about
    ``sample_rays``. Smaller samples rank close candidates differently from
    the full launch (on the 6e8-ray hole a 5e7-ray sample preferred a
    configuration 1.5 % slower at full size)."""
    total = float(sum(it["n_rays"] for it in items))
    return 1.0 if total <= full_up_to else sample_rays / total


def tune(items, fraction=None, min_rays=20000, cache=None, log=None):
    """Best knobs {compact_lanes, waves_per_cu, lds_stack, block_size, leaf_batch, refill_lanes, and
    tree_builder for disc clouds} and their rate."""
    key = signature(items) if cache else None
    if cache and os.path.exists(cache):
        with open(cache) as fh:
            db = json.load(fh)
        if key in db:
            return db[key]
    if fraction is None:
        fraction = default_fraction(items)
    small = [dict(it, n_rays=max(min_rays, int(it["n_rays"] * fraction))) for it in items]
    best = {"compact_lanes": 56, "waves_per_cu": 0, "lds_stack": 8, "block_size": 256, "leaf_batch": 0,
            "refill_lanes": 1}
    discs = all(it["scene"]["kind"] == "discs" for it in items)
    order = [("compact_lanes", COMPACT), ("waves_per_cu", WAVES), ("lds_stack", LDS), ("block_size", BLOCK),
             ("leaf_batch", LEAF), ("refill_lanes", REFILL), ("block_size", BLOCK), ("leaf_batch", LEAF)]
    if discs:
        best["tree_builder"] = 1  # the default for disc clouds, moved only for a clear gain
        order.insert(0, ("tree_builder", TREE))
    trials = []
    passes = {}  # knob -> {value: sample rate} of its last coordinate pass
    with Context() as ctx:
        ctx.batch_set(small)
        ctx.prepare()
        # one pass over all knobs, then the block size and the leaf-phase
        # threshold once more (the two interact most)
        for name, values in order:
            scores = {}
            for v in values:
                knobs = dict(best, **{name: v})
                scores[v] = _rate(ctx, knobs)
                trials.append({**knobs, "segments_per_s": scores[v]})
                if log:
                    log(f"tune {knobs} -> {scores[v]:.3e} segments/s")
            # move only for a clear gain: candidates within the replay-to-replay
            # spread would otherwise steer the descent by noise (waves_per_cu
            # 24 at +0.1 % once hid lds_stack 4, +1.8 % with waves_per_cu 0)
            passes[name] = scores
            top = max(scores, key=scores.get)
            cur = scores.get(best[name])  # None: the starting value is not a candidate
            if cur is None or scores[top] > cur * (1.0 + MIN_GAIN):
                best[name] = top
        # the rate of the knobs returned (the descent can keep a starting
        # value that another trial beat by less than MIN_GAIN)
        rate = _rate(ctx, best)
    full_trials = []
    if fraction < 1.0:
        # The sample does not reproduce the full launch's drain (paths of
        # thousands of segments end it), so the descent's result, the other
        # tree topology (disc clouds) and its single-knob variants that the
        # sample ranks within FINAL_BAND of it are timed once more on the
        # full workload, and the fastest there is kept: on the 64-instance disc batch, leaf tests at each node visit
        # and leaf_batch 128 were 0.2 % apart on a 2e8-ray sample and 6 %
        # apart at its full 6.4e8 rays.
        finalists = _finalists(passes, best)
        if len(finalists) > 1:
            with Context() as ctx:
                ctx.batch_set(items)
                ctx.prepare()
                for cand in finalists:
                    r = _rate(ctx, cand, timed=1)
                    full_trials.append({**cand, "segments_per_s": r})
                    if log:
                        log(f"tune full {cand} -> {r:.3e} segments/s")
            pick = max(full_trials, key=lambda t: t["segments_per_s"])
            best = {k: pick[k] for k in best}
            rate = pick["segments_per_s"]
    out = dict(best, segments_per_s=rate, trials=trials, full_trials=full_trials, signature=key)
    if cache:
        db = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                db = json.load(fh)
        db[key] = out
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as fh:
            json.dump(db, fh, indent=1)
    return out


def knobs(tuned):
    return {k: tuned[k] for k in ("compact_lanes", "waves_per_cu", "lds_stack", "block_size", "leaf_batch",
                                  "refill_lanes", "tree_builder") if k in tuned}
