"""GPU tests at the shapes of the benchmark configurations: the default
bench workload's hole at full size traced under different schedules and
trees, and the configurations that are not the default bench workload:

* the 64-instance disc batch (trenches and holes of aspect 2-20, 1e4-5e5
  discs, sticking 1e-3..1, diffuse or mirror) at its own ray budgets (6.4e8
  rays in total): the graph replay equals a direct launch, every instance
  equals its own single run bit for bit, and instances agree with the oracle
  within 3 sigma;
* the aspect-8 disc trench of the differentiation benchmark (192,001 discs,
  s = 0.05, cos^5 source, 1e8 rays): summed d/ds and d/dn of the adjoint
  against common-random-number central differences (ds = 1e-3, dn = 1e-2)
  over independent seeds;
* the 4.2M-disc rough surface at 1e9 rays: 64-bit fixed point against FP64
  atomic accumulation of the same paths, and the flux against the oracle
  within 3 sigma on a 2e7-ray sample.

Elements compared, and the oracle's standard error: from a third GPU run
with another seed (see _z_against_oracle).
"""
import numpy as np
import pytest

from rti_amd import Context, geometry, run_scene
from tests.helpers import check_z, oracle_run

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def _z_against_oracle(g_raw, g_sigma, n_gpu, o, n_orc, sel, n_sel, min_expected=30):
    """Per-ray means of the GPU run and the oracle run, as z-scores.

    A third, independent GPU run ``sel`` (n_sel rays, another seed) chooses
    the elements (those where the oracle can expect min_expected deposits)
    and gives the per-ray variance for the oracle's standard error. Taking
    either from a compared run itself biases z at a few tens of deposits per
    element: selecting on a run's own counts favours its upward
    fluctuations, and a run's own variance estimate is small exactly when
    its sum fluctuates low."""
    m = sel["counts"] * (n_orc / n_sel) >= min_expected
    var_ray = sel["sigma"][m] ** 2 / n_sel
    a, sa = g_raw[m] / n_gpu, g_sigma[m] / n_gpu
    b, sb = o["s1"][m] / n_orc, np.sqrt(var_ray / n_orc)
    s = np.sqrt(sa ** 2 + sb ** 2)
    ok = s > 0
    return (a[ok] - b[ok]) / s[ok]


def _batch_cases():
    inst = geometry.batch_instances(seed=3)
    return inst, [dict(scene=i["scene"], n_rays=i["n_rays"], sticking=i["sticking"], diffuse_prob=i["diffuse_prob"],
                       bc_x=1, bc_y=1, seed=3) for i in inst]


def test_c3_batch_at_own_budgets():
    inst, cases = _batch_cases()
    total = sum(c["n_rays"] for c in cases)
    nd = sum(c["scene"]["xyzr"].shape[0] for c in cases)
    assert len(cases) == 64 and 6.0e8 < total < 6.6e8
    _say(f"c3 batch: 64 instances, {nd} discs, {total} rays, graph replay and direct launch")
    with Context() as ctx:
        ctx.batch_set(cases)
        ctx.batch_replay(1)
        ctx.finalize()
        replay = [(ctx.batch_flux(i)["raw"], ctx.batch_counts(i)) for i in range(64)]
        seg_replay = int(ctx.counters()[0])
        ctx.batch_run()
        ran = [(ctx.batch_flux(i), ctx.batch_counts(i)) for i in range(64)]
        seg_run = int(ctx.counters()[0])
    assert seg_replay == seg_run
    for k in range(64):
        assert np.array_equal(replay[k][0], ran[k][0]["raw"]), k
        assert np.array_equal(replay[k][1], ran[k][1]), k
    # every instance equals its own run (instance k: Philox key word 1 = k)
    segs = 0
    for k, c in enumerate(cases):
        kw = {key: v for key, v in c.items() if key != "scene"}
        single = run_scene(c["scene"], instance=k, **kw)
        assert np.array_equal(single["raw"], ran[k][0]["raw"]), k
        assert np.array_equal(single["counts"], ran[k][1]), k
        segs += int(single["counters"][0])
    assert segs == seg_run
    _say(f"c3: all 64 instances equal their single runs; {seg_run} segments")
    # three instances against the oracle (independent seeds): a mirror one,
    # a diffuse trench and a diffuse hole, among those with paths short
    # enough for the CPU (s >= 0.05)
    order = sorted(range(64), key=lambda k: cases[k]["scene"]["xyzr"].shape[0])
    picks = []
    for want in (lambda k: cases[k]["diffuse_prob"] == 0.0, lambda k: inst[k]["kind"] == "trench",
                 lambda k: inst[k]["kind"] == "hole"):
        k = next(k for k in order if cases[k]["sticking"] >= 0.05 and want(k) and k not in picks)
        picks.append(k)
    for k in picks:
        c = cases[k]
        n_orc = min(c["n_rays"], 2_000_000)
        o = oracle_run(c["scene"], n_orc, c["sticking"], seed=99 + k, diffuse_prob=c["diffuse_prob"], bc_x=1, bc_y=1)
        f = ran[k][0]
        kw = {key: v for key, v in c.items() if key not in ("scene", "seed", "n_rays")}
        n_sel = 4 * n_orc
        sel = run_scene(c["scene"], n_rays=n_sel, seed=7000 + k, **kw)
        z = _z_against_oracle(f["raw"], f["sigma"], c["n_rays"], o, n_orc, sel, n_sel)
        _say(f"c3 instance {k} ({inst[k]['kind']}, AR {inst[k]['aspect']:.1f}, {c['scene']['xyzr'].shape[0]} discs, "
             f"s={c['sticking']:.3g}, diffuse {c['diffuse_prob']}): {z.size} elements, z stats {check_z(z)}")


def test_c3_batch_independent_of_leaf_scheduling_and_tree():
    """Coplanar discs tie at exactly equal distances; the lower stored
    position wins only if culling never drops a tied partner. At the batch's
    full size (6.4e8 rays, 8.4e9 segments) every instance is bit-identical
    across leaf tests at each node visit, the wave-shared leaf list under two
    schedules, and the binary tree."""
    _, cases = _batch_cases()
    _say("c3 batch: leaf scheduling and tree width, 6.4e8 rays each")
    runs = []
    for knobs in (dict(leaf_batch=-1), dict(leaf_batch=64, compact_lanes=56, refill_lanes=1),
                  dict(leaf_batch=64, compact_lanes=52, refill_lanes=8), dict(bvh_width=2)):
        with Context() as ctx:
            ctx.batch_set(cases)
            ctx.set_params(**knobs)
            ctx.batch_run()
            runs.append(([ctx.batch_flux(i)["raw"] for i in range(64)], [ctx.batch_counts(i) for i in range(64)],
                         int(ctx.counters()[0])))
        _say(f"c3 {knobs}: {runs[-1][2]} segments")
    for r in runs[1:]:
        assert r[2] == runs[0][2]
        for k in range(64):
            assert np.array_equal(r[0][k], runs[0][0][k]), k
            assert np.array_equal(r[1][k], runs[0][1][k]), k


def test_c2_hole_independent_of_leaf_scheduling_and_tree():
    """The benchmark's hole at its full size (six (s, n) cases of 1e8 rays on
    one shared 199,360-triangle mesh): bit-identical sums with the leaf list
    at two thresholds, with leaf tests at each node visit, and on the binary
    tree."""
    sc = geometry.hole_triangles()
    cases = [dict(scene=sc, n_rays=100_000_000, sticking=s, cos_exponent=n, bc_x=1, bc_y=1, seed=2)
             for s in (0.01, 0.1, 1.0) for n in (1.0, 10.0)]
    runs = []
    for knobs in (dict(), dict(leaf_batch=1, compact_lanes=48), dict(leaf_batch=-1), dict(bvh_width=2)):
        with Context() as ctx:
            ctx.batch_set(cases)
            ctx.set_params(**knobs)
            ctx.batch_run()
            runs.append(([ctx.batch_flux(i)["raw"] for i in range(6)], [ctx.batch_counts(i) for i in range(6)],
                         int(ctx.counters()[0])))
        _say(f"c2 hole {knobs}: {runs[-1][2]} segments")
    for r in runs[1:]:
        assert r[2] == runs[0][2]
        for k in range(6):
            assert np.array_equal(r[0][k], runs[0][0][k]), k
            assert np.array_equal(r[1][k], runs[0][1][k]), k


def test_c4_adjoint_vs_crn_finite_differences():
    """Summed d(flux)/ds and d(flux)/dn over the whole trench: adjoint mean
    over 16 seeds against the CRN central differences' mean. Bound: 2 % of
    the difference quotient plus 2 standard errors of the difference (the
    disc cloud's d/dn difference quotient is noisy: a path's deposits jump as
    its hit point crosses a neighbour disc's rim)."""
    sc = geometry.trench_discs(width=1.0, depth=8.0, flat=1.0, ylen=1.0, spacing=0.01)
    assert sc["xyzr"].shape[0] == 192_001
    s0, n0, ds, dn = 0.05, 5.0, 1e-3, 1e-2
    base = dict(n_rays=100_000_000, bc_x=0, bc_y=1, src_z=8.1)
    seeds = 16
    rows = []
    _say(f"c4 trench {sc['xyzr'].shape[0]} discs, 1e8 rays, {seeds} seeds x 5 runs")
    with Context() as ctx:
        ctx.set_scene(sc)

        def total(seed, s, n, adj=0):
            ctx.set_params(**base, sticking=s, cos_exponent=n, seed=seed, adjoint=adj)
            ctx.run()
            return float(ctx.flux()["raw"].sum())
        for k in range(seeds):
            seed = 400 + k
            total(seed, s0, n0, adj=1)
            a_s, a_n = (float(x.sum()) for x in ctx.adjoint())
            f_s = (total(seed, s0 + ds, n0) - total(seed, s0 - ds, n0)) / (2 * ds)
            f_n = (total(seed, s0, n0 + dn) - total(seed, s0, n0 - dn)) / (2 * dn)
            rows.append((a_s, f_s, a_n, f_n))
    r = np.array(rows)
    for name, ia, ifd in (("d/ds", 0, 1), ("d/dn", 2, 3)):
        ma, mf = r[:, ia].mean(), r[:, ifd].mean()
        se = np.std(r[:, ia] - r[:, ifd], ddof=1) / np.sqrt(seeds)
        _say(f"c4 {name}: adjoint {ma:.6g} FD {mf:.6g} rel diff {(ma - mf) / abs(mf):+.4f} "
             f"se of difference {se / abs(mf):.4f}")
        assert abs(ma - mf) <= 0.02 * abs(mf) + 2 * se, name


@pytest.fixture(scope="module")
def rough():
    return geometry.rough_surface_discs()


def test_c5_rough_surface_full_size(rough):
    sc = rough
    n = 1_000_000_000
    kw = dict(n_rays=n, sticking=0.02, diffuse_prob=0.7, bc_x=1, bc_y=1, seed=5)
    _say(f"c5 rough surface {sc['xyzr'].shape[0]} discs, 1e9 rays, fixed point then FP64 atomics")
    fx = run_scene(sc, **kw)
    fp = run_scene(sc, accum=1, **kw)
    assert np.array_equal(fx["counts"], fp["counts"])
    shift = fx["stats"]["fixed_shift"]
    # at most half a grid unit per deposit, plus the FP64 sums' own rounding
    bound = fx["counts"] * 2.0 ** (-shift - 1) + 1e-12 * np.abs(fp["raw"])
    assert np.all(np.abs(fx["raw"] - fp["raw"]) <= bound)
    nz = fp["raw"] > 0
    rel = np.abs(fx["raw"][nz] - fp["raw"][nz]) / fp["raw"][nz]
    m = fx["counts"] >= 30
    rel_s = np.abs(fx["sigma"][m] - fp["sigma"][m]) / fp["sigma"][m]
    _say(f"c5 fixed vs FP64: max rel {rel.max():.3e}, mean {rel.mean():.3e}; sigma max rel {rel_s.max():.3e}")
    assert rel_s.max() < 1e-5
    # against the oracle on a 2e7-ray sample (about 35 deposits per element),
    # independent seed, over the elements expected to get 20 or more
    n_orc = 20_000_000
    o = oracle_run(sc, n_orc, 0.02, seed=8765, diffuse_prob=0.7, bc_x=1, bc_y=1)
    sel = run_scene(sc, **dict(kw, n_rays=200_000_000, seed=6))
    z = _z_against_oracle(fx["raw"], fx["sigma"], n, o, n_orc, sel, 200_000_000, min_expected=20)
    _say(f"c5 vs oracle ({n_orc} rays): {z.size} elements, z stats {check_z(z)}")
