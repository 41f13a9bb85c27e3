"""Accumulation, sharding and shared geometry on the GPU (through the C ABI).

* Shards: a run of rays [0, a) plus a run of [a, N) with the same job size
  (rays_total = N) add up to the run of [0, N) bit for bit (integer sums of
  the same deposits; the reference combines per-thread FP64 copies instead,
  tracer.hpp:110-114, hit_accumulator.hpp:47-80).
* Particle species that share one geometry in a batch give the same bits as
  the same batch with separate copies of the geometry.
* The sum of squared deposits has its own fixed-point scale: at s = 1e-3 the
  per-element sigma and relative error agree with FP64 atomic accumulation of
  the same paths within the grid's rounding bound.
"""
import numpy as np
import pytest

from rti_amd import Context, RtiError, run_scene
from rti_amd._lib import RTI_ERR_INVALID_ARG
from tests.helpers import small_hole_mesh, small_trench

pytestmark = pytest.mark.gpu


def _run(sc, **kw):
    print(f"[gpu] run {kw}", flush=True)
    return run_scene(sc, **kw)


@pytest.mark.parametrize("which", ["trench", "hole_mesh"])
def test_shards_sum_to_the_single_run_bit_for_bit(which):
    sc = small_trench() if which == "trench" else small_hole_mesh()
    n = 3_000_000
    kw = dict(sticking=0.1, bc_x=0 if which == "trench" else 1, bc_y=1, seed=17, rays_total=n)
    full = _run(sc, n_rays=n, **kw)
    cut = [0, 1_000_003, 2_200_000, n]
    parts = [_run(sc, n_rays=b - a, ray_begin=a, **kw) for a, b in zip(cut[:-1], cut[1:])]
    shift = full["stats"]["fixed_shift"]
    assert all(p["stats"]["fixed_shift"] == shift for p in parts)
    # the fixed-point integers of the shards add up to the single run's
    q = lambda r: np.rint(r["raw"] * 2.0 ** shift).astype(np.int64)
    assert np.array_equal(sum(q(p) for p in parts), q(full))
    assert np.array_equal(sum(p["counts"] for p in parts), full["counts"])
    for k in range(4):  # segments, surface hits, misses, truncated
        assert sum(int(p["counters"][k]) for p in parts) == int(full["counters"][k])
    # a shard is not the same as a run of the same size from ray 0
    other = _run(sc, n_rays=cut[2] - cut[1], ray_begin=0, **kw)
    assert not np.array_equal(other["counts"], parts[1]["counts"])


def test_rays_total_below_n_rays_is_rejected():
    with Context() as ctx:
        ctx.set_scene(small_trench())
        with pytest.raises(RtiError) as e:
            ctx.set_params(n_rays=1000, rays_total=999)
        assert e.value.code == RTI_ERR_INVALID_ARG


def _batch(items, knobs=None):
    with Context() as ctx:
        ctx.batch_set(items)
        ctx.set_params(**(knobs or {}))
        ctx.batch_run()
        out = []
        for i in range(len(items)):
            f = ctx.batch_flux(i)
            f["counts"] = ctx.batch_counts(i)
            f["rel"], _ = ctx.batch_error_stats(i, vov=False)
            out.append(f)
        st = ctx.batch_stats()
    return out, st


@pytest.mark.parametrize("which", ["trench", "hole_mesh"])
def test_species_on_one_geometry_equal_separate_copies(which):
    """Three species (sticking, source exponent, reflection) on one geometry
    plus a second geometry in between: the batch that shares the first
    geometry's device copy gives the same bits as the batch that holds three
    copies of it."""
    base = small_trench() if which == "trench" else small_hole_mesh()
    other = small_trench(spacing=0.04, depth=1.0)
    copies = [dict(base, xyzr=base["xyzr"].copy(), normals=base["normals"].copy()) if base["kind"] == "discs"
              else dict(base, verts=base["verts"].copy(), tris=base["tris"].copy()) for _ in range(3)]
    species = [dict(sticking=0.05, cos_exponent=1.0, diffuse_prob=1.0),
               dict(sticking=0.5, cos_exponent=10.0, diffuse_prob=0.3),
               dict(sticking=1.0, cos_exponent=3.0, diffuse_prob=1.0)]
    common = dict(n_rays=300_000, bc_x=1, bc_y=1, seed=5)
    shared_items = [dict(scene=base, **species[0], **common), dict(scene=other, sticking=0.2, **common),
                    dict(scene=base, **species[1], **common), dict(scene=base, **species[2], **common)]
    copy_items = [dict(scene=copies[0], **species[0], **common), dict(scene=other, sticking=0.2, **common),
                  dict(scene=copies[1], **species[1], **common), dict(scene=copies[2], **species[2], **common)]
    a, sa = _batch(shared_items)
    b, sb = _batch(copy_items)
    for x, y in zip(a, b):
        assert np.array_equal(x["raw"], y["raw"])
        assert np.array_equal(x["sigma"], y["sigma"])
        assert np.array_equal(x["counts"], y["counts"])
        assert np.array_equal(x["smoothed"], y["smoothed"])
        assert np.array_equal(x["rel"], y["rel"])
    assert sa["segments"] == sb["segments"]
    # the shared batch holds one tree for the three species
    assert sa["wide_nodes"] < sb["wide_nodes"]
    # the species really differ
    assert not np.array_equal(a[0]["counts"], a[2]["counts"])


def test_shared_geometry_graph_replay_and_adjoint():
    """Shared geometry under the graph replay and with the adjoint sums."""
    base = small_trench()
    items = [dict(scene=base, n_rays=200_000, sticking=s, bc_x=0, bc_y=1, seed=9) for s in (0.05, 0.2)]
    copies = [dict(items[0]), dict(items[1], scene=dict(base, xyzr=base["xyzr"].copy()))]
    res = []
    for its in (items, copies):
        with Context() as ctx:
            ctx.batch_set(its)
            ctx.set_params(adjoint=1)
            ctx.batch_replay(2)
            ctx.finalize()
            replayed = [(ctx.batch_flux(i)["raw"], *ctx.batch_adjoint(i)) for i in range(2)]
            ctx.batch_run()
            ran = [(ctx.batch_flux(i)["raw"], *ctx.batch_adjoint(i)) for i in range(2)]
        for x, y in zip(replayed, ran):
            for u, v in zip(x, y):
                assert np.array_equal(u, v)
        res.append(ran)
    for x, y in zip(*res):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)


def test_sum_of_squares_scale_small_sticking():
    """s = 1e-3: deposits are ~1e-4..1e-3, their squares 1e-8..1e-6. The sum
    of squares has its own grid (2^shift2), so sigma and the relative error
    agree with FP64 atomic accumulation of the same paths: relative
    difference below 1e-5 wherever an element has 30 deposits."""
    sc = small_trench()
    kw = dict(n_rays=2_000_000, sticking=1e-3, bc_x=0, bc_y=1, seed=23)
    fx = _run(sc, **kw)
    fp = _run(sc, accum=1, **kw)
    assert np.array_equal(fx["counts"], fp["counts"])
    st = fx["stats"]
    assert st["fixed_shift2"] > st["fixed_shift"] + 15
    m = fx["counts"] >= 30
    assert m.sum() > 100
    # sums: the rounding bound of the grid, 1/2 unit per deposit
    bound = 0.5 * 2.0 ** -st["fixed_shift"] * fx["counts"]
    assert np.all(np.abs(fx["raw"] - fp["raw"]) <= bound + 1e-300)
    for key in ("sigma", "rel_err"):
        rel = np.abs(fx[key][m] - fp[key][m]) / fp[key][m]
        print(f"[gpu] s=1e-3 {key}: max rel diff {rel.max():.2e}", flush=True)
        assert rel.max() < 1e-5, key


def _caterpillar_x(n=80):
    """Positions growing by 2.5x: the gap to the next primitive exceeds the
    extent of all before it, so a locally ordered clustering (PLOC) merges
    the cluster so far with the next primitive every time and builds a chain
    n levels deep; the Karras radix tree (Morton codes + index) stays within
    64 levels."""
    return (4.0 * 2.5 ** (np.arange(n) - n + 1.0)).astype(np.float64)


def _caterpillar_triangles(n=80):
    x = _caterpillar_x(n)
    w = 0.1 * x
    v = np.zeros((3 * n, 3))
    v[0::3] = np.c_[x, np.full(n, -0.5), np.zeros(n)]
    v[1::3] = np.c_[x + w, np.full(n, -0.5), np.zeros(n)]
    v[2::3] = np.c_[x, np.full(n, 0.5), np.zeros(n)]
    return {"kind": "triangles", "verts": v.astype(np.float32), "tris": np.arange(3 * n, dtype=np.uint32).reshape(n, 3)}


def test_deep_hierarchy_guards():
    """A triangle mesh whose PLOC tree is a chain deeper than 64 levels: the
    binary traversal (64-entry lane stack) refuses it with
    RTI_ERR_INVALID_ARG and a hint, the four-wide traversal (which checks its
    own exact stack bound) traces it, and so does the binary traversal of the
    shallow Karras tree of the same mesh."""
    sc = _caterpillar_triangles()
    with Context() as ctx:
        ctx.set_scene(sc)
        ctx.set_params(n_rays=10_000, sticking=1.0, tree_builder=2)
        ctx.run()
        depth = ctx.stats()["bvh_depth"]
        hits = ctx.counts().sum()
        assert depth > 64, depth
        ctx.set_params(n_rays=10_000, sticking=1.0, tree_builder=2, bvh_width=2)
        with pytest.raises(RtiError) as e:
            ctx.run()
        assert e.value.code == RTI_ERR_INVALID_ARG and "bvh_width = 4" in str(e.value)
        ctx.set_params(n_rays=10_000, sticking=1.0, tree_builder=1, bvh_width=2)
        ctx.run()
        assert ctx.stats()["bvh_depth"] <= 64
        assert ctx.counts().sum() == hits
