"""The on-GPU LBVH and the neighbour graph built from it.

* every primitive is exactly one leaf; every child box encloses the boxes of
  all primitives below it (so no query can miss a primitive);
* the overlap graph equals the oracle's (which equals the reference's
  divide-and-conquer, see test_oracle.py) -- exact, integer index work.
"""
import numpy as np
import pytest

import oracle
from rti_amd import Context, geometry
from tests.helpers import small_trench

pytestmark = pytest.mark.gpu


def _leaf_boxes(scene, perm):
    x = scene["xyzr"][perm].astype(np.float64)
    n = scene["normals"][perm].astype(np.float64)
    ext = x[:, 3:4] * np.sqrt(np.maximum(0.0, 1.0 - n * n))
    return x[:, :3] - ext, x[:, :3] + ext


def _validate(nodes, perm, scene):
    n = len(perm)
    assert sorted(perm.tolist()) == list(range(n))
    lo, hi = _leaf_boxes(scene, perm)
    child = nodes["child"][:, :2]
    seen = np.zeros(n, dtype=np.int64)
    # iterative post-order: subtree box of every node
    sub_lo = np.full((len(nodes), 3), np.inf)
    sub_hi = np.full((len(nodes), 3), -np.inf)
    order, stack = [], [0]
    while stack:
        c = stack.pop()
        order.append(c)
        for cc in set(child[c].tolist()):
            if cc >= 0:
                stack.append(cc)
            else:
                seen[~cc] += 1
    assert np.all(seen == 1), f"{np.sum(seen != 1)} primitives are not exactly one leaf"
    bad = 0
    for c in reversed(order):
        for slot in range(2):
            cc = child[c][slot]
            if cc < 0:
                blo, bhi = lo[~cc], hi[~cc]
            else:
                blo, bhi = sub_lo[cc], sub_hi[cc]
            xy = nodes["xy0"][c] if slot == 0 else nodes["xy1"][c]
            z = nodes["z01"][c][0:2] if slot == 0 else nodes["z01"][c][2:4]
            nlo = np.array([xy[0], xy[2], z[0]])
            nhi = np.array([xy[1], xy[3], z[1]])
            tol = 1e-6 * (1 + np.abs(blo))
            if np.any(nlo > blo + tol) or np.any(nhi < bhi - tol):
                bad += 1
                if bad < 4:
                    print("node", c, "slot", slot, "child", cc, "box", nlo, nhi, "needs", blo, bhi)
            sub_lo[c] = np.minimum(sub_lo[c], blo)
            sub_hi[c] = np.maximum(sub_hi[c], bhi)
    assert bad == 0, f"{bad} child boxes do not enclose their subtree"


@pytest.mark.parametrize("builder", [1, 2], ids=["lbvh", "ploc"])
@pytest.mark.parametrize("which", ["small_trench", "trench_fine", "hole_discs"])
def test_hierarchy_encloses_every_primitive(which, builder):
    sc = {"small_trench": small_trench(),
          "trench_fine": geometry.trench_discs(spacing=0.02, depth=1.0, flat=0.3, ylen=0.3),
          "hole_discs": geometry.hole_discs(spacing=0.05)}[which]
    print(f"[gpu] hierarchy {which} n={sc['xyzr'].shape[0]} tree_builder {builder}", flush=True)
    with Context() as ctx:
        ctx.set_scene(sc)
        ctx.set_params(n_rays=16, sticking=1.0, tree_builder=builder)
        nodes, perm, pa = ctx.hierarchy()
        np.testing.assert_array_equal(pa, sc["xyzr"][perm])
        _validate(nodes, perm, sc)


@pytest.mark.parametrize("which,builder", [("small_trench", 0), ("trench_c1", 0), ("hole_discs", 0), ("rough", 0),
                                           ("trench_c1", 2), ("hole_discs", 2)])
def test_neighbor_graph_equals_oracle(which, builder):
    sc = {"small_trench": small_trench(),
          "trench_c1": geometry.trench_discs(),
          "hole_discs": geometry.hole_discs(spacing=0.04),
          "rough": geometry.rough_surface_discs(nfield=256, cells=1, spacing=0.02)}[which]
    print(f"[gpu] neighbours {which} n={sc['xyzr'].shape[0]} tree_builder {builder}", flush=True)
    with Context() as ctx:
        ctx.set_scene(sc)
        ctx.set_params(n_rays=16, sticking=1.0, tree_builder=builder)
        row, col = ctx.neighbors()
    orow, ocol = oracle.neighbors(sc["xyzr"])
    bad = np.nonzero(np.diff(row.astype(np.int64)) != np.diff(orow.astype(np.int64)))[0]
    assert len(bad) == 0, f"{len(bad)} rows differ in degree, first {bad[:5]}"
    assert np.array_equal(col, ocol)
