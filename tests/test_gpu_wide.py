"""The four-wide traversal tree (quantised child boxes) and its traversal.

* structure: every node reachable once, every primitive exactly one leaf,
  1..4 children per node, internal children first, siblings contiguous; a
  mesh's primitives stored in the tree's depth-first order (each node's
  leaves contiguous from leaf_info), a disc cloud's leaf_info marking the
  internal children with leaves;
* conservativeness: each child's box, decoded exactly as the kernel decodes it
  (org + q * 2^(e - 127) rounded once to float32), encloses the exact bounds
  of every primitive below it -- compared without tolerance;
* the four-wide and the binary traversal return the same closest hits, so
  whole runs agree bit for bit;
* the wide tree visits fewer nodes per segment than the binary one.
"""
import numpy as np
import pytest

from rti_amd import Context, geometry, run_scene
from rti_amd._lib import wide_children, wide_scales
from tests.helpers import small_hole_mesh, small_trench

pytestmark = pytest.mark.gpu


def _prim_bounds(scene, perm):
    if scene["kind"] == "discs":
        x = scene["xyzr"][perm].astype(np.float64)
        n = scene["normals"][perm].astype(np.float64)
        ext = x[:, 3:4] * np.sqrt(np.maximum(0.0, 1.0 - n * n))
        return x[:, :3] - ext, x[:, :3] + ext
    v = scene["verts"].astype(np.float64)[scene["tris"][perm]]
    return v.min(axis=1), v.max(axis=1)


def _decode(node, c):
    """Child c's box planes as the traversal computes them (float32)."""
    lo = np.zeros(3)
    hi = np.zeros(3)
    scales, _ = wide_scales(node)
    for k in range(3):
        s = scales[k]
        assert s == np.float32(np.ldexp(1.0, int(np.frexp(s)[1]) - 1)), "scales are powers of two"
        ql = (int(node["lo"][k]) >> (8 * c)) & 0xFF
        qh = (int(node["hi"][k]) >> (8 * c)) & 0xFF
        # q * s is exact in float32; one rounding for the sum, as fmaf does
        lo[k] = np.float32(node["org"][k]) + np.float32(ql) * s
        hi[k] = np.float32(node["org"][k]) + np.float32(qh) * s
    return lo, hi


def _validate_wide(wide, perm, scene):
    n = len(perm)
    plo, phi = _prim_bounds(scene, perm)
    seen_prim = np.zeros(n, dtype=np.int64)
    seen_node = np.zeros(len(wide), dtype=np.int64)
    order, stack = [], [0]
    leaf_order = []  # leaves in depth-first order, each node's leaves before its subtrees
    mesh = scene["kind"] != "discs"
    while stack:
        v = stack.pop()
        seen_node[v] += 1
        order.append(v)
        _, m = wide_scales(wide[v])
        for c in range(m, 4):
            # empty slot: inverted box
            assert all((int(wide["lo"][v][k]) >> (8 * c)) & 0xFF == 0xFF for k in range(3))
            assert all((int(wide["hi"][v][k]) >> (8 * c)) & 0xFF == 0 for k in range(3))
        kids = wide_children(wide[v])
        internal = [ch for ch in kids if ch >= 0]
        leaves = [~ch for ch in kids if ch < 0]
        # internal children first, then the leaves
        assert kids == internal + [~p for p in leaves]
        for p in leaves:
            seen_prim[p] += 1
        leaf_order.extend(leaves)
        if mesh and leaves:
            assert leaves == list(range(int(wide["leaf_info"][v]), int(wide["leaf_info"][v]) + len(leaves)))
        if not mesh:
            marks = sum(1 << c for c, ch in enumerate(internal) if any(k < 0 for k in wide_children(wide[ch])))
            assert int(wide["leaf_info"][v]) == marks, f"node {v}: leaf marks"
        stack.extend(reversed(internal))
    reached = seen_node > 0
    assert np.all(seen_node[reached] == 1), "every four-wide node is reached exactly once"
    # the slots no node reaches are padding: an inverted box in all four slots
    for v in np.nonzero(~reached)[0]:
        assert all(int(wide["lo"][v][k]) == 0xFFFFFFFF and int(wide["hi"][v][k]) == 0 for k in range(3)), v
    # every family of siblings starts on an even slot (a 128-byte line) and is contiguous
    for v in order:
        kids = [ch for ch in wide_children(wide[v]) if ch >= 0]
        if kids:
            assert min(kids) % 2 == 0, f"node {v}: family at odd slot {min(kids)}"
            assert kids == list(range(min(kids), min(kids) + len(kids))), "siblings are adjacent"
    assert (~reached).sum() <= reached.sum() + 1
    assert np.all(seen_prim == 1), f"{np.sum(seen_prim != 1)} primitives are not exactly one leaf"
    if mesh:
        assert leaf_order == list(range(n)), "a mesh's primitives are stored in the tree's depth-first order"
    sub_lo = np.full((len(wide), 3), np.inf)
    sub_hi = np.full((len(wide), 3), -np.inf)
    bad = 0
    for v in reversed(order):
        for c, ch in enumerate(wide_children(wide[v])):
            blo, bhi = (plo[~ch], phi[~ch]) if ch < 0 else (sub_lo[ch], sub_hi[ch])
            qlo, qhi = _decode(wide[v], c)
            if np.any(qlo > blo) or np.any(qhi < bhi):
                bad += 1
                if bad < 4:
                    print("node", v, "child", c, "->", ch, "box", qlo, qhi, "needs", blo, bhi)
            sub_lo[v] = np.minimum(sub_lo[v], blo)
            sub_hi[v] = np.maximum(sub_hi[v], bhi)
    assert bad == 0, f"{bad} quantised child boxes do not enclose their subtree"


def _one_disc():
    return {"kind": "discs", "xyzr": np.array([[0.0, 0.0, 0.0, 0.5]], np.float32),
            "normals": np.array([[0.0, 0.0, 1.0]], np.float32)}


def _two_discs():
    return {"kind": "discs", "xyzr": np.array([[-0.3, 0.0, 0.0, 0.25], [0.3, 0.0, 0.1, 0.25]], np.float32),
            "normals": np.array([[0.0, 0.0, 1.0], [0.0, 0.6, 0.8]], np.float32)}


SCENES = {
    "one_disc": _one_disc,
    "two_discs": _two_discs,
    "small_trench": small_trench,
    "hole_discs": lambda: geometry.hole_discs(spacing=0.05),
    "hole_mesh": small_hole_mesh,
    "trench_mesh": lambda: geometry.trench_triangles(flat=1.0, ylen=0.5, spacing=0.05),
}


@pytest.mark.parametrize("builder", [1, 2], ids=["lbvh", "ploc"])
@pytest.mark.parametrize("which", list(SCENES))
def test_wide_tree_encloses_every_primitive(which, builder):
    sc = SCENES[which]()
    n = sc["xyzr"].shape[0] if sc["kind"] == "discs" else sc["tris"].shape[0]
    print(f"[gpu] wide tree {which} n={n} tree_builder {builder}", flush=True)
    with Context() as ctx:
        ctx.set_scene(sc)
        ctx.set_params(n_rays=16, sticking=1.0, tree_builder=builder)
        _, perm, _ = ctx.hierarchy()
        wide = ctx.wide_hierarchy()
        st = ctx.stats()
    assert st["wide_nodes"] == len(wide)
    # at most n - 1 nodes plus one padding record per family of siblings
    assert len(wide) <= 2 * max(n - 1, 1) + 2
    assert 0 <= st["wide_stack"] <= 64
    _validate_wide(wide, perm, sc)


@pytest.mark.parametrize("which,kw", [
    ("small_trench", dict(sticking=0.1, bc_x=0, bc_y=1)),
    ("hole_discs", dict(sticking=0.05, bc_x=1, bc_y=1, diffuse_prob=0.6)),
    ("hole_mesh", dict(sticking=0.1, bc_x=1, bc_y=1, cos_exponent=10.0)),
    ("trench_mesh", dict(sticking=0.02, bc_x=0, bc_y=1)),
])
def test_wide_and_binary_traversal_agree_bit_for_bit(which, kw):
    sc = SCENES[which]()
    print(f"[gpu] wide vs binary {which}", flush=True)
    base = dict(n_rays=60_000, seed=11, **kw)
    a = run_scene(sc, bvh_width=4, **base)
    b = run_scene(sc, bvh_width=2, **base)
    b6 = run_scene(sc, bvh_width=2, lds_stack=6, block_size=192, **base)
    c = run_scene(sc, bvh_width=4, kernel=1, **base)
    # leaf tests at each node visit, and postponed leaf tests on the wave's
    # leaf list with a leaf phase as soon as one leaf is listed / only once
    # the listed leaves reach four per lane with a node (small block and LDS
    # stack: the private stack part is used)
    d = run_scene(sc, bvh_width=4, leaf_batch=-1, **base)
    e = run_scene(sc, bvh_width=4, leaf_batch=1, **base)
    f = run_scene(sc, bvh_width=4, leaf_batch=256, block_size=64, lds_stack=4, **base)
    f6 = run_scene(sc, bvh_width=4, leaf_batch=48, block_size=128, lds_stack=6, **base)
    # the other binary-tree topology: the closest hit does not depend on the tree
    g = run_scene(sc, tree_builder=1, **base)
    h = run_scene(sc, tree_builder=2, **base)
    # the counting kernel on the wave-shared leaf list, and with tnear = 0
    # (hit distances of -0 and +0 must order alike)
    k = run_scene(sc, bvh_width=4, leaf_batch=32, kernel=2, **base)
    for other in (b, b6, c, d, e, f, f6, g, h, k):
        assert np.array_equal(a["counts"], other["counts"])
        assert np.array_equal(a["raw"], other["raw"])
        assert a["stats"]["segments"] == other["stats"]["segments"]
    z0 = run_scene(sc, bvh_width=4, leaf_batch=-1, tnear=0.0, **base)
    z1 = run_scene(sc, bvh_width=4, leaf_batch=16, tnear=0.0, **base)
    assert np.array_equal(z0["raw"], z1["raw"]) and np.array_equal(z0["counts"], z1["counts"])
    # adjoint sums with and without postponed leaves
    j0 = run_scene(sc, bvh_width=4, leaf_batch=-1, adjoint=1, **base)
    j1 = run_scene(sc, bvh_width=4, leaf_batch=48, adjoint=1, **base)
    for key in ("raw", "counts", "ds", "dn"):
        assert np.array_equal(j0[key], j1[key]), key


def test_wide_tree_visits_fewer_nodes():
    sc = small_hole_mesh()
    print("[gpu] wide vs binary node visits", flush=True)
    kw = dict(n_rays=50_000, sticking=0.1, bc_x=1, bc_y=1, kernel=2, seed=3)
    a = run_scene(sc, bvh_width=4, **kw)["counters"]
    b = run_scene(sc, bvh_width=2, **kw)["counters"]
    assert a[0] == b[0] and a[7] == b[7]   # same segments, same deposits
    assert a[4] < 0.75 * b[4], f"four-wide {a[4]} vs binary {b[4]} node visits"


def test_invalid_width_is_rejected():
    with Context() as ctx:
        with pytest.raises(Exception):
            ctx.set_params(bvh_width=3)


def test_invalid_leaf_batch_is_rejected():
    with Context() as ctx:
        for v in (-2, 257):
            with pytest.raises(Exception):
                ctx.set_params(leaf_batch=v)
        for v in (-1, 65):
            with pytest.raises(Exception):
                ctx.set_params(refill_lanes=v)


def test_invalid_tree_builder_is_rejected():
    with Context() as ctx:
        for v in (-1, 3):
            with pytest.raises(Exception):
                ctx.set_params(tree_builder=v)
