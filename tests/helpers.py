"""Shared helpers for the parity tests (scenes, oracle runs, statistics)."""
import math

import numpy as np

import oracle
from rti_amd import geometry


def small_trench(spacing=0.05, depth=2.0, flat=0.5, ylen=0.5):
    return geometry.trench_discs(width=1.0, depth=depth, flat=flat, ylen=ylen, spacing=spacing)


def small_hole_mesh():
    return geometry.hole_triangles(radius=0.5, depth=2.0, half=1.0, n_az=64, n_ax=32, n_floor=6, n_top=8)


def oracle_run(scene, n_rays, sticking, seed=1, bc_x=0, bc_y=0, cos_exponent=1.0, diffuse_prob=1.0, src_z=None,
               fp64=False, adjoint=False, roulette=True, threads=0, ray_begin=0, moments=False, source=0,
               tie_rank=None):
    box, sz = geometry.scene_box(scene, src_z)
    return oracle.trace(scene, n_rays=n_rays, box=box, src_z=sz, sticking=sticking, seed=seed, bc_x=bc_x, bc_y=bc_y,
                        cos_exponent=cos_exponent, diffuse_prob=diffuse_prob, fp64=fp64, adjoint=adjoint,
                        roulette=roulette, threads=threads, ray_begin=ray_begin, moments=moments, source=source,
                        tie_rank=tie_rank)


def tie_rank(perm):
    """Rank of each input primitive in the library's stored order (perm maps
    stored position -> input id): the library breaks exact distance ties by
    stored position, and the oracle follows that order when given it."""
    rank = np.empty(len(perm), dtype=np.uint32)
    rank[np.asarray(perm, dtype=np.int64)] = np.arange(len(perm), dtype=np.uint32)
    return rank


def zscores(a, sa, b, sb, min_count=None, ca=None, cb=None):
    """Per-element z = (a - b) / sqrt(sa^2 + sb^2) over elements with signal."""
    s = np.sqrt(sa ** 2 + sb ** 2)
    m = s > 0
    if min_count is not None:
        m &= (ca >= min_count) & (cb >= min_count)
    return (a[m] - b[m]) / s[m]


def check_z(z, max_frac_3sigma=0.01):
    """z-scores of two independent estimates of the same means."""
    assert z.size > 50, f"too few elements with signal ({z.size})"
    frac = float(np.mean(np.abs(z) > 3.0))
    assert frac <= max_frac_3sigma, f"{frac:.4f} of elements outside 3 sigma (expected 0.0027)"
    assert abs(float(np.mean(z))) < 4.0 / np.sqrt(z.size) + 0.05, f"mean z {np.mean(z):.3f}"
    sd = float(np.std(z))
    assert 0.7 < sd < 1.3, f"std of z {sd:.3f}"
    return frac, float(np.mean(z)), sd


def footprint_check(counts, sc, n_rays, centre, radius, margin=1e-3):
    """Near-vertical rays from a disc source onto a flat triangle plate with
    s = 1: a triangle wholly inside the disc receives Binomial(N, A / (pi R^2))
    deposits, one wholly outside none."""
    v, t = sc["verts"].astype(np.float64), sc["tris"]
    d = np.hypot(v[t][:, :, 0] - centre[0], v[t][:, :, 1] - centre[1])
    inside = d.max(1) < radius - margin
    outside = d.min(1) > radius + margin
    e1, e2 = v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]]
    area = 0.5 * np.abs(np.cross(e1, e2)[:, 2])
    p = area[inside] / (math.pi * radius * radius)
    exp = n_rays * p
    z = (counts[inside] - exp) / np.sqrt(exp * (1 - p))
    assert inside.sum() > 300 and outside.sum() > 50
    assert np.all(counts[outside] == 0)
    assert float(np.mean(np.abs(z) > 3)) < 0.01
    assert abs(float(z.mean())) < 4 / math.sqrt(z.size) + 0.05 and 0.7 < float(z.std()) < 1.3


# ---------------------------------------------------------------------------
# Mirror trench and its method-of-images solution.
#
# A triangle-mesh trench (width 1, depth 2, flats of 0.5 on either side, y in
# [0, 1], quads of 0.25) whose surfaces all reflect as mirrors, traced without
# rejection control, reflective x faces at x = +-1, periodic y, source plane
# z = 2.5. Every path is then a straight line in the unfolded space of the
# mirrors: the lateral faces fold x with period 4 above the trench, the walls
# fold it with period 2 inside, the floor folds z. Each surface hit k of a
# path (k = 0, 1, ...) deposits s (1 - s)^k.
# ---------------------------------------------------------------------------
MIRROR = {"width": 1.0, "depth": 2.0, "flat": 0.5, "ylen": 1.0, "spacing": 0.25, "src_z": 2.5}


def mirror_trench():
    m = MIRROR
    return geometry.trench_triangles(width=m["width"], depth=m["depth"], flat=m["flat"], ylen=m["ylen"],
                                     spacing=m["spacing"])


def mirror_quads(scene):
    """(quad of each triangle, number of quads): quads are indexed row-major
    by (y row, profile segment), as trench_triangles lays them out."""
    m = MIRROR
    n_seg = int(round((2 * m["flat"] + m["width"] + 2 * m["depth"]) / m["spacing"]))
    n_row = int(round(m["ylen"] / m["spacing"]))
    nt = scene["tris"].shape[0]
    q = np.arange(nt) % (n_seg * n_row)
    return q, n_seg * n_row, n_seg


def _fold(u, half):
    """Fold the unfolded coordinate u into [-half, half] between mirrors at
    +-half: (folded coordinate, +1 or -1 = direction kept or flipped)."""
    k = np.floor((u + half) / (2 * half))
    even = (k % 2) == 0
    x = np.where(even, u - 2 * half * k, 2 * half * k - u)
    return x, np.where(even, 1.0, -1.0)


def mirror_images(org, dirn, s):
    """Method-of-images hits of the mirror trench for rays (org, dirn) (n x 3).

    Returns flat arrays over all hits (ray index, hit index k, profile arc
    coordinate, y, deposit s (1-s)^k, surface kind 0 flat / 1 wall / 2 floor).
    The arc coordinate runs along the trench profile from the left flat's
    outer edge (0) over the left wall (0.5 .. 2.5), the floor (2.5 .. 3.5)
    and the right wall (3.5 .. 5.5) to the right flat's outer edge (6)."""
    m = MIRROR
    D, w2, L, zs = m["depth"], m["width"] / 2, m["ylen"], m["src_z"]
    half_box = w2 + m["flat"]
    o = np.asarray(org, dtype=np.float64)
    d = np.asarray(dirn, dtype=np.float64)
    n = o.shape[0]
    assert np.all(d[:, 2] < 0)
    t1 = (zs - D) / -d[:, 2]
    x1, sg = _fold(o[:, 0] + d[:, 0] * t1, half_box)
    y1 = o[:, 1] + d[:, 1] * t1
    dx1 = d[:, 0] * sg
    rays, ks, arcs, ys, kinds = [], [], [], [], []
    # flats at z = D: one hit, then the ray leaves upwards
    flat = np.abs(x1) > w2
    idx = np.nonzero(flat)[0]
    rays.append(idx)
    ks.append(np.zeros(idx.size, dtype=np.int64))
    # left flat: arc = x + half_box; right flat: arc = flat + 2 D + width + (x - width / 2)
    arcs.append(np.where(x1[idx] < 0, x1[idx] + half_box, m["flat"] + 2 * D + 2 * w2 + (x1[idx] - w2)))
    ys.append(y1[idx])
    kinds.append(np.zeros(idx.size, dtype=np.int64))
    # trench: wall crossings down to the floor and back up to z = D
    idx = np.nonzero(~flat)[0]
    xa, dxa, ya, dya, dza = x1[idx], dx1[idx], y1[idx], d[idx, 1], d[idx, 2]
    tf = D / -dza
    v_end = xa + dxa * 2 * tf
    # wall planes v = p + 1/2 strictly between the start and the end
    lo = np.minimum(xa, v_end)
    hi = np.maximum(xa, v_end)
    p_first = np.ceil(lo - 0.5).astype(np.int64)
    p_last = np.floor(hi - 0.5).astype(np.int64)
    nwall = np.maximum(p_last - p_first + 1, 0)
    rep = np.repeat(np.arange(idx.size), nwall)
    off = np.arange(rep.size) - np.repeat(np.cumsum(nwall) - nwall, nwall)
    plane = p_first[rep] + off
    tau_w = (plane + 0.5 - xa[rep]) / dxa[rep]
    z_w = np.abs(D + dza[rep] * tau_w)
    right = (plane % 2) == 0
    # right wall: arc = flat + D + width + z ; left wall: arc = flat + (D - z)
    arc_w = np.where(right, m["flat"] + D + 2 * w2 + z_w, m["flat"] + (D - z_w))
    # floor hit at tf
    xf, _ = _fold(xa + dxa * tf, w2)
    arc_f = m["flat"] + D + (xf + w2)
    # order the hits of each ray by distance: the floor at tf
    all_ray = np.concatenate([rep, np.arange(idx.size)])
    all_tau = np.concatenate([tau_w, tf])
    all_arc = np.concatenate([arc_w, arc_f])
    all_y = np.concatenate([ya[rep] + dya[rep] * tau_w, ya + dya * tf])
    all_kind = np.concatenate([np.ones(rep.size, dtype=np.int64), np.full(idx.size, 2, dtype=np.int64)])
    order = np.lexsort((all_tau, all_ray))
    all_ray, all_arc, all_y, all_kind = all_ray[order], all_arc[order], all_y[order], all_kind[order]
    counts = np.bincount(all_ray, minlength=idx.size)
    k = np.arange(all_ray.size) - np.repeat(np.cumsum(counts) - counts, counts)
    rays.append(idx[all_ray])
    ks.append(k)
    arcs.append(all_arc)
    ys.append(all_y)
    kinds.append(all_kind)
    ray = np.concatenate(rays)
    k = np.concatenate(ks)
    dep = s * (1.0 - s) ** k
    y = np.mod(np.concatenate(ys), L)
    return ray, k, np.concatenate(arcs), y, dep, np.concatenate(kinds), n


def mirror_quad_sums(arc, y, dep, n_quads, n_seg):
    """Deposit sums and counts per quad of the mirror trench."""
    h = MIRROR["spacing"]
    i = np.clip(np.floor(arc / h).astype(np.int64), 0, n_seg - 1)
    j = np.clip(np.floor(y / h).astype(np.int64), 0, n_quads // n_seg - 1)
    q = j * n_seg + i
    return np.bincount(q, weights=dep, minlength=n_quads), np.bincount(q, minlength=n_quads)
