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
