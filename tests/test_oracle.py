"""Pins of the CPU oracle (runs without a GPU).

Each oracle function is checked against what fixes it independently of the
oracle's own formula: published known-answer vectors, the reference's own
known-answer tests (tests/golden/*.json), the reference's headers compiled by
oracle/Makefile `ref` (oracle/_ref, built only where the reference checkout
exists), closed forms and brute force.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from rti_amd import geometry
from tests.helpers import footprint_check

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _ref():
    L = oracle.ref_lib()
    if L is None:
        oracle.build_ref()
        L = oracle.ref_lib()
    if L is None:
        pytest.skip("reference checkout not available: oracle/_ref not built")
    return L


_fp = oracle._fp
_u32p = oracle._u32p
_u64p = oracle._u64p


def test_philox_random123_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert [hex(x) for x in oracle.philox([0, 0, 0, 0], [0, 0])] == \
        ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(x) for x in oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)] == \
        ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(x) for x in oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                                          [0xA4093822, 0x299F31D0])] == \
        ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]


def test_local_intersector_known_answers():
    k = json.load(open(os.path.join(GOLD, "local_intersector_kats.json")))
    assert len(k["cases"]) == 9
    for c in k["cases"]:
        assert oracle.local_intersect(c["org"], c["dir"], c["disc"], c["normal"]) == c["hit"], c["name"]


def test_local_intersector_closed_form():
    """Ray-plane hit inside the radius, front face only, t > 0."""
    rng = np.random.default_rng(0)
    for _ in range(2000):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        c = rng.uniform(-1, 1, 3)
        r = rng.uniform(0.1, 1.0)
        o = rng.uniform(-3, 3, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        dn = d @ n
        t = (c - o) @ n / dn if dn != 0 else -1
        h = o + t * d
        expect = dn < -1e-6 and t > 1e-6 and np.linalg.norm(h - c) < r * (1 - 1e-5)
        clear_miss = dn > 1e-6 or t < -1e-6 or np.linalg.norm(h - c) > r * (1 + 1e-5)
        got = oracle.local_intersect(o, d, list(c) + [r], n)
        if expect:
            assert got
        elif clear_miss:
            assert not got


def test_area_inside_known_answers():
    k = json.load(open(os.path.join(GOLD, "area_inside_kats.json")))
    assert len(k["cases"]) == 37
    for c in k["cases"]:
        v = oracle.area_inside(c["box"], c["disc"], c["normal"])
        if c["tol"] == 0:
            assert v == c["expected"], (c["name"], v, c["expected"])
        else:
            assert abs(v - c["expected"]) < c["tol"], (c["name"], v, c["expected"])


def test_area_inside_closed_forms():
    # a disc covering the whole box: area = box area / |n_z| (projection)
    for n in ([0, 0, 1], [0.3, 0.2, 0.9], [-0.5, 0.7, 0.4]):
        nn = np.array(n, dtype=np.float64)
        nn /= np.linalg.norm(nn)
        v = oracle.area_inside((-0.5, -0.25, 0.5, 0.25), [0.01, -0.02, 0.3, 40.0], nn)
        assert abs(v - 0.5 / abs(nn[2])) < 2e-3 * (0.5 / abs(nn[2])), (n, v)
    # circular segment: flat disc crossing one face at distance a from the centre
    for a in (0.1, 0.5, -0.3):
        v = oracle.area_inside((-10, -10, a, 10), [0, 0, 0, 1], [0, 0, 1])
        ang = 2 * math.acos(a)
        assert abs(v - (math.pi - 0.5 * (ang - math.sin(ang)))) < 1e-5


def test_area_inside_matches_compiled_reference():
    L = _ref()
    rng = np.random.default_rng(1)
    box = (-1.0, -2.0, 3.0, 1.5)
    bad = 0
    for _ in range(5000):
        r = np.float32(rng.choice([0.05, 0.3, 1.0, 3.0]))
        c = np.array([rng.uniform(-1.5, 3.5), rng.uniform(-2.5, 2), rng.uniform(-1, 1), r], np.float32)
        n = rng.normal(size=3).astype(np.float32)
        n /= np.linalg.norm(n)
        a = oracle.area_inside(box, c, n)
        b = L.ref_area_inside(*[oracle.ctypes.c_float(v) for v in box], c.ctypes.data_as(_fp), n.ctypes.data_as(_fp))
        if not (a == b or (math.isnan(a) and math.isnan(b))):
            bad += 1
    assert bad == 0


def _ref_neighbors(L, x):
    n = x.shape[0]
    row = np.zeros(n + 1, dtype=np.uint32)
    e = L.ref_neighbors(x.ctypes.data_as(_fp), n, row.ctypes.data_as(_u32p), None)
    col = np.zeros(max(e, 1), dtype=np.uint32)
    L.ref_neighbors(x.ctypes.data_as(_fp), n, row.ctypes.data_as(_u32p), col.ctypes.data_as(_u32p))
    return row, col[:e]


@pytest.mark.parametrize("which", ["trench_c1", "hole", "rough"])
def test_neighbors_match_compiled_reference(which):
    """The reference's divide-and-conquer neighbourhood (run from its own
    header) and the oracle's grid search give the same graph, including the
    FP32-borderline diagonal pairs of the 0.02 grid (2 r = spacing * sqrt 2)."""
    L = _ref()
    if which == "trench_c1":
        sc = geometry.trench_discs()
    elif which == "hole":
        sc = geometry.hole_discs(spacing=0.04)
    else:
        sc = geometry.rough_surface_discs(nfield=256, cells=1, spacing=0.02)
    x = np.ascontiguousarray(sc["xyzr"], dtype=np.float32)
    row, col = _ref_neighbors(L, x)
    orow, ocol = oracle.neighbors(x)
    assert np.array_equal(row, orow)
    assert np.array_equal(col, ocol)


def test_neighbors_brute_force():
    rng = np.random.default_rng(2)
    # a jittered grid puts many pairs near the contact distance
    g = np.stack(np.meshgrid(np.arange(12), np.arange(12), np.arange(3)), -1).reshape(-1, 3) * 0.1
    x = np.zeros((g.shape[0], 4), np.float32)
    x[:, :3] = g + rng.normal(scale=0.003, size=g.shape)
    x[:, 3] = np.float32(0.1 * math.sqrt(2) / 2)
    row, col = oracle.neighbors(x)
    n = x.shape[0]
    for i in range(n):
        d = (x[:, :3] - x[i, :3]).astype(np.float32)
        s = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(np.float32) + d[:, 2] * d[:, 2]
        m = np.sqrt(s.astype(np.float32)) < (x[:, 3] + x[i, 3])
        m[i] = False
        assert np.array_equal(np.nonzero(m)[0], col[row[i]:row[i + 1]])


def test_source_direction_matches_compiled_reference():
    """Same uniforms in, same direction out (power_cosine_direction_z)."""
    L = _ref()
    rng = np.random.default_rng(3)
    for n in (1.0, 5.0, 10.0):
        for _ in range(500):
            a, b = rng.integers(1, 1 << 24, 2)
            draws = np.array([int(a) << 40, int(b) << 40], dtype=np.uint64)
            out = np.zeros(3, np.float32)
            L.ref_power_cosine_direction(n, draws.ctypes.data_as(_u64p), out.ctypes.data_as(_fp))
            mine = oracle.source_dir(a * 2.0 ** -24, b * 2.0 ** -24, n)
            np.testing.assert_allclose(mine, out, atol=2e-6)


def test_source_direction_moments():
    """cos theta = U^(1/(n+1)): E[cos] = (n+1)/(n+2), azimuth uniform."""
    rng = np.random.default_rng(4)
    for n in (1.0, 10.0):
        u = rng.random((40000, 2))
        d = np.array([oracle.source_dir(a, 1.0 - b, n) for a, b in u])
        c = -d[:, 2]
        assert abs(c.mean() - (n + 1) / (n + 2)) < 4 * c.std() / math.sqrt(len(c))
        assert abs(d[:, 0].mean()) < 0.02 and abs(d[:, 1].mean()) < 0.02
        np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)


def test_diffuse_direction_is_cosine_lobe_like_reference():
    """Own orthonormal basis (Duff et al. 2017) vs the reference's: the same
    cosine-weighted lobe about the normal (moments), for reference draws too."""
    L = _ref()
    rng = np.random.default_rng(5)
    for nrm in ([0, 0, 1], [1, 0, 0], [0.3, -0.4, -0.866], [0, 0, -1]):
        n = np.array(nrm, dtype=np.float32)
        n /= np.linalg.norm(n)
        mine, ref = [], []
        for _ in range(20000):
            a, b = rng.integers(0, 1 << 24, 2)
            mine.append(oracle.diffuse_dir(n, a * 2.0 ** -24, b * 2.0 ** -24))
            draws = np.array([int(a) << 40, int(b) << 40], dtype=np.uint64)
            o = np.zeros(3, np.float32)
            L.ref_cos_hemi(n.ctypes.data_as(_fp), draws.ctypes.data_as(_u64p), o.ctypes.data_as(_fp))
            ref.append(o)
        for arr in (np.array(mine), np.array(ref)):
            c = arr @ n
            assert c.min() > -1e-6
            assert abs(c.mean() - 2 / 3) < 0.01       # E[cos] = 2/3 for a Lambertian lobe
            assert abs((c * c).mean() - 0.5) < 0.01   # E[cos^2] = 1/2
            t = arr - np.outer(c, n)
            assert np.abs(t.mean(0)).max() < 0.015    # no tangential bias


def test_roulette_matches_compiled_reference():
    L = _ref()
    for w in (0.5, 0.1, 0.0999, 0.05, 0.01, 0.2999):
        for u in np.linspace(0, 1, 97, endpoint=False):
            draw = int(u * 2 ** 24) << 40
            ww = np.array([w], np.float32)
            rk = L.ref_rejection_control(ww.ctypes.data_as(_fp), draw)
            keep, w2 = oracle.roulette(w, draw / 2.0 ** 64)
            assert keep == bool(rk), (w, u)
            if keep:
                assert w2 == float(ww[0])


def test_roulette_unbiased():
    rng = np.random.default_rng(6)
    for w in (0.09, 0.05, 0.01):
        us = rng.random(200000)
        tot = 0.0
        for u in us[:20000]:
            keep, w2 = oracle.roulette(w, u)
            tot += w2 if keep else 0.0
        assert abs(tot / 20000 - w) < 4 * math.sqrt(w * 0.3 / 20000)


def test_smooth_normalize_closed_form():
    v = np.array([1.0, 2.0, 4.0, 0.0])
    a = np.array([1.0, 2.0, 1.0, 1.0], np.float32)
    row = np.array([0, 1, 3, 4, 4], np.uint32)
    col = np.array([1, 0, 2, 1], np.uint32)
    out = oracle.smooth_normalize(v, a, row, col)
    raw = np.array([(1 + 1) / 2, (1 + 1 + 4) / 3, (4 + 1) / 2, 0.0])
    np.testing.assert_allclose(out, raw / raw.max(), rtol=1e-6)


def test_trace_flat_plate_flux_and_counts():
    """Direct flux on a periodic flat plate is 1 for any cos^n source; the
    number of segments equals rays + boundary crossings."""
    sc = geometry.flat_plate_discs(size=1.0, spacing=0.05)
    box, sz = geometry.scene_box(sc)
    for nexp in (1.0, 10.0):
        r = oracle.trace(sc, n_rays=100000, box=box, src_z=sz, sticking=1.0, bc_x=1, bc_y=1, cos_exponent=nexp)
        areas = np.array([oracle.area_inside((box[0], box[2], box[1], box[3]), sc["xyzr"][i], sc["normals"][i])
                          for i in range(sc["xyzr"].shape[0])])
        flux = r["s1"] / areas * (box[1] - box[0]) * (box[3] - box[2]) / 100000
        assert abs(flux.mean() - 1) < 0.02
        assert r["surface_hits"] + r["misses"] == 100000
        assert r["misses"] < 100


def test_trace_path_weight_identity():
    """Without roulette a history that makes K front hits deposits
    s sum_{k<K} (1-s)^k = 1 - (1-s)^K on the surface (neighbour copies
    excluded: a scene without overlaps), and d/ds of that is K (1-s)^(K-1)."""
    sc = geometry.hole_triangles(radius=0.5, depth=1.0, half=1.0, n_az=32, n_ax=8, n_floor=4, n_top=4)
    box, sz = geometry.scene_box(sc)
    s = 0.3
    r = oracle.trace(sc, n_rays=20000, box=box, src_z=sz, sticking=s, bc_x=1, bc_y=1, roulette=False, adjoint=True)
    # total deposit per ray summed over the surface is 1 - (1-s)^K; its mean
    # equals 1 - E[(1-s)^K], and the total d/ds equals E[K (1-s)^(K-1)].
    # With no roulette both are functions of the per-ray hit count; check the
    # identity d/ds sum = sum over k of d/ds of s (1-s)^k through a second run
    # at s + h with the same paths (paths do not depend on s).
    h = 1e-3
    r2 = oracle.trace(sc, n_rays=20000, box=box, src_z=sz, sticking=s + h, bc_x=1, bc_y=1, roulette=False)
    r1 = oracle.trace(sc, n_rays=20000, box=box, src_z=sz, sticking=s - h, bc_x=1, bc_y=1, roulette=False)
    fd = (r2["s1"].sum() - r1["s1"].sum()) / (2 * h)
    assert abs(r["ds"].sum() - fd) < 1e-3 * abs(fd)
    assert r["surface_hits"] > 0


# ---------------------------------------------------------------------------
# Relative error and variance of the variance (hit_accumulator.hpp:152-249):
# the power-sum formulas against the definitions on explicit sample sets.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed,n,p_hit", [(1, 50, 0.5), (2, 400, 0.1), (3, 2000, 0.02), (4, 37, 1.0)])
def test_error_stats_equal_definitions(seed, n, p_hit):
    rng = np.random.default_rng(seed)
    k = 5  # primitives; each history scores on at most one of them
    who = rng.integers(0, k, n)
    hit = rng.random(n) < p_hit
    val = rng.random(n) * 0.9 + 0.05
    x = np.zeros((k, n))
    x[who[hit], np.nonzero(hit)[0]] = val[hit]
    s = [x.sum(1), (x ** 2).sum(1), (x ** 3).sum(1), (x ** 4).sum(1)]
    cnt = (x > 0).sum(1).astype(np.uint64)
    # N = all deposits of the run; here every nonzero score is one deposit
    total = int(cnt.sum())
    rel, vov = oracle.error_stats(s[0], s[1], cnt, s[2], s[3], total=total)
    for i in range(k):
        if cnt[i] == 0:
            assert rel[i] == np.finfo(np.float64).max and vov[i] == np.finfo(np.float64).max
            continue
        # the same statistics over N = total samples (zeros for other deposits)
        xi = np.zeros(total)
        xi[: int(cnt[i])] = x[i][x[i] > 0]
        d = xi - xi.mean()
        m2 = (d ** 2).sum()
        # relative error of the mean with sqrt(N - 1) ~ sqrt(N)
        assert rel[i] == pytest.approx(np.sqrt(m2) / xi.sum(), rel=1e-10)
        if m2 > 0:
            assert vov[i] == pytest.approx((d ** 4).sum() / m2 ** 2 - 1.0 / total, rel=1e-8, abs=1e-12)


def test_error_stats_undefined_cases():
    big = np.finfo(np.float64).max
    s1 = np.array([0.0, 2.0, 3.0])
    s2 = np.array([0.0, 2.0, 3.0])
    cnt = np.array([0, 2, 3], dtype=np.uint64)
    # element 0: S1 = 0 and no deposit; element 2 holds every deposit of a run of
    # equal values, so N S2 - S1^2 = 0
    rel, vov = oracle.error_stats(s1, s2, cnt, s2, s2, total=3)
    assert rel[0] == big and vov[0] == big
    assert vov[2] == big
    assert rel[2] == pytest.approx(0.0, abs=1e-12)
    # no deposits at all: every relative error is undefined
    rel0, _ = oracle.error_stats(np.zeros(2), np.zeros(2), np.zeros(2, dtype=np.uint64), total=0)
    assert np.all(rel0 == big)


def test_disc_source_origins_uniform_in_inscribed_disc():
    """disc_origin_z (src/rti/ray/disc_origin_z.hpp:17-33) on the disc of
    create_circular_source_from_bounding_box (src/rti/device.hpp:269-280):
    centre of the box, radius half its x extent. Uniform over the disc means
    r^2 / R^2 ~ U(0, 1) and the angle ~ U(0, 2 pi) (a radius drawn uniformly,
    the usual mistake, gives E[r^2] = R^2 / 3); the source plane is exact."""
    from scipy import stats
    box = (-1.5, 0.5, 2.0, 4.0, 0.0, 1.0)  # centre (-0.5, 3), R = 1
    o = oracle.source_origins(box, 2.5, 200_000, source=1, seed=7, instance=3).astype(np.float64)
    dx, dy = o[:, 0] + 0.5, o[:, 1] - 3.0
    r2 = dx * dx + dy * dy
    assert r2.max() <= 1.0 + 1e-6
    assert np.all(o[:, 2] == np.float32(2.5))
    assert stats.kstest(r2, "uniform").pvalue > 1e-3
    assert stats.kstest((np.arctan2(dy, dx) + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3
    assert abs(r2.mean() - 0.5) < 4 * math.sqrt(1 / 12 / r2.size)
    assert abs(dx.mean()) < 4 * 0.5 / math.sqrt(r2.size) and abs(dy.mean()) < 4 * 0.5 / math.sqrt(r2.size)


def test_disc_source_first_draw_layout():
    """A ray whose first Philox block (ray, 0, 0, 0) gives offsets inside the
    disc starts there (the rectangle source uses the same two words); the
    others are redrawn, so no origin equals a rejected candidate."""
    box = (0.0, 2.0, 0.0, 2.0, 0.0, 1.0)
    n = 4000
    o = oracle.source_origins(box, 1.0, n, source=1, seed=5, instance=2)
    inside = 0
    for ray in range(n):
        u = oracle.philox([ray, 0, 0, 0], [5, 2])
        f = (u[:2] >> 8).astype(np.float64) * 2.0 ** -24
        d = (2 * f - 1).astype(np.float32) * np.float32(1.0)
        cand = np.float32(1.0) + d
        if float(d[0]) ** 2 + float(d[1]) ** 2 < 1.0 - 1e-6:
            inside += 1
            np.testing.assert_array_equal(o[ray, :2], cand)
        elif float(d[0]) ** 2 + float(d[1]) ** 2 > 1.0 + 1e-6:
            assert not np.array_equal(o[ray, :2], cand)
    assert abs(inside / n - math.pi / 4) < 4 * math.sqrt(0.25 * 0.75 / n)


def test_trace_disc_source_footprint_on_flat_mesh():
    sc = geometry.flat_plate_triangles(size=1.0, cells=32)
    box = (0.0, 1.0, 0.0, 1.0, 0.0, 0.01)
    n = 200_000
    r = oracle.trace(sc, n_rays=n, box=box, src_z=0.01, sticking=1.0, bc_x=1, bc_y=1, cos_exponent=1e4,
                     source=1)
    assert r["surface_hits"] == n
    footprint_check(r["cnt"].astype(np.float64), sc, n, (0.5, 0.5), 0.5)


def test_tie_rank_orders_equal_distance_hits():
    """Coplanar overlapping discs are hit at exactly equal distance; the
    closest is the one of lower rank (default: lower index). The identity
    rank is the default; another order moves deposits between elements (the
    neighbour deposits follow the winner's neighbour list) and leaves the
    counts of segments and hits alone."""
    sc = small_trench_scene()
    box, sz = geometry.scene_box(sc)
    n = sc["xyzr"].shape[0]
    kw = dict(n_rays=60_000, box=box, src_z=sz, sticking=1.0, bc_x=0, bc_y=1, threads=1)
    a = oracle.trace(sc, **kw)
    b = oracle.trace(sc, tie_rank=np.arange(n, dtype=np.uint32), **kw)
    c = oracle.trace(sc, tie_rank=np.arange(n, dtype=np.uint32)[::-1].copy(), **kw)
    assert np.array_equal(a["cnt"], b["cnt"]) and np.array_equal(a["s1"], b["s1"])
    assert not np.array_equal(a["cnt"], c["cnt"])
    assert (a["segments"], a["surface_hits"]) == (c["segments"], c["surface_hits"])


def small_trench_scene():
    return geometry.trench_discs(width=1.0, depth=2.0, flat=0.5, ylen=0.5, spacing=0.05)


# ---------------------------------------------------------------------------
# Mirror reflection (specular::use, src/rti/reflection/specular.hpp:16-40).
# ---------------------------------------------------------------------------
def test_specular_direction_invariants():
    """d' = d - 2 (d.n) n: |d'| = |d|, d'.n = -d.n, d' - d parallel to n,
    the tangential part unchanged, and d' = H d with the Householder matrix
    H = I - 2 n n^T evaluated in FP64."""
    rng = np.random.default_rng(21)
    for _ in range(2000):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        n32, d32 = n.astype(np.float32), d.astype(np.float32)
        r = oracle.specular_dir(d32, n32).astype(np.float64)
        nn, dd = n32.astype(np.float64), d32.astype(np.float64)
        assert np.linalg.norm(r) == pytest.approx(np.linalg.norm(dd), rel=1e-6)
        assert float(r @ nn) == pytest.approx(-float(dd @ nn), abs=1e-6)
        diff = r - dd
        assert np.linalg.norm(np.cross(diff, nn)) < 1e-6
        tang = dd - (dd @ nn) * nn
        assert np.allclose(r - (r @ nn) * nn, tang, atol=1e-6)
        house = np.eye(3) - 2.0 * np.outer(nn, nn)
        assert np.allclose(r, house @ dd, atol=1e-6)
    # axis-aligned mirrors flip exactly one component, bit for bit
    d = np.array([0.3, -0.4, -0.8660254], dtype=np.float32)
    for ax in range(3):
        n = np.zeros(3, dtype=np.float32)
        n[ax] = 1.0
        want = d.copy()
        want[ax] = -want[ax]
        assert np.array_equal(oracle.specular_dir(d, n), want)


def _mirror_kind(scene):
    """Surface kind of each triangle of the mirror trench: 0 flat, 1 wall, 2 floor."""
    from tests.helpers import MIRROR
    c = scene["verts"][scene["tris"]].mean(1)
    kind = np.ones(c.shape[0], dtype=np.int64)
    kind[c[:, 2] < 1e-6] = 2
    kind[(np.abs(c[:, 2] - MIRROR["depth"]) < 1e-6) & (np.abs(c[:, 0]) > MIRROR["width"] / 2)] = 0
    return kind


def _mirror_setup(n_rays, seed, s=0.3):
    from tests import helpers as H
    sc = H.mirror_trench()
    box, sz = geometry.scene_box(sc, H.MIRROR["src_z"])
    org, dirn = oracle.source_rays(box, sz, n_rays, seed=seed)
    img = H.mirror_images(org, dirn, s)
    return sc, box, sz, img


def _singular_rays(img, n_rays, delta=1e-3):
    """Rays with a hit within delta (arc length) of a corner of the profile:
    there a hit and the next one can lie closer than tnear, which the tracer
    skips by design (tracer.hpp:179), so the straight-line count does not
    apply."""
    ray, arc = img[0], img[2]
    corners = np.array([0.5, 2.5, 3.5, 5.5])
    near = np.min(np.abs(arc[:, None] - corners[None, :]), axis=1) < delta
    sing = np.zeros(n_rays, dtype=bool)
    sing[ray[near]] = True
    return sing


def test_mirror_trench_method_of_images_per_ray():
    """All-mirror trench without rejection control: every path is a straight
    line through the mirror images, so the number of hits of each ray, and
    what it deposits on the flats, walls and floor, follow from the source
    ray alone (computed in FP64 in tests/helpers.py, independently of the
    oracle's reflection and closest-hit code)."""
    n_rays, s = 2000, 0.3
    sc, box, sz, img = _mirror_setup(n_rays, seed=11, s=s)
    ray, k, arc, y, dep, kind_hit, _ = img
    sing = _singular_rays(img, n_rays)
    assert sing.sum() < 0.01 * n_rays
    kind_tri = _mirror_kind(sc)
    checked = 0
    for i in range(n_rays):
        if sing[i]:
            continue
        r = oracle.trace(sc, n_rays=1, ray_begin=i, box=box, src_z=sz, sticking=s, bc_x=0, bc_y=1, diffuse_prob=0.0,
                         roulette=False, seed=11)
        mine = ray == i
        assert int(r["cnt"].sum()) == int(mine.sum()), f"ray {i}: hit count"
        assert r["surface_hits"] == int(mine.sum())
        for kk in range(3):
            want = float(dep[mine & (kind_hit == kk)].sum())
            got = float(r["s1"][kind_tri == kk].sum())
            assert got == pytest.approx(want, rel=1e-5, abs=1e-30), f"ray {i}: deposit on surface kind {kk}"
        checked += 1
    assert checked > 0.99 * n_rays
    # the geometry makes most rays reflect off the walls
    assert np.mean(np.bincount(ray, minlength=n_rays)) > 2.0


def test_mirror_trench_method_of_images_per_quad():
    """The same at 5e4 rays per quad of the mesh: sums and hit totals within
    0.2 % (rays that pass within tnear of a corner lose their later hits)."""
    from tests import helpers as H
    n_rays, s = 50000, 0.3
    sc, box, sz, img = _mirror_setup(n_rays, seed=12, s=s)
    _, _, arc, y, dep, _, _ = img
    q, nq, nseg = H.mirror_quads(sc)
    want_s, want_c = H.mirror_quad_sums(arc, y, dep, nq, nseg)
    r = oracle.trace(sc, n_rays=n_rays, box=box, src_z=sz, sticking=s, bc_x=0, bc_y=1, diffuse_prob=0.0,
                     roulette=False, seed=12)
    got_s = np.bincount(q, weights=r["s1"], minlength=nq)
    got_c = np.bincount(q, weights=r["cnt"].astype(np.float64), minlength=nq)
    assert abs(got_c.sum() - want_c.sum()) <= 2e-3 * want_c.sum()
    assert np.abs(got_s - want_s).sum() <= 2e-3 * want_s.sum()
    assert np.all(np.abs(got_c - want_c) <= 20)


# ---------------------------------------------------------------------------
# Closest hit (rtcIntersect1 semantics, tracer.hpp:176-181: t in [tnear, tfar],
# no back-face culling; boundary faces boundary_x_y.hpp:147-235): the oracle's
# BVH traversal against brute force over every primitive and the four faces.
# ---------------------------------------------------------------------------
def _brute_closest(scene, box, org, dirn, tnear):
    """All primitives and faces per ray, FP32 arithmetic in the order of
    the oracle's expressions (its build has FP contraction off): returns
    (prim, face, t) as oracle.closest_hits does."""
    f = np.float32
    o = org.astype(f)
    d = dirn.astype(f)
    n = o.shape[0]
    tnear = f(tnear)
    # boundary faces, in the order 0 xmax, 2 xmin, 1 ymin, 3 ymax; later ties win
    best = np.full(n, np.finfo(f).max, dtype=f)
    face = np.full(n, -1, dtype=np.int32)
    xmin, xmax, ymin, ymax, zmin, zmax = (f(v) for v in box)
    with np.errstate(divide="ignore", invalid="ignore"):
        for fid, plane, ax, other, lo, hi in ((0, xmax, 0, 1, ymin, ymax), (2, xmin, 0, 1, ymin, ymax),
                                              (1, ymin, 1, 0, xmin, xmax), (3, ymax, 1, 0, xmin, xmax)):
            ok_dir = d[:, ax] != 0
            t = (plane - o[:, ax]) / d[:, ax]
            a = o[:, other] + d[:, other] * t
            z = o[:, 2] + d[:, 2] * t
            ok = ok_dir & (t >= tnear) & (t <= best) & (a >= lo) & (a <= hi) & (z >= zmin) & (z <= zmax)
            best = np.where(ok, t, best)
            face = np.where(ok, fid, face)
        if scene["kind"] == "discs":
            c = scene["xyzr"][:, :3].astype(f)
            rad = scene["xyzr"][:, 3].astype(f)
            nr = scene["normals"].astype(f)
            dn = (d[:, None, 0] * nr[None, :, 0] + d[:, None, 1] * nr[None, :, 1]) + d[:, None, 2] * nr[None, :, 2]
            co = c[None, :, :] - o[:, None, :]
            num = (co[..., 0] * nr[None, :, 0] + co[..., 1] * nr[None, :, 1]) + co[..., 2] * nr[None, :, 2]
            th = num / dn
            hx = (o[:, None, 0] + d[:, None, 0] * th) - c[None, :, 0]
            hy = (o[:, None, 1] + d[:, None, 1] * th) - c[None, :, 1]
            hz = (o[:, None, 2] + d[:, None, 2] * th) - c[None, :, 2]
            inside = ((hx * hx + hy * hy) + hz * hz) < (rad * rad)[None, :]
            valid = (dn != 0) & (th >= tnear) & (th <= best[:, None]) & inside
        else:
            v = scene["verts"].astype(f)
            t3 = scene["tris"]
            v0, v1, v2 = v[t3[:, 0]], v[t3[:, 1]], v[t3[:, 2]]
            e1, e2 = v1 - v0, v2 - v0

            def cross(a, b):
                return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                                 a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                                 a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)

            def dot(a, b):
                return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]

            pv = cross(d[:, None, :], e2[None, :, :])
            det = dot(e1[None, :, :], pv)
            idet = f(1) / det
            tv = o[:, None, :] - v0[None, :, :]
            u = dot(tv, pv) * idet
            qv = cross(tv, e1[None, :, :])
            vv = dot(d[:, None, :], qv) * idet
            th = dot(e2[None, :, :], qv) * idet
            valid = ((det != 0) & (u >= 0) & (u <= 1) & (vv >= 0) & ((u + vv) <= 1) & (th >= tnear)
                     & (th <= best[:, None]))
    thv = np.where(valid, th, np.inf).astype(np.float64)
    prim = np.argmin(thv, axis=1).astype(np.int32)  # first (lowest index) of equal minima
    hit = np.isfinite(thv[np.arange(n), prim])
    t = np.where(hit, th[np.arange(n), prim], best).astype(f)
    return np.where(hit, prim, -1), np.where(hit, -1, face), t


def _random_rays(rng, n, lo, hi):
    org = rng.uniform(lo, hi, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32)


def test_closest_hit_matches_brute_force_discs():
    """Overlapping coplanar discs (exact distance ties: the lower index
    wins), tilted and downward-facing discs (back faces are returned), rays
    from above, below and inside the box in every direction."""
    rng = np.random.default_rng(31)
    g = np.arange(-0.25, 0.2501, 0.1)
    X, Y = np.meshgrid(g, g)
    grid = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    rnd = rng.uniform([-0.45, -0.45, -0.4], [0.45, 0.45, 0.4], size=(40, 3))
    xyz = np.concatenate([grid, rnd])
    nrm = np.concatenate([np.tile([0.0, 0.0, 1.0], (grid.shape[0], 1)), rng.normal(size=(40, 3))])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    rad = np.concatenate([np.full(grid.shape[0], 0.1 * math.sqrt(2) / 2 + 0.01), rng.uniform(0.02, 0.12, 40)])
    sc = {"kind": "discs", "xyzr": np.c_[xyz, rad].astype(np.float32), "normals": nrm.astype(np.float32)}
    box = np.array([-0.5, 0.5, -0.5, 0.5, -0.5, 0.6], dtype=np.float32)
    org, d = _random_rays(rng, 4000, [-0.49, -0.49, -0.45], [0.49, 0.49, 0.55])
    # straight down onto the grid: many exact ties between overlapping discs
    org2 = np.c_[rng.uniform(-0.3, 0.3, (1000, 2)), np.full(1000, 0.5)].astype(np.float32)
    d2 = np.tile(np.array([0.0, 0.0, -1.0], dtype=np.float32), (1000, 1))
    org, d = np.concatenate([org, org2]), np.concatenate([d, d2])
    p, fc, t = oracle.closest_hits(sc, box, org, d, tnear=1e-4)
    bp, bf, bt = _brute_closest(sc, box, org, d, 1e-4)
    assert np.array_equal(p, bp)
    assert np.array_equal(fc, bf)
    assert np.array_equal(t, bt)
    # the cases the test means to cover occur
    hitn = sc["normals"][p[p >= 0]]
    back = np.einsum("ij,ij->i", d[p >= 0], hitn) > 0
    assert back.sum() > 100 and (~back).sum() > 100
    assert (fc >= 0).sum() > 100
    c = sc["xyzr"][:, :3]
    exact_ties = 0
    for i in np.nonzero(p[-1000:] >= 0)[0] + 4000:
        hit_z = org[i, 2] + d[i, 2] * t[i]
        on = np.hypot(c[: grid.shape[0], 0] - org[i, 0], c[: grid.shape[0], 1] - org[i, 1]) < rad[0]
        exact_ties += int(on.sum() > 1 and abs(hit_z) < 1e-6)
    assert exact_ties > 100


def test_closest_hit_matches_brute_force_triangles():
    """Random triangles, a duplicated triangle (exact ties), triangles that
    share an edge, rays from everywhere in every direction."""
    rng = np.random.default_rng(32)
    v0 = rng.uniform(-0.45, 0.45, size=(60, 3))
    e1 = rng.normal(scale=0.15, size=(60, 3))
    e2 = rng.normal(scale=0.15, size=(60, 3))
    verts = np.concatenate([v0, v0 + e1, v0 + e2])
    tris = np.stack([np.arange(60), np.arange(60) + 60, np.arange(60) + 120], 1)
    # a horizontal square of two triangles sharing an edge, and a copy of the first
    sq = np.array([[-0.3, -0.3, 0.05], [0.3, -0.3, 0.05], [0.3, 0.3, 0.05], [-0.3, 0.3, 0.05]])
    base = verts.shape[0]
    verts = np.concatenate([verts, sq])
    tris = np.concatenate([tris, [[base, base + 1, base + 2], [base, base + 2, base + 3], [base, base + 1, base + 2]]])
    sc = {"kind": "triangles", "verts": verts.astype(np.float32), "tris": tris.astype(np.uint32)}
    box = np.array([-0.5, 0.5, -0.5, 0.5, -0.6, 0.6], dtype=np.float32)
    org, d = _random_rays(rng, 4000, [-0.49, -0.49, -0.55], [0.49, 0.49, 0.55])
    org2 = np.c_[rng.uniform(-0.29, 0.29, (600, 2)), np.full(600, 0.5)].astype(np.float32)
    d2 = np.tile(np.array([0.0, 0.0, -1.0], dtype=np.float32), (600, 1))
    org, d = np.concatenate([org, org2]), np.concatenate([d, d2])
    p, fc, t = oracle.closest_hits(sc, box, org, d, tnear=1e-4)
    bp, bf, bt = _brute_closest(sc, box, org, d, 1e-4)
    assert np.array_equal(p, bp)
    assert np.array_equal(fc, bf)
    assert np.array_equal(t, bt)
    # rays through the duplicated triangle take the lower index, never the copy
    assert np.sum(p == 60) > 50 and np.sum(p == 62) == 0
    assert (p >= 0).sum() > 1000 and (fc >= 0).sum() > 500


def test_closest_hit_matches_closed_form():
    """The FP32 disc and triangle distances against the exact ray-plane
    intersection in FP64."""
    rng = np.random.default_rng(33)
    for _ in range(500):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        c = rng.uniform(-0.2, 0.2, 3)
        o = c + rng.normal(size=3) * 0.3 + n * 0.5
        tgt = c + np.cross(n, rng.normal(size=3)) * 0.01
        d = (tgt - o) / np.linalg.norm(tgt - o)
        sc = {"kind": "discs", "xyzr": np.array([[*c, 0.2]], dtype=np.float32),
              "normals": n[None].astype(np.float32)}
        box = np.array([-5, 5, -5, 5, -5, 5], dtype=np.float32)
        p, _, t = oracle.closest_hits(sc, box, o[None], d[None], tnear=1e-4)
        o64, d64, c64, n64 = (x.astype(np.float32).astype(np.float64) for x in (o, d, c, n))
        want = (c64 - o64) @ n64 / (d64 @ n64)
        assert p[0] == 0
        assert float(t[0]) == pytest.approx(want, rel=2e-5)


# ---------------------------------------------------------------------------
# Relative error and variance of the variance against the compiled reference
# hit_accumulator (src/rti/trace/hit_accumulator.hpp:128-249).
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed,n_prims,n_dep", [(41, 7, 50), (42, 40, 5000), (43, 3, 200000)])
def test_error_stats_match_compiled_reference(seed, n_prims, n_dep):
    L = _ref()
    rng = np.random.default_rng(seed)
    ids = rng.integers(0, n_prims, n_dep).astype(np.uint32)
    ids[ids == n_prims - 1] = 0  # one primitive without deposits: undefined statistics
    vals = (rng.random(n_dep) * 0.9 + 0.05).astype(np.float32)
    s1 = np.zeros(n_prims)
    cnt = np.zeros(n_prims, dtype=np.uint64)
    rel_ref = np.zeros(n_prims)
    vov_ref = np.zeros(n_prims)
    L.ref_hit_accumulator(n_prims, ids.ctypes.data_as(_u32p), vals.ctypes.data_as(_fp), n_dep,
                          s1.ctypes.data_as(oracle._dp), cnt.ctypes.data_as(_u64p), rel_ref.ctypes.data_as(oracle._dp),
                          vov_ref.ctypes.data_as(oracle._dp))
    v = vals.astype(np.float64)
    sums = [np.bincount(ids, weights=v ** k, minlength=n_prims) for k in (1, 2, 3, 4)]
    mycnt = np.bincount(ids, minlength=n_prims).astype(np.uint64)
    assert np.array_equal(mycnt, cnt)
    np.testing.assert_allclose(sums[0], s1, rtol=1e-12)
    rel, vov = oracle.error_stats(sums[0], sums[1], mycnt, sums[2], sums[3], total=n_dep)
    big = np.finfo(np.float64).max
    empty = mycnt == 0
    assert np.all(rel[empty] == big) and np.all(rel_ref[empty] == big)
    assert np.all(vov[empty] == big) and np.all(vov_ref[empty] == big)
    np.testing.assert_allclose(rel[~empty], rel_ref[~empty], rtol=1e-9)
    np.testing.assert_allclose(vov[~empty], vov_ref[~empty], rtol=1e-6, atol=1e-12)
