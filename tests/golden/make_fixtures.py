"""Writes the golden fixtures under tests/golden/ (run in the development
container, where the reference tree is mounted; the GPU box only reads the
JSON files).

* local_intersector_kats.json: the nine known-answer cases of the reference's
  neighbour-disc test (test/rti/trace/local_intersector.cpp:14-129): ray
  origin, direction (normalised in FP32 as the test does), disc (x, y, z, r),
  disc normal (normalised), expected hit.
* area_inside_kats.json: the known-answer cases of the exposed-area function
  (test/rti/geo/disc_bounding_box_intersector.cpp:6-406): box, disc, normal,
  expected area and the tolerance the reference test uses (0 = ASSERT_EQ on
  the FP32 value). Closed-form expectations are evaluated here with the
  formulas the reference test states.
* radiosity_ar45.json: the reference's radiosity verification profiles for an
  aspect-45 cylinder, radius 0.5, depth 45, s in {1, 0.1, 0.01}
  (verification/radiosity/analytical_solution/original/*.txt; rti is
  BSD-3-licensed, attribution kept in the file).
* radiosity_hole_bands.json: the FP64 radiosity solution (oracle/radiosity.py,
  itself pinned by radiosity_ar45.json) of the benchmark's hole, radius 0.5,
  depth 5, for s in {0.01, 0.1, 1} x n in {1, 10}: the incident flux on 20
  wall bands of depth 0.25, averaged over the benchmark mesh's wall
  triangles as the benchmark averages its full-size fluxes. Needs no
  reference tree (``python tests/golden/make_fixtures.py radiosity_hole_bands.json``).

* historical_mc_ar45.json: the reference's own recorded Monte Carlo profiles
  of the aspect-45 cylinder (verification/radiosity/data/sp.{1,0.1.v2,0.01}.data),
  shape only: 100 depth bins normalised to the first.

The inputs are transcribed from the reference test sources; the expected
values are those the reference asserts.
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def f32n(v):
    v = np.asarray(v, dtype=np.float32)
    return (v / np.float32(np.sqrt((v * v).sum(dtype=np.float32)))).astype(np.float32).tolist()


def local_intersector():
    cases = [
        ("simple_hit_front", [0, 0, 0], [0, 1, 0], [1, 10, 0, 2], [0, -1, 0], True),
        ("simple_hit_back", [0, 0, 0], [0, 1, 0], [1, 10, 0, 2], [0, 1, 0], False),
        ("simple_miss", [0, 0, 0], [0, 1, 1], [1, 10, 0, 2], [0, -1, 0], False),
        ("simple_parallel_miss", [0, 0, 0], [0, 1, 0], [1, 10, 0, 2], [-1, 0, 0], False),
        ("almost_parallel_miss", [0, 0, 0], [0, 1, 0], [1, 10, 0, 2], [-1, -0.01, 0], False),
        ("some_hit", [2, 10, 1.5], [0.133, -1, 0.05], [3, 5, 2, 2], [-1, 1.5, 0.5], True),
        ("miss_by_shooting_away", [2, 10, 1.5], [-0.2, 1, -0.21], [3, 5, 2, 2], [-1, 1.5, 0.5], False),
        ("tight_miss", [2, 10, 1.5], [-0.125, -1, 0.05], [3, 5, 2, 2], [-1, 1.5, 0.5], False),
        ("miss_by_starting_at_the_wrong_side_and_shooting_away", [1, 1, 1], [0, -1, 0], [3, 10, 2, 2],
         [0, 1, 0], False),
    ]
    out = []
    for name, o, d, disc, n, hit in cases:
        out.append({"name": name, "org": o, "dir": f32n(d), "disc": disc, "normal": f32n(n), "hit": hit})
    return {"source": "test/rti/trace/local_intersector.cpp:14-129", "cases": out}


def area_inside():
    pi = math.pi
    pif = float(np.float32(pi))
    b1 = [-1, -1, 1, 1]
    b2 = [-2, -4, 8, 1]
    b3 = [-100, -100, 0, 0]
    cases = []

    def add(name, box, disc, normal, expected, tol, line):
        cases.append({"name": name, "box": box, "disc": disc, "normal": normal, "expected": expected, "tol": tol,
                      "line": line})

    f = lambda x: float(np.float32(x))
    r1 = 1.0
    full1 = f(np.float32(1.0) * np.float32(1.0) * np.float32(pif))
    half1 = f(np.float32(0.5) * np.float32(1.0) * np.float32(1.0) * np.float32(pif))
    add("disc_fully_inside_r_1", b1, [0, 0, 0, 1], [0, 0, 1], f(np.float32(pif)), 0, 34)
    add("disc_fully_inside_r_0_5", b1, [0.2, 0.2, 0, 0.5], [0, 0, 1],
        f(np.float32(0.5) * np.float32(0.5) * np.float32(pif)), 0, 42)
    add("disc_fully_outside_x_direction", b1, [2, 0, 0, 1], [0, 0, 1], 0.0, 0, 50)
    add("disc_fully_outside_y_direction", b1, [0, 2, 0, 1], [0, 0, 1], 0.0, 0, 58)
    add("disc_inside_parallel_to_bounding_box_x_direction", b1, [0.5, 0, 0, 1], [1, 0, 0], full1, 0, 66)
    add("disc_outside_parallel_to_bounding_box_x_direction", b1, [1.5, 0, 0, 1], [1, 0, 0], 0.0, 0, 74)
    add("disc_fully_inside_tilted_close_to_boundary", b1, [0.2, 0, 0, 1], [1, 0, 1], full1, 0, 82)
    add("disc_half_inside_x_direction_right", b1, [1, 0, 0, 1], [0, 0, 1], half1, 0, 90)

    def seg_case(xc):
        a = 2 * math.acos((1.0 - float(np.float32(xc))) / r1)
        return r1 * r1 * pi - r1 * r1 / 2 * (a - math.sin(a))
    add("disc_a_bit_outside_x_direction_right", b1, [0.5, 0, 0, 1], [0, 0, 1], seg_case(0.5), 1e-7, 98)
    add("disc_almost_entirely_outside_x_direction_right", b1, [1.9, 0, 0, 1], [0, 0, 1], seg_case(1.9), 1e-6, 110)
    add("disc_half_inside_x_direction_left", b1, [-1, 0, 0, 1], [0, 0, 1], half1, 0, 126)
    a = 2 * math.acos((1.0 + float(np.float32(-1.9))) / r1)
    add("disc_almost_entirely_outside_x_direction_left", b1, [-1.9, 0, 0, 1], [0, 0, 1],
        r1 * r1 * pi - r1 * r1 / 2 * (a - math.sin(a)), 1e-6, 134)
    add("disc_half_inside_y_direction_top", b1, [0, 1, 0, 1], [0, 0, 1], half1, 0, 150)
    add("disc_half_inside_y_direction_bottom", b1, [1, 0, 0, 1], [0, 0, 1], half1, 0, 158)
    add("disc_half_inside_tilted_y_direction_bottom", b1, [1, 0, 0, 1], [0, 1, 1], half1, 0, 166)
    add("disc_inside_parallel_to_bounding_box_y_direction", b1, [0, 0.8, 0, 1], [0, 1, 0], full1, 0, 174)
    # fixture 2: expected computed as 0.5 * r * r * (float) pi in double, ASSERT_EQ against the float
    h2 = 0.5 * 1.0 * 1.0 * float(np.float32(pi))
    add("b2_disc_half_inside_x_direction_right", b2, [8, -1, 0, 1], [0, 1, 1], h2, 0, 183)
    add("b2_disc_fully_inside_x_direction_right", b2, [7.18, -1, 0, 1], [1, 1, 1], full1, 0, 194)
    add("b2_disc_half_inside_x_direction_left", b2, [-2, -1, 0, 1], [0, 0, 1], h2, 0, 202)
    add("b2_disc_half_inside_y_direction_bottom", b2, [3, -4, 0, 1], [0, 0, 1], h2, 0, 210)
    add("b2_disc_half_inside_y_direction_top", b2, [3, 1, 0, 1], [0, 0, 1], h2, 0, 218)
    dist = float(np.float32(0.7))
    a = 2 * math.acos(math.sqrt(dist * dist * 2) / 1.0)
    a = 2 * pi - a
    add("b2_disc_mostly_inside_tilted_y_direction_bottom", b2, [3, -4 + 0.7, 0, 1], [0.6, -1, 0.8],
        1.0 * 1.0 / 2 * (a - math.sin(a)), 1e-6, 226)
    q = f(np.float32(0.25) * np.float32(1.0) * np.float32(1.0) * np.float32(pif))
    add("b2_corner_top_right", b2, [8, 1, 0, 1], [0, 0, 1], q, 1e-6, 244)
    add("b2_corner_bottom_right", b2, [8, -4, 0, 1], [0, 0, 1], q, 1e-6, 254)
    add("b2_corner_bottom_left", b2, [-2, -4, 0, 1], [0, 0, 1], q, 1e-6, 264)
    add("b2_corner_top_left", b2, [-2, 1, 0, 1], [0, 0, 1], q, 1e-6, 274)
    add("b3_corner_most_simple", b3, [0, 0, 0, 1], [0, 0, 1], f(np.float32(pif) / np.float32(4)), 1e-6, 285)
    add("b3_corner_complex", b3, [-1, -4, 0, 8], [-1, -1, 1], 94.139436, 2e-3, 296)
    add("b4_top_right_corner_1", b2, [9.35, 1.5, 0, 2], [0, 0, 1], 0.349293, 1e-3, 311)
    add("b4_top_right_corner_2", b2, [9.35, 0.5, 0, 2], [0, 0, 1], 0.978276, 1e-3, 322)
    add("b4_top_right_corner_3", b2, [7.65, 1.5, 0, 2], [0, 0, 1], 2.673503, 1e-3, 333)
    add("b4_top_right_corner_4", b2, [7.65, 0.5, 0, 2], [0, 0, 1], 5.002261, 1e-3, 344)
    add("b4_bottom_right_corner", b2, [7.65, -4.33, 1, 2], [-0.79, 0.43, 0.11], 1.381609, 1e-3, 355)
    add("b4_bottom_left_corner", b2, [-1.82, -4.3, 2, 2], [-0.84, -0.53, -0.21], 0.806725, 1e-3, 365)
    add("b4_top_left_corner", b2, [-2.12, 1.43, 3, 2], [-0.65, 0.47, -0.84], 1.422461, 1e-3, 375)
    add("b4_large_disc", b2, [-12.11, -2.01, 1, 20], [0.23, 0.99, 0.45], 86.310325, 1e-2, 386)
    add("b4_huge_disc", b2, [-12.11, -2.01, 1, 30], [0.23, 0.99, 0.45], 123.516791, 2e-2, 397)
    return {"source": "test/rti/geo/disc_bounding_box_intersector.cpp:6-406",
            "box_layout": "xmin, ymin, xmax, ymax", "cases": cases}


def radiosity():
    base = os.path.join(REF, "verification/radiosity/analytical_solution/original")
    files = {"1": ("position_sp_1_ar_45.txt", "flux_sp_1_ar_45.txt"),
             "0.1": ("position.sp.0.1.ar.45.txt", "flux.sp.0.1.ar.45.txt"),
             "0.01": ("position.sp.0.01.ar.45.txt", "flux.sp.0.01.ar.45.txt")}
    out = {"source": "verification/radiosity/analytical_solution/original/*.txt of alexander-at-github/rti "
                     "(BSD-3-Clause, Copyright the rti authors)",
           "geometry": "cylinder radius 0.5, depth 45 (resources/gmsh-based/cylinder/c45.geo:2)",
           "quantity": "incident flux on wall rings relative to the source flux density", "profiles": {}}
    for s, (pf, ff) in files.items():
        pos = [float(x) for x in open(os.path.join(base, pf)).read().split()]
        flux = [float(x) for x in open(os.path.join(base, ff)).read().split()]
        out["profiles"][s] = {"position": pos, "flux": flux}
    return out


def hole_bands():
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from oracle import radiosity as rad
    from rti_amd import geometry
    sc = geometry.hole_triangles()
    edges = np.linspace(0.0, 5.0, 21)
    out = {"source": "oracle/radiosity.py hole_mesh_band_flux (FP64 radiosity of the cylinder, 2000 wall "
                     "elements and 20 floor annuli, cos^n source, integrated over the mesh's wall triangles)",
           "geometry": "rti_amd.geometry.hole_triangles(): radius 0.5, depth 5, 320 x 256 wall facets",
           "quantity": "area-weighted mean incident flux over the wall triangles whose centroid depth lies in "
                       "each band, relative to the source flux density",
           "band_edges_depth": edges.tolist(), "cases": []}
    for s in (0.01, 0.1, 1.0):
        for n in (1.0, 10.0):
            flux = rad.hole_mesh_band_flux(sc["verts"], sc["tris"], 0.5, 5.0, s, n, edges)
            out["cases"].append({"sticking": s, "cos_exponent": n, "wall_band_flux": flux.tolist()})
    return out


def historical_mc():
    """The reference's recorded Monte Carlo profiles of the aspect-45 cylinder
    (verification/radiosity/data/sp.1.data, sp.0.1.v2.data, sp.0.01.data,
    written by an older rti with triangle_geometry_from_gmsh on
    resources/cylinder/c45.msh): 100 depth bins (x = depth / 45), flux
    normalised to the first bin, and the run lines (rays, hits, non-hits)."""
    import re
    base = os.path.join(REF, "verification", "radiosity", "data")
    num = re.compile(r"^\s*[0-9.eE+-]+\s+[0-9.eE+-]+\s*$")
    run = re.compile(r"(\d+)rays (\d+)hits (\d+)nonhits")
    out = {"source": "rti verification/radiosity/data (BSD-3-Clause, Copyright (c) the rti authors)",
           "geometry": "cylinder radius 0.5, depth 45 (resources/gmsh-based/cylinder/c45.geo:2), triangle mesh",
           "quantity": "mean flux per depth bin of 0.45, divided by the first bin's", "profiles": {}}
    for s, name in (("1", "sp.1.data"), ("0.1", "sp.0.1.v2.data"), ("0.01", "sp.0.01.data")):
        text = open(os.path.join(base, name)).read().splitlines()
        rows = [ln.split() for ln in text if num.match(ln)]
        m = next(run.search(ln) for ln in text if run.search(ln))
        out["profiles"][s] = {"file": name, "x": [float(r[0]) for r in rows], "flux": [float(r[1]) for r in rows],
                              "rays": int(m.group(1)), "hits": int(m.group(2)), "nonhits": int(m.group(3))}
    return out


if __name__ == "__main__":
    import sys
    jobs = [("local_intersector_kats.json", local_intersector), ("area_inside_kats.json", area_inside),
            ("radiosity_ar45.json", radiosity), ("radiosity_hole_bands.json", hole_bands),
            ("historical_mc_ar45.json", historical_mc)]
    if len(sys.argv) > 1:
        jobs = [j for j in jobs if j[0] in sys.argv[1:]]
    for name, fn in jobs:
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump(fn(), fh, indent=1)
        print("wrote", name)
