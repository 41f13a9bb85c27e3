"""The disc source (rti_params.source = RTI_SOURCE_DISC): disc_origin_z
(src/rti/ray/disc_origin_z.hpp:17-33) on the disc that
create_circular_source_from_bounding_box (src/rti/device.hpp:269-280) places
on the box: centre of the box, radius half its x extent. The rectangle
source stays the default (what device::run uses, device.hpp:93)."""
import numpy as np
import pytest

from rti_amd import Context, RtiError, geometry, run_scene
from tests.helpers import check_z, footprint_check, oracle_run, small_hole_mesh, small_trench, tie_rank, zscores

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def test_disc_source_footprint_on_flat_mesh():
    sc = geometry.flat_plate_triangles(size=1.0, cells=32)
    n = 4_000_000
    _say("disc source footprint on a flat mesh")
    g = run_scene(sc, n_rays=n, sticking=1.0, bc_x=1, bc_y=1, cos_exponent=1e4, src_z=0.01, source=1)
    assert g["stats"]["surface_hits"] == n
    footprint_check(g["counts"].astype(np.float64), sc, n, (0.5, 0.5), 0.5)


@pytest.mark.parametrize("s", [1.0, 0.1])
def test_disc_source_common_random_numbers(s):
    """Same Philox streams as the oracle, including the redraws of rejected
    offsets, and the library's tie order: per-element sums far inside the MC
    error, hit counts equal for nearly all elements. On this trench (box 2
    wide in x, 0.5 in y) most origins lie outside the box in y, so the
    boundary faces are also met from outside."""
    sc = small_trench()
    n = 200_000
    _say(f"disc source, common random numbers, s={s}")
    g = run_scene(sc, n_rays=n, sticking=s, bc_x=0, bc_y=1, seed=1, source=1)
    o = oracle_run(sc, n, s, bc_x=0, bc_y=1, source=1, tie_rank=tie_rank(g["perm"]))
    scale = np.maximum(np.sqrt(g["sigma"] ** 2 + o["s2"] - o["s1"] ** 2 / n), 1e-12)
    d = np.abs(g["raw"] - o["s1"]) / scale
    assert np.quantile(d, 0.99) < 0.5, np.quantile(d, [0.5, 0.9, 0.99, 1.0])
    same = np.mean(g["counts"] == o["cnt"])
    assert same > (0.995 if s == 1.0 else 0.9), same
    assert g["stats"]["segments"] == o["segments"] or s != 1.0
    # the disc source lights a different part of the trench than the rectangle
    r = run_scene(sc, n_rays=n, sticking=s, bc_x=0, bc_y=1, seed=1)
    assert not np.array_equal(r["counts"], g["counts"])


def test_disc_source_independent_seeds_within_3sigma():
    sc = small_hole_mesh()
    n = 400_000
    _say("disc source on the hole mesh vs oracle, independent seeds")
    g = run_scene(sc, n_rays=n, sticking=0.1, bc_x=1, bc_y=1, seed=3, source=1)
    o = oracle_run(sc, n, 0.1, seed=99, bc_x=1, bc_y=1, source=1)
    so = np.sqrt(np.maximum(o["s2"] - o["s1"] ** 2 / n, 0))
    z = zscores(g["raw"], g["sigma"], o["s1"], so, min_count=30, ca=g["counts"], cb=o["cnt"])
    check_z(z)


def test_batch_with_both_sources_equals_single_runs():
    items = [
        dict(scene=small_trench(), n_rays=100_000, sticking=0.1, bc_x=0, bc_y=1, seed=4, source=1),
        dict(scene=small_trench(), n_rays=100_000, sticking=0.1, bc_x=0, bc_y=1, seed=4),
        dict(scene=small_hole_mesh(), n_rays=80_000, sticking=0.05, cos_exponent=5.0, bc_x=1, bc_y=1, seed=4,
             source=1),
    ]
    _say("batch mixing disc and rectangle sources")
    with Context() as ctx:
        ctx.batch_set(items)
        ctx.batch_run()
        got = [ctx.batch_flux(i)["raw"] for i in range(len(items))]
    for k, it in enumerate(items):
        kw = {key: v for key, v in it.items() if key != "scene"}
        single = run_scene(it["scene"], instance=k, **kw)
        assert np.array_equal(got[k], single["raw"]), k


def test_unknown_source_is_rejected():
    with Context() as ctx:
        ctx.set_scene(small_trench())
        with pytest.raises(RtiError, match="source"):
            ctx.set_params(source=2)
