"""Per-primitive relative error and variance of the variance on the GPU
(the reference's hit_accumulator::get_relative_error / get_vov,
src/rti/trace/hit_accumulator.hpp:152-249, as written to its VTP output).

* s = 1: every deposit is exactly 1, so S1 = S2 = S3 = S4 = count and both
  statistics have closed forms in the counts -- checked to 1e-12;
* s = 0.1: the GPU statistics against the oracle's on the same seed (common
  random numbers: the same paths up to rare floating-point forks);
* a batch gives each instance the statistics of its single run;
* the variance of the variance needs the moment sums (moments = 1).
"""
import numpy as np
import pytest

import oracle
from rti_amd import Context, run_scene
from tests.helpers import oracle_run, small_trench

pytestmark = pytest.mark.gpu
BIG = np.finfo(np.float64).max


def test_closed_form_with_unit_deposits():
    sc = small_trench()
    print("[gpu] error stats s=1", flush=True)
    g = run_scene(sc, n_rays=200_000, sticking=1.0, bc_x=0, bc_y=1, seed=3, moments=1)
    c = g["counts"].astype(np.float64)
    N = c.sum()
    assert np.array_equal(g["raw"], c)  # unit deposits: S1 = count
    m = c > 0
    assert np.all(g["rel_err"][~m] == BIG) and np.all(g["vov"][~m] == BIG)
    np.testing.assert_allclose(g["rel_err"][m], np.sqrt(1.0 / c[m] - 1.0 / N), rtol=1e-12)
    num = c - 4 * c * c / N + 8 * c * c * c / N ** 2 - 4 * c ** 4 / N ** 3 - c * c / N
    den = (N * c - c * c) ** 2
    np.testing.assert_allclose(g["vov"][m], num[m] / den[m] * N * N, rtol=1e-9)


def test_matches_oracle_on_common_random_numbers():
    sc = small_trench()
    print("[gpu] error stats vs oracle", flush=True)
    kw = dict(n_rays=100_000, sticking=0.1, bc_x=0, bc_y=1, seed=21)
    g = run_scene(sc, moments=1, **kw)
    o = oracle_run(sc, moments=True, **kw)
    rel, vov = oracle.error_stats(o["s1"], o["s2"], o["cnt"], o["s3"], o["s4"])
    m = (g["counts"] >= 50) & (o["cnt"] >= 50)
    assert m.sum() > 300
    rd = np.abs(g["rel_err"][m] - rel[m]) / rel[m]
    vd = np.abs(g["vov"][m] - vov[m]) / np.abs(vov[m])
    # common random numbers: most elements agree closely, a few paths fork
    assert np.median(rd) < 1e-3 and np.quantile(rd, 0.95) < 0.05, (np.median(rd), np.quantile(rd, 0.95))
    assert np.median(vd) < 1e-2 and np.quantile(vd, 0.9) < 0.2, (np.median(vd), np.quantile(vd, 0.9))


def test_batch_instances_have_single_run_statistics():
    sc = small_trench()
    print("[gpu] error stats batch", flush=True)
    items = [dict(scene=sc, n_rays=40_000, sticking=s, bc_x=0, bc_y=1, seed=5) for s in (0.05, 0.5)]
    with Context() as ctx:
        ctx.batch_set(items)
        ctx.set_params(moments=1)
        ctx.batch_run()
        got = [ctx.batch_error_stats(k) for k in range(2)]
    for k, it in enumerate(items):
        single = run_scene(sc, instance=k, moments=1, **{q: v for q, v in it.items() if q != "scene"})
        assert np.array_equal(got[k][0], single["rel_err"])  # from the fixed-point sums: bit-exact
        np.testing.assert_allclose(got[k][1], single["vov"], rtol=1e-9)  # FP64 atomics: order-dependent bits


def test_vov_needs_moments():
    sc = small_trench()
    with Context() as ctx:
        ctx.set_scene(sc)
        ctx.set_params(n_rays=1000, sticking=0.5)
        ctx.run()
        rel, _ = ctx.error_stats(vov=False)
        assert np.isfinite(rel).any()
        with pytest.raises(Exception):
            ctx.error_stats(vov=True)
