"""rti_trace_noalloc (the trace entry point that allocates no device memory)
against rti_trace + rti_finalize on the same context: bit-identical sums and
counts, on two calls in a row, the second with a smaller scene and fewer
rays (the static scratch keeps the first call's contents).

The caller's accumulator buffer is allocated here through the HIP runtime
the library itself is linked against (the test is the caller; the entry
point allocates nothing)."""
import ctypes

import numpy as np
import pytest

from rti_amd import _lib, geometry
from rti_amd._lib import Context

pytestmark = pytest.mark.gpu

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        _lib.lib()  # loads libamdhip64 through librti_hip.so first
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _HIP = h
    return _HIP


HIP_D2H = 2


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        # garbage that the entry point must clear
        assert hip().hipMemset(self.ptr, 0x5A, self.nbytes) == 0

    def to_host_u64(self):
        out = np.zeros(self.nbytes // 8, dtype=np.uint64)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, HIP_D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def noalloc_trace(ctx, buf, words):
    rc = _lib.lib().rti_trace_noalloc(ctx._h, buf.ptr, int(words))
    ctx._check(rc)


def check_against_trace(scene, n_rays, params, buf):
    """rti_trace + rti_finalize, then rti_trace_noalloc into `buf` on the
    same built context: sums and counts must agree bit for bit."""
    n = scene["xyzr"].shape[0]
    with Context(0) as ctx:
        ctx.set_scene(scene)
        ctx.set_params(n_rays=n_rays, **params)
        ctx.run()
        ref = ctx.flux()
        ref_cnt = ctx.counts()
        ref_counters = ctx.counters().copy()
        shift = ctx.stats()["fixed_shift"]
        perm = ctx.perm()
        assert hip().hipMemset(buf.ptr, 0x5A, buf.nbytes) == 0
        noalloc_trace(ctx, buf, buf.nbytes // 8)
        counters = ctx.counters().copy()
    acc = buf.to_host_u64()
    assert np.all(acc[4 * n:] == np.uint64(0x5A5A5A5A5A5A5A5A)), "wrote past the scene's records"
    rec = acc[: 4 * n].reshape(n, 4)
    raw = np.zeros(n)
    cnt = np.zeros(n, dtype=np.uint64)
    raw[perm] = rec[:, 0].astype(np.int64).astype(np.float64) * 2.0 ** -shift
    cnt[perm] = rec[:, 2]
    assert np.all(rec[:, 3] == 0)
    np.testing.assert_array_equal(raw, ref["raw"])
    np.testing.assert_array_equal(cnt, ref_cnt)
    np.testing.assert_array_equal(counters[:4], ref_counters[:4])
    assert ref_counters[0] > n_rays


def test_noalloc_matches_trace_twice_second_smaller():
    big = geometry.trench_discs(width=1.0, depth=2.0, flat=0.5, ylen=0.5, spacing=0.04)
    small = geometry.trench_discs(width=1.0, depth=1.0, flat=0.25, ylen=0.25, spacing=0.05)
    nb, ns = big["xyzr"].shape[0], small["xyzr"].shape[0]
    assert ns < nb
    print(f"[no_alloc] big trench {nb} discs, small trench {ns} discs", flush=True)
    params = dict(sticking=0.1, bc_x=0, bc_y=1, seed=7)
    buf = DeviceBuffer(8 * 4 * nb + 4096)
    try:
        check_against_trace(big, 300_000, params, buf)
        print("[no_alloc] first call ok", flush=True)
        check_against_trace(small, 100_000, dict(params, sticking=0.3, seed=8), buf)
        print("[no_alloc] second call ok", flush=True)
    finally:
        buf.free()


def test_noalloc_rejects():
    sc = geometry.trench_discs(width=1.0, depth=1.0, flat=0.25, ylen=0.25, spacing=0.05)
    n = sc["xyzr"].shape[0]
    buf = DeviceBuffer(8 * 4 * n)
    try:
        with Context(0) as ctx:
            ctx.set_scene(sc)
            ctx.set_params(n_rays=1000)
            # not built yet
            assert _lib.lib().rti_trace_noalloc(ctx._h, buf.ptr, 4 * n) == _lib.RTI_ERR_STATE
            ctx.prepare()
            assert _lib.lib().rti_trace_noalloc(ctx._h, buf.ptr, 4 * n - 1) == _lib.RTI_ERR_INVALID_ARG
            assert b"needs" in _lib.lib().rti_last_error(ctx._h)
            assert _lib.lib().rti_trace_noalloc(ctx._h, None, 4 * n) == _lib.RTI_ERR_INVALID_ARG
            ctx.set_params(n_rays=1000, adjoint=1)
            assert _lib.lib().rti_trace_noalloc(ctx._h, buf.ptr, 4 * n) == _lib.RTI_ERR_INVALID_ARG
            ctx.set_params(n_rays=1000)
            assert _lib.lib().rti_trace_noalloc(ctx._h, buf.ptr, 4 * n) == _lib.RTI_OK
            # the context's own accumulators were not written: finalize refuses
            assert _lib.lib().rti_finalize(ctx._h) == _lib.RTI_ERR_STATE
        # more instances than the static scratch holds
        items = [dict(scene=sc, n_rays=10, sticking=0.5)] * 257
        with Context(0) as ctx:
            ctx.batch_set(items)
            ctx.prepare()
            words = 4 * n * 257
            big = DeviceBuffer(8 * words)
            try:
                assert _lib.lib().rti_trace_noalloc(ctx._h, big.ptr, words) == _lib.RTI_ERR_INVALID_ARG
                assert b"at most 256" in _lib.lib().rti_last_error(ctx._h)
            finally:
                big.free()
    finally:
        buf.free()
