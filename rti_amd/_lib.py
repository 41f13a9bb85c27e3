"""ctypes binding of librti_hip.so (the C ABI in include/rti/rti_c.h).

The library is required: importing works without it, but every call that
needs it raises ``RuntimeError`` -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librti_hip.so")

RTI_OK = 0
RTI_ERR_INVALID_ARG = 1
RTI_ERR_NO_GEOMETRY = 2
RTI_ERR_HIP = 3
RTI_ERR_OVERFLOW = 4
RTI_ERR_STATE = 5
RTI_ERR_ALLOC = 6

BC_REFLECTIVE = 0
BC_PERIODIC = 1
KERNEL_PERSISTENT = 0
KERNEL_PLAIN = 1
KERNEL_COUNT = 2
ACCUM_FIXED64 = 0
ACCUM_FP64 = 1
SOURCE_RECTANGLE = 0
SOURCE_DISC = 1
TREE_LBVH = 1
TREE_PLOC = 2

# Every symbol the header declares; tests check the library exports them all.
EXPORTS = [
    "rti_default_params", "rti_ctx_create", "rti_ctx_destroy", "rti_last_error", "rti_set_discs",
    "rti_set_triangles", "rti_set_params", "rti_run", "rti_prepare", "rti_trace", "rti_finalize", "rti_get_flux",
    "rti_get_counts", "rti_get_areas", "rti_get_neighbors", "rti_get_stats", "rti_get_adjoint", "rti_get_counters",
    "rti_get_hierarchy", "rti_get_wide_hierarchy",
    "rti_batch_set", "rti_batch_run", "rti_batch_replay", "rti_batch_get_flux", "rti_batch_get_counts",
    "rti_batch_get_stats", "rti_get_error_stats", "rti_batch_get_error_stats", "rti_batch_get_adjoint",
    "rti_trace_noalloc", "rti_device_info",
]


class RtiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rti error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("n_rays", ctypes.c_uint64), ("seed", ctypes.c_uint32), ("instance", ctypes.c_uint32),
        ("bc_x", ctypes.c_int32), ("bc_y", ctypes.c_int32), ("sticking", ctypes.c_float),
        ("cos_exponent", ctypes.c_float), ("diffuse_prob", ctypes.c_float), ("tnear", ctypes.c_float),
        ("roulette", ctypes.c_int32), ("src_z", ctypes.c_float), ("kernel", ctypes.c_int32),
        ("adjoint", ctypes.c_int32), ("accum", ctypes.c_int32), ("block_size", ctypes.c_int32),
        ("lds_stack", ctypes.c_int32), ("waves_per_cu", ctypes.c_int32), ("compact_lanes", ctypes.c_int32),
        ("bvh_width", ctypes.c_int32), ("moments", ctypes.c_int32), ("source", ctypes.c_int32),
        ("leaf_batch", ctypes.c_int32), ("tree_builder", ctypes.c_int32), ("refill_lanes", ctypes.c_int32),
        ("ray_begin", ctypes.c_uint64), ("rays_total", ctypes.c_uint64),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("n_prims", ctypes.c_uint64), ("n_rays", ctypes.c_uint64), ("segments", ctypes.c_uint64),
        ("surface_hits", ctypes.c_uint64), ("misses", ctypes.c_uint64), ("truncated", ctypes.c_uint64),
        ("n_edges", ctypes.c_uint64), ("bvh_depth", ctypes.c_uint32), ("fixed_shift", ctypes.c_int32),
        ("trace_ms", ctypes.c_double), ("build_ms", ctypes.c_double), ("post_ms", ctypes.c_double),
        ("box", ctypes.c_float * 6), ("src_z", ctypes.c_float), ("wide_stack", ctypes.c_uint32),
        ("wide_nodes", ctypes.c_uint64), ("fixed_shift2", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("max_segments", ctypes.c_uint64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "box"}
        d["box"] = list(self.box)
        return d


# Four-wide traversal-tree record (64 bytes): child c's box planes are
# fmaf(byte c of the plane word, |s|, org) with s the axis's power-of-two
# scale; the signs of s_y and s_z hold nchild - 1 (bits 0 and 1).
WIDE_NODE = np.dtype([("org", "<f4", 3), ("meta", "<u4"), ("lo", "<u4", 3), ("hi", "<u4", 3), ("int_base", "<u4"),
                      ("leaf_info", "<u4"), ("leaf", "<u4", 4)])


def wide_scales(node):
    """(s_x, s_y, s_z) = 2^(e - 127) and the child count of a four-wide record."""
    meta = int(node["meta"])
    s = np.array([np.ldexp(1.0, ((meta >> (8 * k)) & 0xFF) - 127) for k in range(3)], dtype=np.float32)
    return s, 1 + ((meta >> 24) & 3)


def wide_children(node):
    """Child references of a four-wide record in slot order: a node index
    (>= 0) for internal children, ~stored position for leaves."""
    meta = int(node["meta"])
    m, nint = 1 + ((meta >> 24) & 3), (meta >> 26) & 7
    return [int(node["int_base"]) + c if c < nint else ~int(node["leaf"][c]) for c in range(m)]

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class Instance(ctypes.Structure):
    _fields_ = [
        ("xyzr", _fp), ("nrm", _fp), ("verts", _fp), ("n_verts", ctypes.c_uint64), ("tris", _u32p),
        ("n_prims", ctypes.c_uint64), ("params", Params),
    ]


_LIB = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rti_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _LIB = L
    return _LIB


def _declare(L):
    c = ctypes
    vp = c.c_void_p
    L.rti_default_params.argtypes = [c.POINTER(Params)]
    L.rti_default_params.restype = None
    L.rti_ctx_create.argtypes = [c.POINTER(vp), c.c_int]
    L.rti_ctx_destroy.argtypes = [vp]
    L.rti_last_error.argtypes = [vp]
    L.rti_last_error.restype = c.c_char_p
    L.rti_set_discs.argtypes = [vp, _fp, _fp, c.c_uint64]
    L.rti_set_triangles.argtypes = [vp, _fp, c.c_uint64, _u32p, c.c_uint64]
    L.rti_set_params.argtypes = [vp, c.POINTER(Params)]
    for name in ("rti_run", "rti_prepare", "rti_trace", "rti_finalize", "rti_batch_run"):
        getattr(L, name).argtypes = [vp]
    L.rti_get_flux.argtypes = [vp, _fp, _dp, _dp]
    L.rti_get_counts.argtypes = [vp, _u64p]
    L.rti_get_areas.argtypes = [vp, _fp]
    L.rti_get_neighbors.argtypes = [vp, _u32p, _u32p, c.c_uint64, _u64p]
    L.rti_get_stats.argtypes = [vp, c.POINTER(Stats)]
    L.rti_get_adjoint.argtypes = [vp, _dp, _dp]
    L.rti_get_counters.argtypes = [vp, _u64p]
    L.rti_get_hierarchy.argtypes = [vp, c.c_void_p, c.c_uint64, _u32p, _u64p, _fp]
    L.rti_get_wide_hierarchy.argtypes = [vp, c.c_void_p, c.c_uint64, _u64p]
    L.rti_batch_set.argtypes = [vp, c.POINTER(Instance), c.c_uint32]
    L.rti_batch_replay.argtypes = [vp, c.c_uint32]
    L.rti_batch_get_flux.argtypes = [vp, c.c_uint32, _fp, _dp, _dp]
    L.rti_batch_get_counts.argtypes = [vp, c.c_uint32, _u64p]
    L.rti_batch_get_stats.argtypes = [vp, c.POINTER(Stats)]
    L.rti_get_error_stats.argtypes = [vp, _dp, _dp]
    L.rti_batch_get_error_stats.argtypes = [vp, c.c_uint32, _dp, _dp]
    L.rti_batch_get_adjoint.argtypes = [vp, c.c_uint32, _dp, _dp]
    L.rti_trace_noalloc.argtypes = [vp, c.c_void_p, c.c_uint64]
    L.rti_device_info.argtypes = [c.c_int, c.c_char_p, c.c_uint64, c.POINTER(c.c_int)]
    for name in EXPORTS:
        f = getattr(L, name)
        if name not in ("rti_default_params", "rti_last_error"):
            f.restype = c.c_int


def device_info(device=0):
    """(architecture name, compute units) of a HIP device."""
    buf = ctypes.create_string_buffer(64)
    cus = ctypes.c_int(0)
    rc = lib().rti_device_info(int(device), buf, 64, ctypes.byref(cus))
    if rc != RTI_OK:
        raise RtiError(rc, "rti_device_info failed (no HIP device?)")
    return buf.value.decode(errors="replace"), int(cus.value)


def default_params(**kw):
    p = Params()
    lib().rti_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Context:
    """One rti_ctx: geometry, parameters, last results."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        rc = lib().rti_ctx_create(ctypes.byref(self._h), int(device))
        if rc != RTI_OK:
            raise RtiError(rc, "rti_ctx_create failed (no HIP device?)")
        self._keep = []
        self.n = 0
        self.params = default_params()

    def close(self):
        if self._h:
            lib().rti_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != RTI_OK:
            raise RtiError(rc, lib().rti_last_error(self._h).decode(errors="replace"))

    # --- inputs ---
    def set_discs(self, xyzr, normals):
        x = np.ascontiguousarray(xyzr, dtype=np.float32).reshape(-1, 4)
        n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        if x.shape[0] != n.shape[0]:
            raise ValueError("xyzr and normals differ in length")
        self._check(lib().rti_set_discs(self._h, _ptr(x, _fp), _ptr(n, _fp), x.shape[0]))
        self.n = x.shape[0]

    def set_triangles(self, verts, tris):
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
        self._check(lib().rti_set_triangles(self._h, _ptr(v, _fp), v.shape[0], _ptr(t, _u32p), t.shape[0]))
        self.n = t.shape[0]

    def set_scene(self, scene):
        if scene["kind"] == "discs":
            self.set_discs(scene["xyzr"], scene["normals"])
        else:
            self.set_triangles(scene["verts"], scene["tris"])

    def set_params(self, **kw):
        p = default_params(**kw)
        self._check(lib().rti_set_params(self._h, ctypes.byref(p)))
        self.params = p

    # --- execution ---
    def run(self):
        self._check(lib().rti_run(self._h))

    def prepare(self):
        self._check(lib().rti_prepare(self._h))

    def trace(self):
        self._check(lib().rti_trace(self._h))

    def finalize(self):
        self._check(lib().rti_finalize(self._h))

    # --- outputs ---
    def flux(self):
        n = self.n
        sm = np.zeros(n, dtype=np.float32)
        raw = np.zeros(n)
        sig = np.zeros(n)
        self._check(lib().rti_get_flux(self._h, _ptr(sm, _fp), _ptr(raw, _dp), _ptr(sig, _dp)))
        return {"smoothed": sm, "raw": raw, "sigma": sig}

    def counts(self):
        c = np.zeros(self.n, dtype=np.uint64)
        self._check(lib().rti_get_counts(self._h, _ptr(c, _u64p)))
        return c

    def areas(self):
        a = np.zeros(self.n, dtype=np.float32)
        self._check(lib().rti_get_areas(self._h, _ptr(a, _fp)))
        return a

    def neighbors(self):
        row = np.zeros(self.n + 1, dtype=np.uint32)
        ne = ctypes.c_uint64(0)
        self._check(lib().rti_get_neighbors(self._h, _ptr(row, _u32p), None, 0, ctypes.byref(ne)))
        col = np.zeros(max(int(ne.value), 1), dtype=np.uint32)
        self._check(lib().rti_get_neighbors(self._h, _ptr(row, _u32p), _ptr(col, _u32p), col.size,
                                            ctypes.byref(ne)))
        return row, col[: int(ne.value)]

    def hierarchy(self):
        """(nodes structured array, perm, prim_a) of the built hierarchy."""
        nn = ctypes.c_uint64(0)
        self._check(lib().rti_get_hierarchy(self._h, None, 0, None, ctypes.byref(nn), None))
        dt = np.dtype([("xy0", "<f4", 4), ("xy1", "<f4", 4), ("z01", "<f4", 4), ("child", "<i4", 4)])
        nodes = np.zeros(int(nn.value), dtype=dt)
        perm = np.zeros(self.n, dtype=np.uint32)
        pa = np.zeros((self.n, 4), dtype=np.float32)
        self._check(lib().rti_get_hierarchy(self._h, nodes.ctypes.data, nodes.size, _ptr(perm, _u32p),
                                            ctypes.byref(nn), _ptr(pa, _fp)))
        return nodes, perm, pa

    def perm(self):
        """Stored position -> input id of the (first) geometry: the library's
        order, which also breaks exact distance ties (lower position wins)."""
        nn = ctypes.c_uint64(0)
        perm = np.zeros(self.n, dtype=np.uint32)
        self._check(lib().rti_get_hierarchy(self._h, None, 0, _ptr(perm, _u32p), ctypes.byref(nn), None))
        return perm

    def wide_hierarchy(self):
        """Structured array of the four-wide traversal tree (root first)."""
        nn = ctypes.c_uint64(0)
        self._check(lib().rti_get_wide_hierarchy(self._h, None, 0, ctypes.byref(nn)))
        nodes = np.zeros(int(nn.value), dtype=WIDE_NODE)
        self._check(lib().rti_get_wide_hierarchy(self._h, nodes.ctypes.data, nodes.size, ctypes.byref(nn)))
        return nodes

    def stats(self):
        s = Stats()
        self._check(lib().rti_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def counters(self):
        c = np.zeros(8, dtype=np.uint64)
        self._check(lib().rti_get_counters(self._h, _ptr(c, _u64p)))
        return c

    def error_stats(self, vov=True):
        """(relative error, variance of the variance or None) per primitive;
        the variance of the variance needs a run with moments=1."""
        rel = np.zeros(self.n)
        v = np.zeros(self.n) if vov else None
        self._check(lib().rti_get_error_stats(self._h, _ptr(rel, _dp), _ptr(v, _dp) if vov else None))
        return rel, v

    def batch_error_stats(self, i, vov=True):
        n = self.batch_n[i]
        rel = np.zeros(n)
        v = np.zeros(n) if vov else None
        self._check(lib().rti_batch_get_error_stats(self._h, i, _ptr(rel, _dp), _ptr(v, _dp) if vov else None))
        return rel, v

    def batch_adjoint(self, i):
        n = self.batch_n[i]
        ds = np.zeros(n)
        dn = np.zeros(n)
        self._check(lib().rti_batch_get_adjoint(self._h, i, _ptr(ds, _dp), _ptr(dn, _dp)))
        return ds, dn

    def adjoint(self):
        ds = np.zeros(self.n)
        dn = np.zeros(self.n)
        self._check(lib().rti_get_adjoint(self._h, _ptr(ds, _dp), _ptr(dn, _dp)))
        return ds, dn

    # --- batch ---
    def batch_set(self, items):
        """items: list of dicts with keys scene, and rti_params fields."""
        arr = (Instance * len(items))()
        keep = []
        converted = {}  # one buffer pair per scene object: items on the same scene share its device copy
        self.batch_n = []
        for k, it in enumerate(items):
            sc = it["scene"]
            kw = {key: val for key, val in it.items() if key not in ("scene",) and hasattr(Params, key)}
            arr[k].params = default_params(**kw)
            if sc["kind"] == "discs":
                if id(sc) not in converted:
                    converted[id(sc)] = (np.ascontiguousarray(sc["xyzr"], dtype=np.float32),
                                         np.ascontiguousarray(sc["normals"], dtype=np.float32))
                x, n = converted[id(sc)]
                keep += [x, n]
                arr[k].xyzr = _ptr(x, _fp)
                arr[k].nrm = _ptr(n, _fp)
                arr[k].n_prims = x.shape[0]
            else:
                if id(sc) not in converted:
                    converted[id(sc)] = (np.ascontiguousarray(sc["verts"], dtype=np.float32),
                                         np.ascontiguousarray(sc["tris"], dtype=np.uint32))
                v, t = converted[id(sc)]
                keep += [v, t]
                arr[k].verts = _ptr(v, _fp)
                arr[k].n_verts = v.shape[0]
                arr[k].tris = _ptr(t, _u32p)
                arr[k].n_prims = t.shape[0]
            self.batch_n.append(int(arr[k].n_prims))
        self._check(lib().rti_batch_set(self._h, arr, len(items)))
        self.n = self.batch_n[0]

    def batch_run(self):
        self._check(lib().rti_batch_run(self._h))

    def batch_replay(self, times=1):
        self._check(lib().rti_batch_replay(self._h, int(times)))

    def batch_flux(self, i):
        n = self.batch_n[i]
        sm = np.zeros(n, dtype=np.float32)
        raw = np.zeros(n)
        sig = np.zeros(n)
        self._check(lib().rti_batch_get_flux(self._h, i, _ptr(sm, _fp), _ptr(raw, _dp), _ptr(sig, _dp)))
        return {"smoothed": sm, "raw": raw, "sigma": sig}

    def batch_counts(self, i):
        c = np.zeros(self.batch_n[i], dtype=np.uint64)
        self._check(lib().rti_batch_get_counts(self._h, i, _ptr(c, _u64p)))
        return c

    def batch_stats(self):
        s = Stats()
        self._check(lib().rti_batch_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()


def run_scene(scene, device=0, **params):
    """Trace one scene and return flux, counts, areas and stats."""
    with Context(device) as ctx:
        ctx.set_scene(scene)
        ctx.set_params(**params)
        ctx.run()
        out = ctx.flux()
        out["counts"] = ctx.counts()
        out["areas"] = ctx.areas()
        out["stats"] = ctx.stats()
        out["counters"] = ctx.counters()
        out["perm"] = ctx.perm()
        if params.get("adjoint"):
            out["ds"], out["dn"] = ctx.adjoint()
        out["rel_err"], out["vov"] = ctx.error_stats(vov=bool(params.get("moments")))
        return out

