"""CPU oracle for the rti Monte Carlo flux path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg import this package, and only as the checker (or as the timed CPU
baseline). The product (``librti_hip.so``, ``include/rti``, ``rti_amd``)
never imports, links or calls it.

``liboracle.so`` is built from ``rti_oracle.cpp`` by ``oracle/Makefile``; each
function there cites the reference file:line it restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def build_ref(reference="/root/reference"):
    """Compile the reference's Embree-free headers into oracle/_ref (only
    where the reference checkout exists)."""
    if os.path.isdir(os.path.join(reference, "src", "rti")):
        subprocess.run(["make", "-s", "-C", _HERE, "ref", f"REF={reference}"], check=True)
        return True
    return False


def ref_lib():
    """The compiled reference headers (oracle/_ref/libref.so) or None."""
    path = os.path.join(_HERE, "_ref", "libref.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    L.ref_area_inside.argtypes = [ctypes.c_float] * 4 + [_fp, _fp]
    L.ref_area_inside.restype = ctypes.c_float
    L.ref_neighbors.argtypes = [_fp, ctypes.c_uint64, _u32p, _u32p]
    L.ref_neighbors.restype = ctypes.c_int64
    L.ref_orthonormal_basis.argtypes = [_fp, _fp]
    L.ref_cos_hemi.argtypes = [_fp, _u64p, _fp]
    L.ref_power_cosine_direction.argtypes = [ctypes.c_float, _u64p, _fp]
    L.ref_rectangle_origin.argtypes = [ctypes.c_float] * 5 + [_u64p, _fp]
    L.ref_rejection_control.argtypes = [_fp, ctypes.c_uint64]
    L.ref_rejection_control.restype = ctypes.c_int
    L.ref_hit_accumulator.argtypes = [ctypes.c_uint64, _u32p, _fp, ctypes.c_uint64, _dp, _u64p, _dp, _dp]
    return L


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
        _declare(_LIB)
    return _LIB


_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class Scene(ctypes.Structure):
    _fields_ = [
        ("prim_kind", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("n_prims", ctypes.c_uint64),
        ("xyzr", _fp), ("nrm", _fp),
        ("verts", _fp), ("n_verts", ctypes.c_uint64), ("tris", _u32p),
        ("nb_ptr", _u32p), ("nb_col", _u32p), ("tie_rank", _u32p),
    ]


class Params(ctypes.Structure):
    _fields_ = [
        ("n_rays", ctypes.c_uint64), ("ray_begin", ctypes.c_uint64),
        ("seed", ctypes.c_uint32), ("instance", ctypes.c_uint32),
        ("bc_x", ctypes.c_int32), ("bc_y", ctypes.c_int32),
        ("sticking", ctypes.c_float), ("cos_exponent", ctypes.c_float),
        ("diffuse_prob", ctypes.c_float), ("tnear", ctypes.c_float),
        ("roulette", ctypes.c_int32), ("fp64", ctypes.c_int32),
        ("box", ctypes.c_float * 6), ("src_z", ctypes.c_float),
        ("threads", ctypes.c_int32), ("adjoint", ctypes.c_int32), ("moments", ctypes.c_int32),
        ("source", ctypes.c_int32),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("s1", _dp), ("s2", _dp), ("cnt", _u64p), ("dsum_ds", _dp), ("dsum_dn", _dp),
        ("segments", ctypes.c_uint64), ("surface_hits", ctypes.c_uint64),
        ("misses", ctypes.c_uint64), ("seconds", ctypes.c_double), ("s3", _dp), ("s4", _dp),
    ]


def _declare(L):
    L.orc_error_stats.argtypes = [_dp, _dp, _dp, _dp, _u64p, ctypes.c_uint64, ctypes.c_uint64, _dp, _dp]
    L.orc_error_stats.restype = None
    L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
    L.orc_local_intersect.argtypes = [_fp, _fp, _fp, _fp]
    L.orc_local_intersect.restype = ctypes.c_int
    L.orc_area_inside.argtypes = [ctypes.c_float] * 4 + [_fp, _fp]
    L.orc_area_inside.restype = ctypes.c_float
    L.orc_neighbors.argtypes = [_fp, ctypes.c_uint64, _u32p, _u32p]
    L.orc_neighbors.restype = ctypes.c_int64
    L.orc_smooth_normalize.argtypes = [_dp, _fp, _u32p, _u32p, ctypes.c_uint64, _fp]
    L.orc_trace.argtypes = [ctypes.POINTER(Scene), ctypes.POINTER(Params), ctypes.POINTER(Result)]
    L.orc_source_dir.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, _fp]
    L.orc_source_origins.argtypes = [ctypes.POINTER(Params), ctypes.c_uint64, ctypes.c_uint64, _fp]
    L.orc_source_origins.restype = None
    L.orc_diffuse_dir.argtypes = [_fp, ctypes.c_float, ctypes.c_float, _fp]
    L.orc_source_rays.argtypes = [ctypes.POINTER(Params), ctypes.c_uint64, ctypes.c_uint64, _fp, _fp]
    L.orc_source_rays.restype = None
    L.orc_specular_dir.argtypes = [_fp, _fp, _fp]
    L.orc_specular_dir.restype = None
    L.orc_closest_hits.argtypes = [ctypes.POINTER(Scene), ctypes.POINTER(Params), ctypes.c_uint64, _fp, _fp,
                                   ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), _fp]
    L.orc_closest_hits.restype = None
    L.orc_roulette.argtypes = [_fp, ctypes.c_float]
    L.orc_roulette.restype = ctypes.c_int
    L.orc_trace.restype = ctypes.c_int


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(c, _u32p), _ptr(k, _u32p), _ptr(out, _u32p))
    return out


def local_intersect(org, dirn, disc, nrm):
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    o, d, c, n = f(org), f(dirn), f(disc), f(nrm)
    return bool(lib().orc_local_intersect(_ptr(o, _fp), _ptr(d, _fp), _ptr(c, _fp), _ptr(n, _fp)))


def area_inside(box, disc, nrm):
    """box = (xmin, ymin, xmax, ymax)."""
    c = np.ascontiguousarray(disc, dtype=np.float32)
    n = np.ascontiguousarray(nrm, dtype=np.float32)
    return float(lib().orc_area_inside(*[ctypes.c_float(v) for v in box], _ptr(c, _fp), _ptr(n, _fp)))


def neighbors(xyzr):
    """Overlap graph (CSR, columns ascending) with the reference's predicate."""
    x = np.ascontiguousarray(xyzr, dtype=np.float32).reshape(-1, 4)
    n = x.shape[0]
    row = np.zeros(n + 1, dtype=np.uint32)
    e = lib().orc_neighbors(_ptr(x, _fp), n, _ptr(row, _u32p), None)
    col = np.zeros(max(int(e), 1), dtype=np.uint32)
    lib().orc_neighbors(_ptr(x, _fp), n, _ptr(row, _u32p), _ptr(col, _u32p))
    return row, col[: int(e)]


def source_origins(box, src_z, n, source=0, seed=1, instance=0, ray_begin=0):
    """Primary-ray origins (n x 3, float32) of rays [ray_begin, ray_begin + n):
    rectangle over the box (source 0) or the disc inscribed in it (source 1)."""
    p = Params()
    p.seed, p.instance, p.source = int(seed), int(instance), int(source)
    for i in range(6):
        p.box[i] = float(box[i])
    p.src_z = float(src_z)
    out = np.zeros((int(n), 3), dtype=np.float32)
    lib().orc_source_origins(ctypes.byref(p), int(ray_begin), int(n), _ptr(out, _fp))
    return out


def source_rays(box, src_z, n, cos_exponent=1.0, source=0, seed=1, instance=0, ray_begin=0):
    """Origins and directions (each n x 3, float32) of the primary rays
    [ray_begin, ray_begin + n) of a run with these parameters."""
    p = Params()
    p.seed, p.instance, p.source = int(seed), int(instance), int(source)
    p.cos_exponent = float(cos_exponent)
    for i in range(6):
        p.box[i] = float(box[i])
    p.src_z = float(src_z)
    org = np.zeros((int(n), 3), dtype=np.float32)
    dirn = np.zeros((int(n), 3), dtype=np.float32)
    lib().orc_source_rays(ctypes.byref(p), int(ray_begin), int(n), _ptr(org, _fp), _ptr(dirn, _fp))
    return org, dirn


def specular_dir(d, nrm):
    dd = np.ascontiguousarray(d, dtype=np.float32)
    nn = np.ascontiguousarray(nrm, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    lib().orc_specular_dir(_ptr(dd, _fp), _ptr(nn, _fp), _ptr(out, _fp))
    return out


def _scene_struct(scene, tie_rank=None):
    keep = []
    sc = Scene()
    if scene["kind"] == "discs":
        xyzr = np.ascontiguousarray(scene["xyzr"], dtype=np.float32)
        nrm = np.ascontiguousarray(scene["normals"], dtype=np.float32)
        keep += [xyzr, nrm]
        sc.prim_kind = 0
        sc.xyzr = _ptr(xyzr, _fp)
        sc.nrm = _ptr(nrm, _fp)
        sc.n_prims = xyzr.shape[0]
    else:
        verts = np.ascontiguousarray(scene["verts"], dtype=np.float32)
        tris = np.ascontiguousarray(scene["tris"], dtype=np.uint32)
        keep += [verts, tris]
        sc.prim_kind = 1
        sc.verts = _ptr(verts, _fp)
        sc.n_verts = verts.shape[0]
        sc.tris = _ptr(tris, _u32p)
        sc.n_prims = tris.shape[0]
    if tie_rank is not None:
        tr = np.ascontiguousarray(tie_rank, dtype=np.uint32)
        keep.append(tr)
        sc.tie_rank = _ptr(tr, _u32p)
    return sc, keep


def closest_hits(scene, box, org, dirn, tnear=1e-4, tie_rank=None):
    """One closest-hit query per ray (the oracle's BVH plus the boundary
    faces of ``box``): (prim, face, t) arrays; prim -1 when no primitive is
    hit, face -1 unless the boundary is returned."""
    sc, keep = _scene_struct(scene, tie_rank)
    p = Params()
    for i in range(6):
        p.box[i] = float(box[i])
    p.tnear = float(tnear)
    o = np.ascontiguousarray(org, dtype=np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirn, dtype=np.float32).reshape(-1, 3)
    n = o.shape[0]
    prim = np.zeros(n, dtype=np.int32)
    face = np.zeros(n, dtype=np.int32)
    t = np.zeros(n, dtype=np.float32)
    i32p = ctypes.POINTER(ctypes.c_int32)
    lib().orc_closest_hits(ctypes.byref(sc), ctypes.byref(p), n, _ptr(o, _fp), _ptr(d, _fp), prim.ctypes.data_as(i32p),
                           face.ctypes.data_as(i32p), _ptr(t, _fp))
    del keep
    return prim, face, t


def source_dir(r1, r2, n):
    out = np.zeros(3, dtype=np.float32)
    lib().orc_source_dir(float(r1), float(r2), float(n), _ptr(out, _fp))
    return out


def diffuse_dir(nrm, r1, r2):
    nn = np.ascontiguousarray(nrm, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    lib().orc_diffuse_dir(_ptr(nn, _fp), float(r1), float(r2), _ptr(out, _fp))
    return out


def roulette(w, u):
    ww = np.array([w], dtype=np.float32)
    keep = lib().orc_roulette(_ptr(ww, _fp), float(u))
    return bool(keep), float(ww[0])


def smooth_normalize(values, areas, row, col):
    v = np.ascontiguousarray(values, dtype=np.float64)
    a = np.ascontiguousarray(areas, dtype=np.float32)
    out = np.zeros(v.shape[0], dtype=np.float32)
    r = np.ascontiguousarray(row, dtype=np.uint32) if row is not None else None
    c = np.ascontiguousarray(col, dtype=np.uint32) if col is not None and len(col) else np.zeros(1, np.uint32)
    lib().orc_smooth_normalize(_ptr(v, _dp), _ptr(a, _fp), _ptr(r, _u32p), _ptr(c, _u32p) if r is not None else None,
                               v.shape[0], _ptr(out, _fp))
    return out


def trace(scene, *, n_rays, box, src_z, sticking, seed=1, instance=0, bc_x=0, bc_y=0, cos_exponent=1.0,
          diffuse_prob=1.0, tnear=1e-4, roulette=True, fp64=False, threads=0, ray_begin=0, adjoint=False,
          neighbors=None, moments=False, source=0, tie_rank=None):
    """Run the oracle trace. ``scene`` is a dict from ``rti_amd.geometry``.

    Returns a dict with s1, s2, cnt (per primitive), s3 and s4 with
    ``moments``, and the counters. ``tie_rank`` (one entry per primitive):
    of hits at exactly equal distance the primitive of lower rank is the
    closest (default: lower index).
    """
    keep = []
    sc = Scene()
    if scene["kind"] == "discs":
        xyzr = np.ascontiguousarray(scene["xyzr"], dtype=np.float32)
        nrm = np.ascontiguousarray(scene["normals"], dtype=np.float32)
        keep += [xyzr, nrm]
        n = xyzr.shape[0]
        sc.prim_kind = 0
        sc.xyzr = _ptr(xyzr, _fp)
        sc.nrm = _ptr(nrm, _fp)
        if neighbors is None:
            neighbors = globals()["neighbors"](xyzr)
        row, col = neighbors
        row = np.ascontiguousarray(row, dtype=np.uint32)
        col = np.ascontiguousarray(col if len(col) else np.zeros(1), dtype=np.uint32)
        keep += [row, col]
        sc.nb_ptr = _ptr(row, _u32p)
        sc.nb_col = _ptr(col, _u32p)
    else:
        verts = np.ascontiguousarray(scene["verts"], dtype=np.float32)
        tris = np.ascontiguousarray(scene["tris"], dtype=np.uint32)
        keep += [verts, tris]
        n = tris.shape[0]
        sc.prim_kind = 1
        sc.verts = _ptr(verts, _fp)
        sc.n_verts = verts.shape[0]
        sc.tris = _ptr(tris, _u32p)
    sc.n_prims = n
    if tie_rank is not None:
        tr = np.ascontiguousarray(tie_rank, dtype=np.uint32)
        assert tr.shape == (n,)
        keep.append(tr)
        sc.tie_rank = _ptr(tr, _u32p)
    p = Params()
    p.n_rays = int(n_rays)
    p.ray_begin = int(ray_begin)
    p.seed = int(seed)
    p.instance = int(instance)
    p.bc_x, p.bc_y = int(bc_x), int(bc_y)
    p.sticking = float(sticking)
    p.cos_exponent = float(cos_exponent)
    p.diffuse_prob = float(diffuse_prob)
    p.tnear = float(tnear)
    p.roulette = 1 if roulette else 0
    p.fp64 = 1 if fp64 else 0
    for i in range(6):
        p.box[i] = float(box[i])
    p.src_z = float(src_z)
    p.threads = int(threads)
    p.adjoint = 1 if adjoint else 0
    p.moments = 1 if moments else 0
    p.source = int(source)
    out = {"s1": np.zeros(n), "s2": np.zeros(n), "cnt": np.zeros(n, dtype=np.uint64)}
    if moments:
        out["s3"] = np.zeros(n)
        out["s4"] = np.zeros(n)
    if adjoint:
        out["ds"] = np.zeros(n)
        out["dn"] = np.zeros(n)
    r = Result()
    r.s1 = _ptr(out["s1"], _dp)
    r.s2 = _ptr(out["s2"], _dp)
    r.cnt = _ptr(out["cnt"], _u64p)
    if adjoint:
        r.dsum_ds = _ptr(out["ds"], _dp)
        r.dsum_dn = _ptr(out["dn"], _dp)
    if moments:
        r.s3 = _ptr(out["s3"], _dp)
        r.s4 = _ptr(out["s4"], _dp)
    rc = lib().orc_trace(ctypes.byref(sc), ctypes.byref(p), ctypes.byref(r))
    if rc != 0:
        raise RuntimeError(f"orc_trace failed with code {rc}")
    out.update(segments=int(r.segments), surface_hits=int(r.surface_hits), misses=int(r.misses),
               seconds=float(r.seconds))
    return out


def error_stats(s1, s2, cnt, s3=None, s4=None, total=None):
    """Relative error and (with s3, s4) variance of the variance per primitive,
    from the power sums; total = all deposits of the run (default cnt.sum()).
    Restates hit_accumulator::get_relative_error / get_vov
    (src/rti/trace/hit_accumulator.hpp:152-249)."""
    s1 = np.ascontiguousarray(s1, dtype=np.float64)
    s2 = np.ascontiguousarray(s2, dtype=np.float64)
    cnt = np.ascontiguousarray(cnt, dtype=np.uint64)
    n = s1.shape[0]
    total = int(cnt.sum()) if total is None else int(total)
    rel = np.zeros(n)
    vov = np.zeros(n) if s3 is not None else None
    z = np.zeros(1)
    a3 = np.ascontiguousarray(s3, dtype=np.float64) if s3 is not None else z
    a4 = np.ascontiguousarray(s4, dtype=np.float64) if s4 is not None else z
    lib().orc_error_stats(_ptr(s1, _dp), _ptr(s2, _dp), _ptr(a3, _dp), _ptr(a4, _dp), _ptr(cnt, _u64p), n, total,
                          _ptr(rel, _dp), _ptr(vov, _dp) if vov is not None else None)
    return rel, vov
