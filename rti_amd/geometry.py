"""Synthetic, seeded scene generators for the flux trace.

Every generator returns a plain dict:

* discs:     ``{"kind": "discs", "xyzr": (N,4) float32, "normals": (N,3) float32, ...}``
  -- an oriented-disc point cloud in the layout of the reference's Embree
  ORIENTED_DISC_POINT buffers (x, y, z, radius) + unit normal
  (src/rti/geo/point_cloud_disc_geometry.hpp:178-222);
* triangles: ``{"kind": "triangles", "verts": (V,3) float32, "tris": (T,3) uint32, ...}``
  with the winding normal (v1-v0) x (v2-v0) pointing into the open region.

Positions are computed in float64 and rounded once to float32.
"""
import math

import numpy as np


def _finish_discs(xyz, nrm, radius, **meta):
    nrm = np.asarray(nrm, dtype=np.float64)
    nrm = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    xyzr = np.empty((xyz.shape[0], 4), dtype=np.float32)
    xyzr[:, :3] = xyz
    xyzr[:, 3] = radius
    d = {"kind": "discs", "xyzr": xyzr, "normals": nrm.astype(np.float32)}
    d.update(meta)
    return d


def _profile_points(vertices, normals, h):
    """Sample a 2-D polyline (x, z) at arc-length step h.

    ``normals[k]`` is the normal of segment k. Points that fall on a polyline
    vertex between two segments get the normalised mean of both normals.
    Returns (x, z, nx, nz) arrays.
    """
    xs, zs, nxs, nzs = [], [], [], []
    seglen = [math.dist(vertices[k], vertices[k + 1]) for k in range(len(vertices) - 1)]
    total = sum(seglen)
    npts = int(round(total / h))
    acc = np.concatenate([[0.0], np.cumsum(seglen)])
    for i in range(npts + 1):
        s = min(i * h, total)
        k = int(np.searchsorted(acc, s, side="right") - 1)
        k = min(k, len(seglen) - 1)
        t = (s - acc[k]) / seglen[k]
        p0, p1 = vertices[k], vertices[k + 1]
        x = p0[0] + (p1[0] - p0[0]) * t
        z = p0[1] + (p1[1] - p0[1]) * t
        n = np.array(normals[k], dtype=np.float64)
        if abs(t) < 1e-9 and k > 0:
            n = n + np.array(normals[k - 1])
        elif abs(t - 1) < 1e-9 and k + 1 < len(seglen):
            n = n + np.array(normals[k + 1])
        n /= np.linalg.norm(n)
        xs.append(x); zs.append(z); nxs.append(n[0]); nzs.append(n[1])
    return np.array(xs), np.array(zs), np.array(nxs), np.array(nzs)


def trench_discs(width=1.0, depth=4.0, flat=5.5, ylen=1.0, spacing=0.02, radius=None):
    """2-D-extruded rectangular trench as an oriented-disc cloud.

    The floor is at z = 0 and the top flats at z = depth. x spans
    [-width/2 - flat, width/2 + flat]; y spans [0, ylen] with rows every
    ``spacing``. Radius defaults to spacing * sqrt(2) / 2 (discs overlap their
    4 axis neighbours and touch the diagonal ones). With the defaults this is
    the 1e6-ray CPU-runnable trench of the benchmark (51,051 discs).
    """
    if radius is None:
        radius = spacing * math.sqrt(2.0) / 2.0
    w2 = width / 2.0
    verts = [(-w2 - flat, depth), (-w2, depth), (-w2, 0.0), (w2, 0.0), (w2, depth), (w2 + flat, depth)]
    nrms = [(0, 1), (1, 0), (0, 1), (-1, 0), (0, 1)]
    px, pz, nx, nz = _profile_points(verts, nrms, spacing)
    ny_rows = int(round(ylen / spacing)) + 1
    ys = np.arange(ny_rows, dtype=np.float64) * spacing
    X = np.tile(px, ny_rows)
    Z = np.tile(pz, ny_rows)
    Y = np.repeat(ys, px.shape[0])
    xyz = np.stack([X, Y, Z], axis=1)
    nrm = np.stack([np.tile(nx, ny_rows), np.zeros_like(X), np.tile(nz, ny_rows)], axis=1)
    return _finish_discs(xyz, nrm, radius, width=width, depth=depth, flat=flat, ylen=ylen, spacing=spacing,
                         profile_x=px, profile_z=pz)


def trench_triangles(width=1.0, depth=4.0, flat=5.5, ylen=1.0, spacing=0.02):
    """The trench of ``trench_discs`` as a triangle mesh: the same profile
    points, extruded along y in rows of ``spacing``, two triangles per quad,
    winding normals into the open region."""
    sc = trench_discs(width=width, depth=depth, flat=flat, ylen=ylen, spacing=spacing)
    px, pz = sc["profile_x"], sc["profile_z"]
    pn = sc["normals"][: px.size].astype(np.float64)
    ny_rows = int(round(ylen / spacing)) + 1
    ys = np.arange(ny_rows, dtype=np.float64) * spacing
    P = px.size
    verts = np.stack([np.tile(px, ny_rows), np.repeat(ys, P), np.tile(pz, ny_rows)], axis=1)
    j, i = np.meshgrid(np.arange(ny_rows - 1), np.arange(P - 1), indexing="ij")
    a = (j * P + i).ravel()
    b = a + 1
    c = a + P + 1
    d = a + P
    tris = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    seg = np.concatenate([i.ravel(), i.ravel()])
    want = pn[seg] + pn[seg + 1]
    want[:, 1] = 0.0
    tris = _orient(verts, tris, want)
    return {"kind": "triangles", "verts": verts.astype(np.float32), "tris": tris.astype(np.uint32),
            "width": width, "depth": depth, "flat": flat, "ylen": ylen, "spacing": spacing}


def hole_discs(radius=0.5, depth=5.0, half=1.0, spacing=0.02, disc_radius=None):
    """Cylindrical hole (axis z, floor z = 0, top z = depth) in a
    [-half, half]^2 cell, as an oriented-disc cloud with inward normals."""
    if disc_radius is None:
        disc_radius = spacing * math.sqrt(2.0) / 2.0
    pts, nrm = [], []
    # wall
    n_az = max(8, int(round(2 * math.pi * radius / spacing)))
    n_z = max(1, int(round(depth / spacing)))
    phi = np.arange(n_az) * (2 * math.pi / n_az)
    zz = (np.arange(n_z + 1) / n_z) * depth
    P, Zw = np.meshgrid(phi, zz)
    P, Zw = P.ravel(), Zw.ravel()
    pts.append(np.stack([radius * np.cos(P), radius * np.sin(P), Zw], 1))
    nrm.append(np.stack([-np.cos(P), -np.sin(P), np.zeros_like(P)], 1))
    # floor and top on a square grid
    g = np.arange(-half, half + 0.5 * spacing, spacing)
    GX, GY = np.meshgrid(g, g)
    GX, GY = GX.ravel(), GY.ravel()
    rr = np.hypot(GX, GY)
    fl = rr < radius - 0.5 * spacing
    pts.append(np.stack([GX[fl], GY[fl], np.zeros(fl.sum())], 1))
    nrm.append(np.tile([0.0, 0.0, 1.0], (fl.sum(), 1)))
    tp = rr > radius + 0.5 * spacing
    pts.append(np.stack([GX[tp], GY[tp], np.full(tp.sum(), depth)], 1))
    nrm.append(np.tile([0.0, 0.0, 1.0], (tp.sum(), 1)))
    return _finish_discs(np.concatenate(pts), np.concatenate(nrm), disc_radius, hole_radius=radius, depth=depth,
                         half=half, spacing=spacing)


def _orient(verts, tris, want):
    """Flip triangles whose winding normal disagrees with ``want`` (T,3)."""
    v = verts[tris]
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    bad = (n * want).sum(1) < 0
    tris = tris.copy()
    tris[bad] = tris[bad][:, [0, 2, 1]]
    return tris


def hole_triangles(radius=0.5, depth=5.0, half=1.0, n_az=320, n_ax=256, n_floor=16, n_top=40):
    """Triangle mesh of a cylindrical hole in a [-half, half]^2 periodic cell.

    Uniform azimuthal x axial tessellation of the wall, floor rings, and a flat
    top annulus that reaches the cell edges (n_az must be a multiple of 8 so
    the cell corners are vertices). Defaults give 199,360 triangles.
    """
    assert n_az % 8 == 0
    phi = np.arange(n_az) * (2 * math.pi / n_az)
    c, s = np.cos(phi), np.sin(phi)
    V, T, W = [], [], []
    base = 0

    def grid(P, nu, nv, want, periodic_u=True):
        nonlocal base
        # P: (nv+1, nu, 3) vertex grid, quads between consecutive rows / columns
        V.append(P.reshape(-1, 3))
        i = np.arange(nu)
        ip = (i + 1) % nu if periodic_u else i + 1
        tris = []
        for j in range(nv):
            a = base + j * nu + i
            b = base + j * nu + ip
            cc = base + (j + 1) * nu + ip
            d = base + (j + 1) * nu + i
            tris.append(np.stack([a, b, cc], 1))
            tris.append(np.stack([a, cc, d], 1))
        tris = np.concatenate(tris)
        T.append(tris)
        W.append(np.tile(want, (tris.shape[0], 1)) if np.ndim(want) == 1 else want)
        base += P.reshape(-1, 3).shape[0]

    # wall: rows along z
    zz = np.linspace(0.0, depth, n_ax + 1)
    P = np.stack([np.broadcast_to(radius * c, (n_ax + 1, n_az)), np.broadcast_to(radius * s, (n_ax + 1, n_az)),
                  np.broadcast_to(zz[:, None], (n_ax + 1, n_az))], -1)
    nu = n_az
    # wall_want must follow the triangle order of grid(): per row j, n_az "abc" then n_az "acd"
    wall_want = np.concatenate([np.stack([-c, -s, np.zeros_like(c)], 1)] * (2 * n_ax))
    grid(P, nu, n_ax, wall_want)
    # floor: centre fan + rings
    rr = np.linspace(0.0, radius, n_floor + 1)[1:]
    P = np.stack([rr[:, None] * c, rr[:, None] * s, np.zeros((n_floor, n_az))], -1)
    start = base
    grid(P, n_az, n_floor - 1, np.array([0.0, 0.0, 1.0]))
    V.append(np.array([[0.0, 0.0, 0.0]]))
    centre = base
    base += 1
    i = np.arange(n_az)
    fan = np.stack([np.full(n_az, centre), start + i, start + (i + 1) % n_az], 1)
    T.append(fan)
    W.append(np.tile([0.0, 0.0, 1.0], (n_az, 1)))
    # top annulus from the rim to the square cell boundary
    sq = 1.0 / np.maximum(np.abs(c), np.abs(s)) * half
    m = np.linspace(0.0, 1.0, n_top + 1)[:, None]
    inner = np.stack([radius * c, radius * s], 1)
    outer = np.stack([sq * c, sq * s], 1)
    xy = inner[None] + (outer - inner)[None] * m[..., None]
    P = np.concatenate([xy, np.full((n_top + 1, n_az, 1), depth)], -1)
    grid(P, n_az, n_top, np.array([0.0, 0.0, 1.0]))
    verts = np.concatenate(V)
    tris = np.concatenate(T).astype(np.int64)
    want = np.concatenate(W)
    tris = _orient(verts, tris, want)
    return {"kind": "triangles", "verts": verts.astype(np.float32), "tris": tris.astype(np.uint32),
            "radius": radius, "depth": depth, "half": half}


def flat_plate_discs(size=1.0, spacing=0.02, z=0.0):
    """Flat periodic plate: every disc sees only the source (flux 1)."""
    g = np.arange(0.0, size + 0.5 * spacing, spacing)
    GX, GY = np.meshgrid(g, g)
    xyz = np.stack([GX.ravel(), GY.ravel(), np.full(GX.size, z)], 1)
    nrm = np.tile([0.0, 0.0, 1.0], (GX.size, 1))
    return _finish_discs(xyz, nrm, spacing * math.sqrt(2.0) / 2.0, spacing=spacing)


def flat_plate_triangles(size=1.0, cells=32, z=0.0):
    """Flat plate of cells x cells squares, two triangles each, normals +z."""
    g = np.linspace(0.0, size, cells + 1)
    GX, GY = np.meshgrid(g, g)
    verts = np.stack([GX.ravel(), GY.ravel(), np.full(GX.size, z)], 1)
    i, j = np.meshgrid(np.arange(cells), np.arange(cells))
    v00 = (j * (cells + 1) + i).ravel()
    v10, v01, v11 = v00 + 1, v00 + cells + 1, v00 + cells + 2
    tris = np.concatenate([np.stack([v00, v10, v01], 1), np.stack([v10, v11, v01], 1)])
    return {"kind": "triangles", "verts": verts.astype(np.float32), "tris": tris.astype(np.uint32)}


def scene_box(scene, src_z=None):
    """Boundary box and source height the library derives from a scene.

    Discs: x/y extent of the centres; z from the lowest centre to the highest
    centre + 2 * max radius (src/rti/geo/point_cloud_disc_geometry.hpp:83-87
    and src/rti/device.hpp:88-90, 224-237); the source plane sits at that top
    (device.hpp:257-267). Triangles: vertex extent, source 1% of the lateral
    extent above the top. An explicit ``src_z`` moves the source plane and
    with it the top of the boundary faces.
    Returns (box[6] = xmin, xmax, ymin, ymax, zmin, zmax, src_z) in float32.
    """
    f = np.float32
    if scene["kind"] == "discs":
        p = scene["xyzr"]
        lo = p[:, :3].min(0)
        hi = p[:, :3].max(0)
        rmax = p[:, 3].max()
        ztop = f(f(hi[2] + rmax) + rmax)
    else:
        v = scene["verts"]
        lo = v.min(0)
        hi = v.max(0)
        ztop = f(hi[2] + f(0.01) * f(max(hi[0] - lo[0], hi[1] - lo[1])))
    sz = ztop if src_z is None else f(src_z)
    box = np.array([lo[0], hi[0], lo[1], hi[1], lo[2], sz], dtype=np.float32)
    return box, f(sz)


def rough_surface_discs(seed=5, nfield=4096, cells=4, hole_radius=0.15, aspect=6.0, spacing=0.0033,
                        n_bumps=256):
    """Periodic rough surface over a cells x cells array of holes.

    A periodic height field of ``nfield``^2 FP32 samples (sum of ``n_bumps``
    Gaussian bumps, amplitude U[0, 0.3], width U[0.02, 0.2]) covers a
    [0, cells]^2 domain; discs on a ``spacing`` grid sample it bilinearly,
    with normals from the field gradient. Each unit cell holds a hole of
    radius ``hole_radius`` and depth aspect * 2 * radius (walls and floor as
    discs). Defaults give about 4M discs.
    """
    rng = np.random.default_rng(seed)
    L = float(cells)
    amp = rng.uniform(0.0, 0.3, n_bumps)
    wid = rng.uniform(0.02, 0.2, n_bumps)
    cx = rng.uniform(0.0, L, n_bumps)
    cy = rng.uniform(0.0, L, n_bumps)
    g = (np.arange(nfield, dtype=np.float64) + 0.5) * (L / nfield)
    # each bump is separable, a(b) gy_b(y) gx_b(x): the field is one matrix
    # product GY^T diag(a) GX (FP64, rounded once to FP32)
    dx = (g[None, :] - cx[:, None] + L / 2) % L - L / 2
    dy = (g[None, :] - cy[:, None] + L / 2) % L - L / 2
    gx = np.exp(-0.5 * (dx / wid[:, None]) ** 2)
    gy = np.exp(-0.5 * (dy / wid[:, None]) ** 2)
    field = ((gy * amp[:, None]).T @ gx).astype(np.float32)
    # gradient (periodic central differences)
    hstep = L / nfield
    fx = (np.roll(field, -1, 1) - np.roll(field, 1, 1)) / np.float32(2 * hstep)
    fy = (np.roll(field, -1, 0) - np.roll(field, 1, 0)) / np.float32(2 * hstep)

    def sample(arr, x, y):
        u = x / hstep - 0.5
        v = y / hstep - 0.5
        i0 = np.floor(u).astype(np.int64)
        j0 = np.floor(v).astype(np.int64)
        tu = (u - i0).astype(np.float32)
        tv = (v - j0).astype(np.float32)
        i0 %= nfield
        j0 %= nfield
        i1 = (i0 + 1) % nfield
        j1 = (j0 + 1) % nfield
        return ((1 - tv) * ((1 - tu) * arr[j0, i0] + tu * arr[j0, i1]) +
                tv * ((1 - tu) * arr[j1, i0] + tu * arr[j1, i1]))

    depth = aspect * 2.0 * hole_radius
    r_d = spacing * math.sqrt(2.0) / 2.0
    gg = np.arange(0.0, L + 0.5 * spacing, spacing)
    pts, nrm = [], []
    for row in range(gg.size):
        y = np.full(gg.size, gg[row])
        x = gg
        ccx = np.floor(x) + 0.5
        ccy = np.floor(np.minimum(y, L - 1e-9)) + 0.5
        ccx = np.minimum(ccx, L - 0.5)
        rr = np.hypot(x - ccx, y - ccy)
        keep = rr > hole_radius + 0.5 * spacing
        z = sample(field, x[keep], y[keep])
        gx = sample(fx, x[keep], y[keep])
        gy = sample(fy, x[keep], y[keep])
        pts.append(np.stack([x[keep], y[keep], z], 1))
        nrm.append(np.stack([-gx, -gy, np.ones_like(gx)], 1))
    # holes
    n_az = max(8, int(round(2 * math.pi * hole_radius / spacing)))
    phi = np.arange(n_az) * (2 * math.pi / n_az)
    for hx in range(cells):
        for hy in range(cells):
            x0, y0 = hx + 0.5, hy + 0.5
            rx = x0 + hole_radius * np.cos(phi)
            ry = y0 + hole_radius * np.sin(phi)
            ztop = sample(field, rx, ry).astype(np.float64)
            nz = int(round((depth + ztop.max()) / spacing))
            for k in range(nz + 1):
                z = -depth + k * spacing
                m = z <= ztop
                if not m.any():
                    continue
                pts.append(np.stack([rx[m], ry[m], np.full(m.sum(), z)], 1))
                nrm.append(np.stack([-np.cos(phi[m]), -np.sin(phi[m]), np.zeros(m.sum())], 1))
            fg = np.arange(-hole_radius, hole_radius + 0.5 * spacing, spacing)
            FX, FY = np.meshgrid(fg, fg)
            m = np.hypot(FX, FY) < hole_radius - 0.5 * spacing
            pts.append(np.stack([x0 + FX[m], y0 + FY[m], np.full(m.sum(), -depth)], 1))
            nrm.append(np.tile([0.0, 0.0, 1.0], (m.sum(), 1)))
    return _finish_discs(np.concatenate(pts), np.concatenate(nrm), r_d, spacing=spacing, cells=cells,
                         hole_radius=hole_radius, depth=depth)


def batch_instances(seed=3, count=64, total_rays=None):
    """The 64-instance batch: trenches and holes as disc clouds.

    Per instance: type trench/hole, aspect ratio U[2, 20], disc count
    log-uniform in [1e4, 5e5] (grid spacing clamped to [0.01, 0.05]),
    sticking log-uniform in [1e-3, 1], reflection diffuse or specular, and a
    ray budget proportional to the surface area, clamped to [1e6, 5e7].
    ``total_rays`` rescales the budgets (keeping proportions) when given.
    Returns a list of dicts with keys scene, sticking, diffuse_prob, n_rays.
    """
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = "trench" if rng.random() < 0.5 else "hole"
        ar = rng.uniform(2.0, 20.0)
        ntarget = math.exp(rng.uniform(math.log(1e4), math.log(5e5)))
        s = math.exp(rng.uniform(math.log(1e-3), math.log(1.0)))
        diffuse = 1.0 if rng.random() < 0.5 else 0.0
        if kind == "trench":
            width, depth, flat, ylen = 1.0, ar * 1.0, 0.5, 1.0
            area = (2 * flat + 2 * depth + width) * ylen
            h = float(np.clip(math.sqrt(area / ntarget), 0.01, 0.05))
            sc = trench_discs(width=width, depth=depth, flat=flat, ylen=ylen, spacing=h)
        else:
            radius, half = 0.5, 1.0
            depth = ar * 2 * radius
            area = 2 * math.pi * radius * depth + 4 * half * half
            h = float(np.clip(math.sqrt(area / ntarget), 0.01, 0.05))
            sc = hole_discs(radius=radius, depth=depth, half=half, spacing=h)
        out.append({"scene": sc, "sticking": s, "diffuse_prob": diffuse, "area": area, "kind": kind, "aspect": ar})
    areas = np.array([o["area"] for o in out])
    rays = np.clip(areas / areas.mean() * 1e7, 1e6, 5e7)
    if total_rays is not None:
        rays = rays * (total_rays / rays.sum())
    for o, r in zip(out, rays):
        o["n_rays"] = int(max(1, round(r)))
    return out
