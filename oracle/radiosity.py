"""FP64 radiosity reference for the flux on diffusely reflecting cavities.

TEST INFRASTRUCTURE ONLY (the checker for per-element flux; nothing in the
product imports it).

Model (the continuum limit of the reference's Monte Carlo, for s in (0, 1],
diffuse reflection): an element i receives the incident flux

    G_i = G0_i + (1 - s) * sum_j F_ij G_j,

where F_ij is the diffuse view factor from i to j and G0_i the direct flux
from the source; G is relative to the source flux density through a
horizontal plane, which is what the Monte Carlo estimates as
raw_i / (s A_i) * A_src / N_rays. The source plane is laterally infinite
(periodic / reflective domain), so the flux entering a cavity equals that of
its opening emitting with the source's angular distribution: per unit area
and solid angle (n + 1) / (2 pi) cos^n(theta) (power_cosine_direction_z,
src/rti/ray/power_cosine_direction_z.hpp:17-32; Lambertian for n = 1).

* Cylindrical hole (axisymmetric): wall bands and floor annuli. View factors
  are closed forms built from the coaxial disc-to-disc factor (Howell's
  catalogue, C-41) with the window (isothermal enclosure) identity; the
  direct flux for n != 1 is a Gauss-Legendre quadrature over the opening.
  The reference's AR-45 solution (verification/radiosity) pins this solver.
* Trench (2-D, extruded along y): wall and floor strips; view factors by
  Hottel's crossed strings; the direct flux for n != 1 integrates the
  extruded source over y in closed form and over the opening numerically.
"""
import math

import numpy as np


# ---------------------------------------------------------------------------
# coaxial discs
# ---------------------------------------------------------------------------
def disc_to_disc(r1, r2, h):
    """View factor from a disc of radius r1 to a parallel coaxial disc of
    radius r2 at distance h (h > 0)."""
    r1 = np.asarray(r1, dtype=np.float64)
    r2 = np.asarray(r2, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    R1 = r1 / h
    R2 = r2 / h
    S = 1.0 + (1.0 + R2 * R2) / (R1 * R1)
    return 0.5 * (S - np.sqrt(np.maximum(S * S - 4.0 * (r2 / r1) ** 2, 0.0)))


def _D(a, h):
    """Disc-to-disc factor for equal radii a at distance h (h = 0 gives 1)."""
    h = np.asarray(h, dtype=np.float64)
    out = np.ones_like(h)
    m = h > 0
    out[m] = disc_to_disc(a, a, h[m])
    return out


def hole_elements(radius, depth, n_wall, n_floor):
    """Element layout: wall bands (depth below the opening) then floor annuli."""
    z = np.linspace(0.0, depth, n_wall + 1)
    rho = np.linspace(0.0, radius, n_floor + 1)
    return z, rho


def hole_view_factors(radius, depth, z, rho):
    a = float(radius)
    H = float(depth)
    Ad = math.pi * a * a
    nw = len(z) - 1
    nf = len(rho) - 1
    n = nw + nf
    F = np.zeros((n, n))
    Aw = 2.0 * math.pi * a * np.diff(z)
    Af = math.pi * np.diff(rho ** 2)
    z1 = z[:-1][:, None]
    z2 = z[1:][:, None]
    z3 = z[:-1][None, :]
    z4 = z[1:][None, :]
    # band i above band j (z2 <= z3): A_i F_ij = Ad [D(z3-z2) - D(z3-z1) - D(z4-z2) + D(z4-z1)]
    below = z3 >= z2
    AF = Ad * (_D(a, np.abs(z3 - z2)) - _D(a, np.abs(z3 - z1)) - _D(a, np.abs(z4 - z2)) + _D(a, np.abs(z4 - z1)))
    Fww = np.where(below, AF, 0.0)
    Fww = Fww + np.triu(Fww, 1).T * 0  # placeholder, filled by reciprocity below
    Fww_full = np.zeros((nw, nw))
    iu = np.triu_indices(nw, 1)
    Fww_full[iu] = AF[iu]
    Fww_full = Fww_full + Fww_full.T  # A_i F_ij is symmetric
    # self: 1 - 2 (Ad / A_i)(1 - D(dz))
    dz = np.diff(z)
    selfF = 1.0 - 2.0 * (Ad / Aw) * (1.0 - _D(a, dz))
    F[:nw, :nw] = Fww_full / Aw[:, None]
    F[np.arange(nw), np.arange(nw)] = selfF
    # band -> floor annulus [r1, r2]:
    # A_i F = Ad [ (Fdd(a,r2,H-z2) - Fdd(a,r1,H-z2)) - (Fdd(a,r2,H-z1) - Fdd(a,r1,H-z1)) ]
    def Fdd(r, h):
        r = np.asarray(r, dtype=np.float64)
        h = np.asarray(h, dtype=np.float64)
        out = np.zeros(np.broadcast(r, h).shape)
        rr = np.broadcast_to(r, out.shape)
        hh = np.broadcast_to(h, out.shape)
        m = (rr > 0) & (hh > 0)
        out[m] = disc_to_disc(a, rr[m], hh[m])
        # at h = 0 the annulus of the floor plane itself sees nothing of the
        # disc plane; r = 0 is an empty disc
        m0 = (rr > 0) & (hh <= 0)
        out[m0] = np.minimum(rr[m0] / a, 1.0) ** 2
        return out
    r1 = rho[:-1][None, :]
    r2 = rho[1:][None, :]
    hz1 = (H - z[:-1])[:, None]
    hz2 = (H - z[1:])[:, None]
    AFwf = Ad * ((Fdd(r2, hz2) - Fdd(r1, hz2)) - (Fdd(r2, hz1) - Fdd(r1, hz1)))
    F[:nw, nw:] = AFwf / Aw[:, None]
    F[nw:, :nw] = AFwf.T / Af[:, None]
    # opening factors
    Fw_open = (Ad / Aw) * (_D(a, z[:-1]) - _D(a, z[1:]))
    Ff_open = (Ad / Af) * (Fdd(rho[1:], np.full(nf, H)) - Fdd(rho[:-1], np.full(nf, H)))
    return F, np.concatenate([Fw_open, Ff_open]), np.concatenate([Aw, Af])


def _gauss(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def hole_direct_flux(radius, depth, z, rho, cos_exponent, nq=48):
    """Direct flux on wall bands and floor annuli from the opening emitting
    (n + 1) / (2 pi) cos^n per unit area and solid angle."""
    n = float(cos_exponent)
    a = float(radius)
    H = float(depth)
    c = (n + 1.0) / (2.0 * math.pi)
    xr, wr = _gauss(nq)
    xp, wp = _gauss(2 * nq)
    # source points on the opening disc (polar)
    rs = 0.5 * a * (xr + 1.0)
    wrs = 0.5 * a * wr * rs
    ph = math.pi * (xp + 1.0)
    wph = math.pi * wp
    sx = (rs[:, None] * np.cos(ph)[None, :]).ravel()
    sy = (rs[:, None] * np.sin(ph)[None, :]).ravel()
    sw = (wrs[:, None] * wph[None, :]).ravel()
    out = []
    xz, wz = _gauss(8)
    for k in range(len(z) - 1):
        # receiver on the wall at (a, 0, -zz), inward normal (-1, 0, 0); average over the band
        acc = 0.0
        for t, wt in zip(0.5 * (xz + 1), 0.5 * wz):
            zz = z[k] + (z[k + 1] - z[k]) * t
            dx = sx - a
            dy = sy
            dz = zz
            r2 = dx * dx + dy * dy + dz * dz
            cs = dz / np.sqrt(r2)            # cos at the source (normal -z into the hole)
            cx = -dx / np.sqrt(r2)           # cos at the wall point
            acc += wt * np.sum(sw * c * cs ** n * cx / r2)
        out.append(acc)
    for k in range(len(rho) - 1):
        acc = 0.0
        area = 0.0
        for t, wt in zip(0.5 * (xz + 1), 0.5 * wz):
            rr = rho[k] + (rho[k + 1] - rho[k]) * t
            dx = sx - rr
            dy = sy
            r2 = dx * dx + dy * dy + H * H
            cs = H / np.sqrt(r2)
            acc += wt * rr * np.sum(sw * c * cs ** n * cs / r2)
            area += wt * rr
        out.append(acc / area)
    return np.array(out)


def solve_hole(radius, depth, sticking, cos_exponent=1.0, n_wall=300, n_floor=20, direct="closed"):
    """Incident flux on the hole's wall bands and floor annuli.

    Returns dict(z (band edges, depth below the opening), rho (annulus
    edges), wall, floor, F, G0)."""
    z, rho = hole_elements(radius, depth, n_wall, n_floor)
    F, Fopen, A = hole_view_factors(radius, depth, z, rho)
    if cos_exponent == 1.0 and direct == "closed":
        G0 = Fopen
    else:
        G0 = hole_direct_flux(radius, depth, z, rho, cos_exponent)
    M = np.eye(F.shape[0]) - (1.0 - sticking) * F
    G = np.linalg.solve(M, G0)
    nw = n_wall
    return {"z": z, "rho": rho, "wall": G[:nw], "floor": G[nw:], "F": F, "G0": G0, "areas": A}


# ---------------------------------------------------------------------------
# 2-D trench (crossed strings)
# ---------------------------------------------------------------------------
def _strips(width, depth, n_wall, n_floor):
    """Strips of the trench cavity in the x-z plane: left wall (top to
    bottom), floor (left to right), right wall (bottom to top). Floor at
    z = 0, opening at z = depth, walls at x = -/+ width / 2."""
    w2 = width / 2.0
    zw = np.linspace(depth, 0.0, n_wall + 1)
    xf = np.linspace(-w2, w2, n_floor + 1)
    segs = []
    for k in range(n_wall):
        segs.append(((-w2, zw[k]), (-w2, zw[k + 1]), (1.0, 0.0)))
    for k in range(n_floor):
        segs.append(((xf[k], 0.0), (xf[k + 1], 0.0), (0.0, 1.0)))
    for k in range(n_wall):
        segs.append(((w2, zw[n_wall - k]), (w2, zw[n_wall - k - 1]), (-1.0, 0.0)))
    return segs


def _crossed(a, b):
    (p1, p2), (q1, q2) = a, b
    d = lambda u, v: math.hypot(u[0] - v[0], u[1] - v[1])
    crossed = d(p1, q2) + d(p2, q1)
    uncrossed = d(p1, q1) + d(p2, q2)
    return abs(crossed - uncrossed) / 2.0


def solve_trench(width, depth, sticking, cos_exponent=1.0, n_wall=200, n_floor=50):
    """Incident flux on the trench's strips (2-D radiosity, crossed strings).

    Returns dict(strips: list of ((x0,z0),(x1,z1),normal), flux, G0)."""
    segs = _strips(width, depth, n_wall, n_floor)
    n = len(segs)
    L = np.array([math.dist(s[0], s[1]) for s in segs])
    F = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            ni, nj = segs[i][2], segs[j][2]
            if ni == nj:   # coplanar strips of one face do not see each other
                continue
            lf = _crossed((segs[i][0], segs[i][1]), (segs[j][0], segs[j][1]))
            F[i, j] = lf / L[i]
            F[j, i] = lf / L[j]
    w2 = width / 2.0
    opening = ((-w2, depth), (w2, depth))
    Fopen = np.array([_crossed((s[0], s[1]), opening) / L[k] for k, s in enumerate(segs)])
    if cos_exponent == 1.0:
        G0 = Fopen
    else:
        G0 = trench_direct_flux(segs, width, depth, cos_exponent)
    G = np.linalg.solve(np.eye(n) - (1.0 - sticking) * F, G0)
    return {"strips": segs, "flux": G, "G0": G0, "F": F, "lengths": L}


def trench_direct_flux(segs, width, depth, cos_exponent, nq=64):
    """Direct flux from the opening line (extruded along y) for cos^n.

    Integrating (n + 1) / (2 pi) cos^n(theta_s) cos(theta_x) / r^2 over y in
    closed form leaves a 1-D integral over the opening, done by Gauss-Legendre."""
    n = float(cos_exponent)
    c = (n + 1.0) / (2.0 * math.pi)
    w2 = width / 2.0
    xs, ws = _gauss(nq)
    sx = w2 * xs
    sw = w2 * ws
    # int_{-inf}^{inf} dy / (rho^2 + y^2)^{(n+3)/2} = sqrt(pi) Gamma((n+2)/2) / Gamma((n+3)/2) / rho^(n+2)
    k = math.sqrt(math.pi) * math.gamma((n + 2) / 2.0) / math.gamma((n + 3) / 2.0)
    out = []
    xg, wg = _gauss(8)
    for (p0, p1, nr) in segs:
        acc = 0.0
        for t, wt in zip(0.5 * (xg + 1), 0.5 * wg):
            px = p0[0] + (p1[0] - p0[0]) * t
            pz = p0[1] + (p1[1] - p0[1]) * t
            dx = sx - px
            dz = depth - pz
            rho2 = dx * dx + dz * dz
            # cos_s = dz / r, cos_x = (n . (dx, 0, dz)) / r with r^2 = rho^2 + y^2
            num = dz ** n * (nr[0] * dx + nr[1] * dz)
            acc += wt * np.sum(sw * c * num * k / rho2 ** ((n + 2) / 2.0))
        out.append(acc)
    return np.array(out)


# ---------------------------------------------------------------------------
# band means over the triangles of a faceted mesh
# ---------------------------------------------------------------------------
def _triangle_points(k=8):
    """Barycentric points of a uniform subdivision of the triangle into k^2
    congruent sub-triangles (their centroids): an equal-weight quadrature."""
    pts = []
    for i in range(k):
        for j in range(k - i):
            pts.append(((i + 1.0 / 3.0) / k, (j + 1.0 / 3.0) / k))
            if i + j < k - 1:
                pts.append(((i + 2.0 / 3.0) / k, (j + 2.0 / 3.0) / k))
    return np.array(pts)


def hole_wall_mask(verts, tris, radius, depth):
    """Wall triangles of a hole mesh (centroid strictly between the opening
    and the floor, on the wall radius) and their centroid depth below the
    opening, as the Monte Carlo band averages select them."""
    v = np.asarray(verts, dtype=np.float64)[np.asarray(tris)]
    c = v.mean(1)
    top = float(np.asarray(verts)[:, 2].max())
    d = top - c[:, 2]
    wall = (d > 1e-6) & (d < depth - 1e-6) & (np.abs(np.hypot(c[:, 0], c[:, 1]) - radius) < 0.01)
    return wall, d


def hole_mesh_band_flux(verts, tris, radius, depth, sticking, cos_exponent, edges, n_wall=2000, n_floor=20, k=8):
    """Mean incident flux per depth band over the wall triangles of a hole
    mesh, as the Monte Carlo measures it: each triangle belongs to the band
    of its centroid depth, band means are weighted by triangle area, and a
    triangle's flux is the mean of the radiosity solution over its surface
    (equal-weight quadrature, linear interpolation between the solution's
    elements). Mesh rows do not line up with the band edges, so this differs
    from the solution's mean over [edge_k, edge_k+1] by up to 0.4 % where the
    flux is steep."""
    ref = solve_hole(radius, depth, sticking, cos_exponent=cos_exponent, n_wall=n_wall, n_floor=n_floor)
    zc = 0.5 * (ref["z"][:-1] + ref["z"][1:])
    wall, d = hole_wall_mask(verts, tris, radius, depth)
    v = np.asarray(verts, dtype=np.float64)[np.asarray(tris)][wall]
    top = float(np.asarray(verts)[:, 2].max())
    area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    p = _triangle_points(k)
    zq = top - (v[:, 0, 2][:, None] * (1 - p[:, 0] - p[:, 1])[None] + v[:, 1, 2][:, None] * p[:, 0][None]
                + v[:, 2, 2][:, None] * p[:, 1][None])
    tri = np.interp(zq, zc, ref["wall"]).mean(1)
    dw = d[wall]
    out = []
    for b in range(len(edges) - 1):
        m = (dw >= edges[b]) & (dw < edges[b + 1])
        out.append(float(np.sum(tri[m] * area[m]) / area[m].sum()))
    return np.array(out)
