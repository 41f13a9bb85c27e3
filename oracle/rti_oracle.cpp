// CPU oracle for the rti Monte Carlo flux path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product (librti_hip.so, include/rti,
// rti_amd) links, loads or calls this file. Only tests/, the smoke check and
// bench.py's cpu_baseline leg use it, and only as the checker or as the timed
// CPU baseline ("restated reference CPU path").
//
// It restates, in plain scalar C++ (OpenMP over rays), the algorithm of the
// reference's tracer (src/rti/trace/tracer.hpp:66-330) and the functions it
// calls. Each function below cites the reference file:line it follows.
// Random numbers are Philox4x32-10 (Salmon et al., SC'11) keyed by
// (seed, instance) with counter (ray id, event index); the reference uses
// per-thread mt19937_64 streams (tracer.hpp:136-143), so the oracle matches
// the reference in distribution, not stream.
//
// Parity pins (see the tests that exercise each function):
//   philox          : Random123 published known-answer vectors
//   local_intersect : the 9 KATs of test/rti/trace/local_intersector.cpp:14-129
//   area_inside     : the KATs of test/rti/geo/disc_bounding_box_intersector.cpp,
//                     closed forms, and the compiled reference header (oracle/_ref)
//   neighbors       : brute force over all pairs, and the compiled reference
//                     divide-and-conquer (oracle/_ref)
//   trace           : moment checks of the samplers, analytic flux on a flat
//                     plate, the radiosity fixture of the reference
//                     (verification/radiosity), and the (1-s)^k path-weight identity.
//
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off). Contraction is
// off so that FP32 results do not depend on whether the host has FMA.

#include <cmath>
#include <limits>
#include <cstdint>
#include <cstring>
#include <cfloat>
#include <vector>
#include <algorithm>
#include <chrono>
#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11). Multipliers and Weyl constants are the published ones.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t v[4]; };

inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

inline u32x4 philox(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(M0, c.v[0], hi0, lo0);
    mulhilo(M1, c.v[2], hi1, lo1);
    u32x4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// [0,1) with 24 random bits, exactly representable in float.
inline float u01(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }
// (0,1] with 24 random bits.
inline float u01_open0(uint32_t x) { return (float)((x >> 8) + 1u) * 5.9604644775390625e-08f; }

template <typename R> struct v3 { R x, y, z; };
template <typename R> inline v3<R> mk(R x, R y, R z) { return {x, y, z}; }
template <typename R> inline v3<R> add(v3<R> a, v3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> inline v3<R> sub(v3<R> a, v3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> inline v3<R> mul(v3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> inline R dot(v3<R> a, v3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> inline v3<R> cross(v3<R> a, v3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R> inline R len(v3<R> a) { return std::sqrt(dot(a, a)); }
template <typename R> inline v3<R> normalize(v3<R> a) { R l = len(a); return {a.x / l, a.y / l, a.z / l}; }

// ---------------------------------------------------------------------------
// Neighbour-disc test. Restates local_intersector::intersect
// (src/rti/trace/local_intersector.hpp:41-97): front face only (d.n <= 0,
// :58), not parallel (|d.n| >= 1e-9, :65-66), t > 0 (:77), and the hit point
// strictly inside the radius (:93). No tnear/tfar bound.
// ---------------------------------------------------------------------------
template <typename R>
inline bool local_intersect(v3<R> o, v3<R> d, v3<R> c, R radius, v3<R> n) {
  R pod = n.x * d.x + n.y * d.y + n.z * d.z;
  if (pod > 0) return false;
  if (std::fabs(pod) < (R)1e-9) return false;
  R ddneg = c.x * n.x + c.y * n.y + c.z * n.z;
  R ttnom = -(n.x * o.x + n.y * o.y + n.z * o.z) + ddneg;
  R tt = ttnom / pod;
  if (tt <= 0) return false;
  v3<R> h = {o.x + d.x * tt, o.y + d.y * tt, o.z + d.z * tt};
  v3<R> e = {h.x - c.x, h.y - c.y, h.z - c.z};
  R dist = std::sqrt(e.x * e.x + e.y * e.y + e.z * e.z);
  return radius > dist;
}

// ---------------------------------------------------------------------------
// Exposed disc area inside the x-y box. Restates
// disc_bounding_box_intersector::area_inside
// (src/rti/geo/disc_bounding_box_intersector.hpp:42-76) with its helpers:
// per-face closest approach (:372-441), circular segments per face (:116-125)
// and the corner-overlap correction for adjacent faces (:127-261). Faces are
// visited in the reference's order right, bottom, left, top (:349-356).
// ---------------------------------------------------------------------------
inline float face_approach(int f, float xmin, float ymin, float xmax, float ymax,
                           const float c[3], float r, const float n[3]) {
  // f: 0 right (x = xmax), 1 bottom (y = ymin), 2 left (x = xmin), 3 top (y = ymax).
  // The reference maps every face onto "x = high" by a swap/reflection of
  // coordinates; this evaluates the same expression directly per face.
  float pos, plane, nperp;  // disc coordinate across the face, face coordinate, other in-plane normal component
  switch (f) {
    case 0: pos = c[0];  plane = xmax;  nperp = n[1]; break;
    case 1: pos = -c[1]; plane = -ymin; nperp = n[0]; break;
    case 2: pos = -c[0]; plane = -xmin; nperp = n[1]; break;
    default: pos = c[1]; plane = ymax;  nperp = n[0]; break;
  }
  float nz = n[2];
  float xterm = r * std::sqrt(nz * nz + nperp * nperp);
  if (pos + xterm <= plane) return FLT_MAX;        // inside w.r.t. this face
  if (pos - xterm >= plane) return -FLT_MAX;       // outside w.r.t. this face
  if (xterm <= 1e-9f) return FLT_MAX;              // disc parallel to the face plane
  return (plane - pos) * r / xterm;
}

float area_inside(float xmin, float ymin, float xmax, float ymax, const float disc[4], const float nrm_in[3]) {
  float c[3] = {disc[0], disc[1], disc[2]};
  float r = disc[3];
  float nl = std::sqrt(nrm_in[0] * nrm_in[0] + nrm_in[1] * nrm_in[1] + nrm_in[2] * nrm_in[2]);
  float n[3] = {nrm_in[0] / nl, nrm_in[1] / nl, nrm_in[2] / nl};
  if (r <= 0) return 0;
  const float pi_f = (float)M_PI;
  float full = r * r * pi_f;
  if ((xmin <= c[0] - r && c[0] + r <= xmax) && (ymin <= c[1] - r && c[1] + r <= ymax)) return full;
  if ((c[0] + r <= xmin || xmax <= c[0] - r) || (c[1] + r <= ymin || ymax <= c[1] - r)) return 0;
  float a[4];
  for (int f = 0; f < 4; ++f) {
    a[f] = face_approach(f, xmin, ymin, xmax, ymax, c, r, n);
    if (a[f] < -r) return 0;
  }
  // Inward normals and the corner each face contributes as the second face of
  // an adjacent pair (right->bottom: bottom-right corner, bottom->left:
  // bottom-left, left->top: top-left, top->right: top-right).
  const float fn[4][2] = {{-1, 0}, {0, 1}, {1, 0}, {0, -1}};
  const float fc[4][2] = {{xmax, ymax}, {xmax, ymin}, {xmin, ymin}, {xmin, ymax}};
  float area = 0;
  for (int f = 0; f < 4; ++f) {
    if (-r < a[f] && a[f] < r) {
      double ang = 2 * std::acos((double)a[f] / r);
      double seg = r * r / 2 * (ang - std::sin(ang));
      area += seg;
    }
  }
  v3<float> cp = {c[0], c[1], c[2]};
  v3<float> dn = {n[0], n[1], n[2]};
  for (int f = 0; f < 4; ++f) {
    int g = (f + 1) % 4;
    float d1 = a[f], d2 = a[g];
    if (!(-r < d1 && d1 < r && -r < d2 && d2 < r)) continue;
    v3<float> n1 = {fn[f][0], fn[f][1], 0.0f};
    v3<float> n2 = {fn[g][0], fn[g][1], 0.0f};
    v3<float> i1 = normalize(cross(dn, n1));
    v3<float> i2 = normalize(cross(dn, n2));
    if (dot(i1, n2) >= 0) i1 = mul(i1, -1.0f);
    if (dot(i2, n1) >= 0) i2 = mul(i2, -1.0f);
    float xi = fc[g][0], yi = fc[g][1];
    float zi = (dn.x * cp.x + dn.y * cp.y + dn.z * cp.z - dn.x * xi - dn.y * yi) / dn.z;
    v3<float> ip = {xi, yi, zi};
    if (len(sub(cp, ip)) >= r) continue;
    float ca1 = dot(sub(cp, ip), i1);
    v3<float> q1 = add(add(ip, mul(i1, ca1)), mul(i1, std::sqrt(r * r - d1 * d1)));
    float ca2 = dot(sub(cp, ip), i2);
    v3<float> q2 = add(add(ip, mul(i2, ca2)), mul(i2, std::sqrt(r * r - d2 * d2)));
    v3<float> e1 = sub(q1, cp), e2 = sub(q2, cp);
    float ang = std::acos(dot(e1, e2) / len(e1) / len(e2));
    float seg = r * r / 2 * (ang - std::sin(ang));
    float tri = 0.5f * len(cross(sub(q1, ip), sub(q2, ip)));
    area -= seg + tri;
  }
  return full - area;
}

// ---------------------------------------------------------------------------
// Overlap predicate of the disc neighbourhood. Restates
// disc_neighborhood::check_dist (src/rti/geo/disc_neighborhood.hpp:190-220):
// per-axis rejection with >= sum of radii, then Euclidean distance (FP32,
// sum in x,y,z order) strictly below the sum of radii.
// ---------------------------------------------------------------------------
inline bool overlap(const float* p1, const float* p2) {
  float sr = p1[3] + p2[3];
  if (std::fabs(p2[0] - p1[0]) >= sr) return false;
  if (std::fabs(p1[1] - p2[1]) >= sr) return false;
  if (std::fabs(p1[2] - p2[2]) >= sr) return false;
  float d1 = p1[0] - p2[0], d2 = p1[1] - p2[1], d3 = p1[2] - p2[2];
  float dist = std::sqrt(d1 * d1 + d2 * d2 + d3 * d3);
  return dist < p1[3] + p2[3];
}

// ---------------------------------------------------------------------------
// Scene, BVH (binned SAH, own code) and closest-hit queries.
// ---------------------------------------------------------------------------
struct aabb {
  float lo[3], hi[3];
};

struct bvh_node {
  aabb box;
  int32_t left;   // internal: index of left child, right = left + 1; leaf: -1
  int32_t first;  // leaf: first index into prim order
  int32_t count;  // leaf: number of prims
};

struct bvh {
  std::vector<bvh_node> nodes;
  std::vector<uint32_t> order;
};

void build_bvh(const std::vector<aabb>& boxes, bvh& out) {
  size_t n = boxes.size();
  out.order.resize(n);
  for (size_t i = 0; i < n; ++i) out.order[i] = (uint32_t)i;
  out.nodes.clear();
  out.nodes.reserve(2 * n + 1);
  std::vector<float> cen(3 * n);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) cen[3 * i + k] = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
  struct item { int node, first, count, depth; };
  std::vector<item> work;
  out.nodes.push_back({});
  work.push_back({0, 0, (int)n, 0});
  const int NB = 16;
  while (!work.empty()) {
    item it = work.back();
    work.pop_back();
    aabb b{{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
    aabb cb = b;
    for (int i = it.first; i < it.first + it.count; ++i) {
      uint32_t p = out.order[i];
      for (int k = 0; k < 3; ++k) {
        b.lo[k] = std::min(b.lo[k], boxes[p].lo[k]);
        b.hi[k] = std::max(b.hi[k], boxes[p].hi[k]);
        cb.lo[k] = std::min(cb.lo[k], cen[3 * p + k]);
        cb.hi[k] = std::max(cb.hi[k], cen[3 * p + k]);
      }
    }
    bvh_node& nd = out.nodes[it.node];
    nd.box = b;
    nd.left = -1;
    nd.first = it.first;
    nd.count = it.count;
    if (it.count <= 4) continue;
    // binned SAH over the widest centroid axis
    int ax = 0;
    float ext[3];
    for (int k = 0; k < 3; ++k) ext[k] = cb.hi[k] - cb.lo[k];
    if (ext[1] > ext[ax]) ax = 1;
    if (ext[2] > ext[ax]) ax = 2;
    if (ext[ax] <= 0) continue;
    int cnt[NB] = {0};
    aabb bb[NB];
    for (int j = 0; j < NB; ++j) bb[j] = {{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
    auto bin_of = [&](uint32_t p) {
      int j = (int)((cen[3 * p + ax] - cb.lo[ax]) / ext[ax] * NB);
      return std::min(std::max(j, 0), NB - 1);
    };
    for (int i = it.first; i < it.first + it.count; ++i) {
      uint32_t p = out.order[i];
      int j = bin_of(p);
      cnt[j]++;
      for (int k = 0; k < 3; ++k) {
        bb[j].lo[k] = std::min(bb[j].lo[k], boxes[p].lo[k]);
        bb[j].hi[k] = std::max(bb[j].hi[k], boxes[p].hi[k]);
      }
    }
    auto area = [](const aabb& x) {
      float d0 = std::max(0.0f, x.hi[0] - x.lo[0]), d1 = std::max(0.0f, x.hi[1] - x.lo[1]),
            d2 = std::max(0.0f, x.hi[2] - x.lo[2]);
      return d0 * d1 + d1 * d2 + d2 * d0;
    };
    float best = FLT_MAX;
    int bestj = -1;
    for (int s = 1; s < NB; ++s) {
      aabb l{{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}}, r = l;
      int nl = 0, nr = 0;
      for (int j = 0; j < s; ++j) {
        nl += cnt[j];
        for (int k = 0; k < 3; ++k) { l.lo[k] = std::min(l.lo[k], bb[j].lo[k]); l.hi[k] = std::max(l.hi[k], bb[j].hi[k]); }
      }
      for (int j = s; j < NB; ++j) {
        nr += cnt[j];
        for (int k = 0; k < 3; ++k) { r.lo[k] = std::min(r.lo[k], bb[j].lo[k]); r.hi[k] = std::max(r.hi[k], bb[j].hi[k]); }
      }
      if (nl == 0 || nr == 0) continue;
      float cost = area(l) * nl + area(r) * nr;
      if (cost < best) { best = cost; bestj = s; }
    }
    int mid;
    if (bestj < 0 || it.depth >= 48) {  // median split keeps the depth bounded
      mid = it.first + it.count / 2;
      std::nth_element(out.order.begin() + it.first, out.order.begin() + mid, out.order.begin() + it.first + it.count,
                       [&](uint32_t a, uint32_t b2) { return cen[3 * a + ax] < cen[3 * b2 + ax]; });
    } else {
      auto pm = std::partition(out.order.begin() + it.first, out.order.begin() + it.first + it.count,
                               [&](uint32_t p) { return bin_of(p) < bestj; });
      mid = (int)(pm - out.order.begin());
    }
    int l = (int)out.nodes.size();
    out.nodes.push_back({});
    out.nodes.push_back({});
    out.nodes[it.node].left = l;
    out.nodes[it.node].count = 0;
    work.push_back({l, it.first, mid - it.first, it.depth + 1});
    work.push_back({l + 1, mid, it.first + it.count - mid, it.depth + 1});
  }
}

template <typename R>
inline bool slab(const aabb& b, v3<R> o, v3<R> inv, R tmin, R tmax) {
  R t0 = ((R)b.lo[0] - o.x) * inv.x, t1 = ((R)b.hi[0] - o.x) * inv.x;
  R lo = std::min(t0, t1), hi = std::max(t0, t1);
  t0 = ((R)b.lo[1] - o.y) * inv.y; t1 = ((R)b.hi[1] - o.y) * inv.y;
  lo = std::max(lo, std::min(t0, t1)); hi = std::min(hi, std::max(t0, t1));
  t0 = ((R)b.lo[2] - o.z) * inv.z; t1 = ((R)b.hi[2] - o.z) * inv.z;
  lo = std::max(lo, std::min(t0, t1)); hi = std::min(hi, std::max(t0, t1));
  return std::max(lo, tmin) <= std::min(hi, tmax);
}

}  // namespace

extern "C" {

struct orc_scene {
  int32_t prim_kind;        // 0 = oriented discs, 1 = triangles
  int32_t pad0;
  uint64_t n_prims;
  const float* xyzr;        // discs: n_prims * 4
  const float* nrm;         // discs: n_prims * 3, unit length
  const float* verts;       // triangles: n_verts * 3
  uint64_t n_verts;
  const uint32_t* tris;     // triangles: n_prims * 3
  const uint32_t* nb_ptr;   // neighbour CSR row pointers (n_prims + 1), discs only; may be null
  const uint32_t* nb_col;   // neighbour CSR columns
  const uint32_t* tie_rank; // optional: of hits at equal distance the primitive of lower rank
                            // wins; null = lower index. Embree leaves such ties to its
                            // traversal order; coplanar overlapping discs make them common.
};

struct orc_params {
  uint64_t n_rays;
  uint64_t ray_begin;       // first global ray id traced (sub-sampling a run)
  uint32_t seed;
  uint32_t instance;
  int32_t bc_x, bc_y;       // 0 = reflective, 1 = periodic
  float sticking;
  float cos_exponent;
  float diffuse_prob;       // probability of a diffuse (vs specular) bounce
  float tnear;
  int32_t roulette;         // 1 = rejection control on (reference behaviour)
  int32_t fp64;             // 1 = trace in double precision
  float box[6];             // xmin, xmax, ymin, ymax, zmin, zmax of the boundary faces
  float src_z;
  int32_t threads;          // 0 = OpenMP default
  int32_t adjoint;          // 1 = also accumulate d/ds and d/dn estimators
  int32_t moments;          // 1 = also accumulate sums of deposit^3 and deposit^4
  int32_t source;           // 0 = rectangle over the box, 1 = disc inscribed in it
};

struct orc_result {
  double* s1;               // sum of deposits per primitive
  double* s2;               // sum of squared deposits per primitive
  uint64_t* cnt;            // number of deposits per primitive
  double* dsum_ds;          // adjoint: d(s1)/ds estimator (may be null)
  double* dsum_dn;          // adjoint: d(s1)/dn estimator (may be null)
  uint64_t segments;        // closest-hit queries
  uint64_t surface_hits;    // front-face primitive hits
  uint64_t misses;
  double seconds;           // wall time of the trace region
  double* s3;               // sum of cubed deposits (moments only; may be null)
  double* s4;               // sum of deposits^4 (moments only; may be null)
};

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  u32x4 c{{ctr[0], ctr[1], ctr[2], ctr[3]}};
  u32x4 r = philox(c, key[0], key[1]);
  std::memcpy(out, r.v, 16);
}

int orc_local_intersect(const float org[3], const float dir[3], const float disc[4], const float nrm[3]) {
  return local_intersect<float>({org[0], org[1], org[2]}, {dir[0], dir[1], dir[2]}, {disc[0], disc[1], disc[2]},
                                disc[3], {nrm[0], nrm[1], nrm[2]})
             ? 1
             : 0;
}

float orc_area_inside(float xmin, float ymin, float xmax, float ymax, const float disc[4], const float nrm[3]) {
  return area_inside(xmin, ymin, xmax, ymax, disc, nrm);
}

// Overlap graph by a uniform grid (cell = 2 * max radius) with the reference's
// predicate. Two passes: pass 1 (col == null) fills row_ptr, pass 2 fills col,
// columns sorted ascending per row. Returns the number of directed edges.
int64_t orc_neighbors(const float* xyzr, uint64_t n, uint32_t* row_ptr, uint32_t* col) {
  if (n == 0) { row_ptr[0] = 0; return 0; }
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX}, rmax = 0;
  for (uint64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], xyzr[4 * i + k]); hi[k] = std::max(hi[k], xyzr[4 * i + k]); }
    rmax = std::max(rmax, xyzr[4 * i + 3]);
  }
  double cell = std::max(2.0 * rmax, 1e-6);
  long dim[3];
  for (int k = 0; k < 3; ++k) dim[k] = std::max(1L, std::min(1L << 20, (long)((hi[k] - lo[k]) / cell) + 1));
  // guard against huge grids: coarsen until the cell count is moderate
  while ((double)dim[0] * dim[1] * dim[2] > 8.0 * n + 1024) {
    cell *= 1.5;
    for (int k = 0; k < 3; ++k) dim[k] = std::max(1L, (long)((hi[k] - lo[k]) / cell) + 1);
  }
  auto cell_of = [&](uint64_t i, int k) {
    long c = (long)((xyzr[4 * i + k] - lo[k]) / cell);
    return std::min(std::max(c, 0L), dim[k] - 1);
  };
  size_t ncell = (size_t)dim[0] * dim[1] * dim[2];
  std::vector<uint32_t> start(ncell + 1, 0), items(n);
  for (uint64_t i = 0; i < n; ++i) start[(cell_of(i, 2) * dim[1] + cell_of(i, 1)) * dim[0] + cell_of(i, 0) + 1]++;
  for (size_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
  std::vector<uint32_t> fill(start.begin(), start.end() - 1);
  for (uint64_t i = 0; i < n; ++i) items[fill[(cell_of(i, 2) * dim[1] + cell_of(i, 1)) * dim[0] + cell_of(i, 0)]++] = (uint32_t)i;
  std::vector<uint32_t> deg(n, 0);
  auto visit = [&](uint64_t i, bool write) {
    long cx = cell_of(i, 0), cy = cell_of(i, 1), cz = cell_of(i, 2);
    uint32_t w = write ? row_ptr[i] : 0;
    uint32_t d = 0;
    std::vector<uint32_t> found;
    for (long z = std::max(0L, cz - 1); z <= std::min(dim[2] - 1, cz + 1); ++z)
      for (long y = std::max(0L, cy - 1); y <= std::min(dim[1] - 1, cy + 1); ++y)
        for (long x = std::max(0L, cx - 1); x <= std::min(dim[0] - 1, cx + 1); ++x) {
          size_t c = (z * dim[1] + y) * dim[0] + x;
          for (uint32_t q = start[c]; q < start[c + 1]; ++q) {
            uint32_t j = items[q];
            if (j == i) continue;
            // the reference tests the pair with the lower coordinate first
            const float* a = &xyzr[4 * i];
            const float* b = &xyzr[4 * (uint64_t)j];
            if (overlap(a, b)) {
              if (write) found.push_back(j);
              ++d;
            }
          }
        }
    if (write) {
      std::sort(found.begin(), found.end());
      for (uint32_t j : found) col[w++] = j;
    }
    return d;
  };
  if (col == nullptr) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint64_t i = 0; i < n; ++i) deg[i] = visit(i, false);
    row_ptr[0] = 0;
    for (uint64_t i = 0; i < n; ++i) row_ptr[i + 1] = row_ptr[i] + deg[i];
    return row_ptr[n];
  }
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t i = 0; i < n; ++i) visit(i, true);
  return row_ptr[n];
}

// Neighbourhood-smoothed, max-normalised estimate. Restates
// device::extract_mc_estimates_normalized_smoothed (src/rti/device.hpp:153-187).
void orc_smooth_normalize(const double* values, const float* areas, const uint32_t* nb_ptr, const uint32_t* nb_col,
                          uint64_t n, float* out) {
  std::vector<double> v(n);
  double maxv = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    double vv = values[i] / areas[i];
    uint32_t b = nb_ptr ? nb_ptr[i] : 0, e = nb_ptr ? nb_ptr[i + 1] : 0;
    for (uint32_t q = b; q < e; ++q) vv += values[nb_col[q]] / areas[nb_col[q]];
    vv /= (double)(e - b + 1);
    v[i] = vv;
    if (maxv < vv) maxv = vv;
  }
  for (uint64_t i = 0; i < n; ++i) out[i] = (float)(v[i] / maxv);
}

}  // extern "C"

namespace {

template <typename R>
struct scene_view {
  int kind;
  uint64_t n;
  std::vector<v3<R>> c;   // disc centre / triangle v0
  std::vector<R> rad;     // disc radius
  std::vector<v3<R>> nrm; // unit normal (disc normal / triangle winding normal)
  std::vector<v3<R>> e1, e2;
  const uint32_t* nb_ptr;
  const uint32_t* nb_col;
  const uint32_t* rank;   // tie order (null = index order)
  bvh tree;
};

template <typename R>
void make_scene(const orc_scene* s, scene_view<R>& v) {
  v.kind = s->prim_kind;
  v.n = s->n_prims;
  v.rank = s->tie_rank;
  v.nb_ptr = s->nb_ptr;
  v.nb_col = s->nb_col;
  v.c.resize(v.n);
  v.nrm.resize(v.n);
  std::vector<aabb> boxes(v.n);
  if (v.kind == 0) {
    v.rad.resize(v.n);
    for (uint64_t i = 0; i < v.n; ++i) {
      const float* p = &s->xyzr[4 * i];
      const float* q = &s->nrm[3 * i];
      v.c[i] = {(R)p[0], (R)p[1], (R)p[2]};
      v.rad[i] = (R)p[3];
      v.nrm[i] = {(R)q[0], (R)q[1], (R)q[2]};
      for (int k = 0; k < 3; ++k) {
        float nk = q[k];
        float ext = p[3] * std::sqrt(std::max(0.0f, 1.0f - nk * nk)) + p[3] * 1e-5f;
        boxes[i].lo[k] = p[k] - ext;
        boxes[i].hi[k] = p[k] + ext;
      }
    }
  } else {
    v.e1.resize(v.n);
    v.e2.resize(v.n);
    for (uint64_t i = 0; i < v.n; ++i) {
      const uint32_t* t = &s->tris[3 * i];
      v3<R> a = {(R)s->verts[3 * t[0]], (R)s->verts[3 * t[0] + 1], (R)s->verts[3 * t[0] + 2]};
      v3<R> b = {(R)s->verts[3 * t[1]], (R)s->verts[3 * t[1] + 1], (R)s->verts[3 * t[1] + 2]};
      v3<R> c = {(R)s->verts[3 * t[2]], (R)s->verts[3 * t[2] + 1], (R)s->verts[3 * t[2] + 2]};
      v.c[i] = a;
      v.e1[i] = sub(b, a);
      v.e2[i] = sub(c, a);
      v.nrm[i] = normalize(cross(v.e1[i], v.e2[i]));
      for (int k = 0; k < 3; ++k) {
        float x0 = s->verts[3 * t[0] + k], x1 = s->verts[3 * t[1] + k], x2 = s->verts[3 * t[2] + k];
        float lo = std::min(x0, std::min(x1, x2)), hi = std::max(x0, std::max(x1, x2));
        float pad = 1e-6f * std::max(1.0f, std::max(std::fabs(lo), std::fabs(hi)));
        boxes[i].lo[k] = lo - pad;
        boxes[i].hi[k] = hi + pad;
      }
    }
  }
  build_bvh(boxes, v.tree);
}

// Closest primitive hit with t in [tnear, tmax]; ties resolved to the lower id.
template <typename R>
inline bool closest_hit(const scene_view<R>& s, v3<R> o, v3<R> d, R tnear, R& tbest, uint32_t& pbest) {
  v3<R> inv;
  const R tiny = (R)1e-20;
  inv.x = (R)1 / (std::fabs(d.x) > tiny ? d.x : std::copysign(tiny, d.x));
  inv.y = (R)1 / (std::fabs(d.y) > tiny ? d.y : std::copysign(tiny, d.y));
  inv.z = (R)1 / (std::fabs(d.z) > tiny ? d.z : std::copysign(tiny, d.z));
  bool found = false;
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  const bvh& t = s.tree;
  while (sp > 0) {
    const bvh_node& nd = t.nodes[stack[--sp]];
    if (!slab(nd.box, o, inv, tnear, tbest)) continue;
    if (nd.left >= 0) {
      stack[sp++] = nd.left + 1;
      stack[sp++] = nd.left;
      continue;
    }
    for (int q = nd.first; q < nd.first + nd.count; ++q) {
      uint32_t p = t.order[q];
      R th;
      if (s.kind == 0) {
        v3<R> n = s.nrm[p];
        R dn = dot(d, n);
        if (dn == 0) continue;
        th = dot(sub(s.c[p], o), n) / dn;
        if (!(th >= tnear && th <= tbest)) continue;
        v3<R> h = sub(add(o, mul(d, th)), s.c[p]);
        if (!(dot(h, h) < s.rad[p] * s.rad[p])) continue;
      } else {
        // Moller-Trumbore (J. Graphics Tools 2(1), 1997)
        v3<R> pv = cross(d, s.e2[p]);
        R det = dot(s.e1[p], pv);
        if (det == 0) continue;
        R idet = (R)1 / det;
        v3<R> tv = sub(o, s.c[p]);
        R u = dot(tv, pv) * idet;
        if (u < 0 || u > 1) continue;
        v3<R> qv = cross(tv, s.e1[p]);
        R vv = dot(d, qv) * idet;
        if (vv < 0 || u + vv > 1) continue;
        th = dot(s.e2[p], qv) * idet;
        if (!(th >= tnear && th <= tbest)) continue;
      }
      if (th < tbest || (th == tbest && (!found || (s.rank ? s.rank[p] < s.rank[pbest] : p < pbest)))) {
        tbest = th;
        pbest = p;
        found = true;
      }
    }
  }
  return found;
}

// Boundary faces x = xmin/xmax and y = ymin/ymax spanning z in [zmin, zmax]
// (src/rti/geo/boundary_x_y.hpp:147-235, two triangles per face, no top or
// bottom). Returns the face (0 xmax, 1 ymin, 2 xmin, 3 ymax) or -1.
template <typename R>
inline int boundary_hit(const float* box, v3<R> o, v3<R> d, R tnear, R& tb) {
  int face = -1;
  R best = (R)FLT_MAX;
  const R xmin = box[0], xmax = box[1], ymin = box[2], ymax = box[3], zmin = box[4], zmax = box[5];
  auto test = [&](int f, R t, R a, R alo, R ahi, R z) {
    if (t >= tnear && t <= best && a >= alo && a <= ahi && z >= zmin && z <= zmax) {
      best = t;
      face = f;
    }
  };
  if (d.x != 0) {
    R t = (xmax - o.x) / d.x;
    test(0, t, o.y + d.y * t, ymin, ymax, o.z + d.z * t);
    t = (xmin - o.x) / d.x;
    test(2, t, o.y + d.y * t, ymin, ymax, o.z + d.z * t);
  }
  if (d.y != 0) {
    R t = (ymin - o.y) / d.y;
    test(1, t, o.x + d.x * t, xmin, xmax, o.z + d.z * t);
    t = (ymax - o.y) / d.y;
    test(3, t, o.x + d.x * t, xmin, xmax, o.z + d.z * t);
  }
  tb = best;
  return face;
}

template <typename R>
inline v3<R> source_dir(float r1, float r2, float n, float& lncos) {
  // power_cosine_direction_z::get (src/rti/ray/power_cosine_direction_z.hpp:17-32),
  // cosine_direction_z::get (src/rti/ray/cosine_direction_z.hpp:13-27) for n = 1.
  R tt = (n == 1.0f) ? (R)r2 : (R)std::pow((R)r2, (R)2 / ((R)n + 1));
  R two_pi = (R)(2.0 * M_PI);
  R st = std::sqrt((R)1 - tt);
  v3<R> d = {std::cos(two_pi * (R)r1) * st, std::sin(two_pi * (R)r1) * st, -std::sqrt(tt)};
  lncos = std::log(r2) / (n + 1.0f);
  return normalize(d);
}

// Orthonormal basis around a unit vector: Duff et al., "Building an
// Orthonormal Basis, Revisited", JCGT 6(1), 2017. The reference picks a
// different (deterministic) perpendicular (src/rti/util/utils.hpp:207-244);
// the cosine lobe it generates is the same distribution.
template <typename R>
inline void onb(v3<R> n, v3<R>& b1, v3<R>& b2) {
  R sign = std::copysign((R)1, n.z);
  R a = (R)-1 / (sign + n.z);
  R b = n.x * n.y * a;
  b1 = {(R)1 + sign * n.x * n.x * a, sign * b, -sign * n.x};
  b2 = {b, sign + n.y * n.y * a, -n.y};
}

template <typename R>
inline v3<R> diffuse_dir(v3<R> n, float r1, float r2) {
  // cos_hemi::get (src/rti/ray/cos_hemi.hpp:10-45)
  v3<R> b1, b2;
  onb(n, b1, b2);
  R two_pi = (R)(2.0 * M_PI);
  R c1 = std::sqrt((R)r2);
  R st = std::sqrt((R)1 - (R)r2);
  R c2 = std::cos(two_pi * (R)r1) * st;
  R c3 = std::sin(two_pi * (R)r1) * st;
  return normalize(add(add(mul(n, c1), mul(b1, c2)), mul(b2, c3)));
}

template <typename R>
inline v3<R> specular_dir(v3<R> d, v3<R> n) {
  // specular::use (src/rti/reflection/specular.hpp:16-40)
  return sub(d, mul(n, (R)2 * dot(d, n)));
}

// Safety bound on segments per ray; a ray that reaches it is dropped (it
// could only be an exactly horizontal ray bouncing between boundary faces).
constexpr uint32_t kMaxSegments = 1u << 20;

struct thread_acc {
  std::vector<double> s1, s2, s3, s4, ds, dn;
  std::vector<uint64_t> cnt;
  uint64_t segments = 0, hits = 0, misses = 0;
};

template <typename R>
v3<R> source_origin(const orc_params& p, const u32x4& u, uint32_t clo, uint32_t chi) {
  const R xmin = p.box[0], xmax = p.box[1], ymin = p.box[2], ymax = p.box[3];
  if (p.source == 0) {  // rectangle_origin_z::get (src/rti/ray/rectangle_origin_z.hpp:23-37)
    return {xmin + (xmax - xmin) * (R)u01(u.v[0]), ymin + (ymax - ymin) * (R)u01(u.v[1]), (R)p.src_z};
  }
  // disc_origin_z::get (src/rti/ray/disc_origin_z.hpp:17-33) on the disc of
  // device::create_circular_source_from_bounding_box (src/rti/device.hpp:269-280):
  // centre of the box, radius half its x extent; offsets uniform in [-r, r]^2,
  // redrawn until inside the disc. Attempt a > 0 draws the Philox block
  // (ray, 0, a); after 64 rejections the ray starts at the centre.
  const R cx = (xmin + xmax) * (R)0.5, cy = (ymin + ymax) * (R)0.5, rad = (xmax - xmin) * (R)0.5;
  uint32_t ux = u.v[0], uy = u.v[1];
  for (uint32_t a = 1;; ++a) {
    const R dx = ((R)2 * (R)u01(ux) - (R)1) * rad;
    const R dy = ((R)2 * (R)u01(uy) - (R)1) * rad;
    if (std::sqrt(std::fma(dx, dx, dy * dy)) <= rad) return {cx + dx, cy + dy, (R)p.src_z};
    if (a == 64) return {cx, cy, (R)p.src_z};
    const u32x4 v = philox({{clo, chi, 0u, a}}, p.seed, p.instance);
    ux = v.v[0];
    uy = v.v[1];
  }
}

template <typename R>
void trace_range(const scene_view<R>& s, const orc_params& p, thread_acc& acc, uint64_t r0, uint64_t r1) {
  const R tnear = (R)p.tnear;
  const R xmin = p.box[0], xmax = p.box[1], ymin = p.box[2], ymax = p.box[3];
  const float sk = p.sticking;
  const float nexp = p.cos_exponent;
  auto deposit = [&](uint32_t prim, R drop, double dds, double ddn) {
    acc.s1[prim] += (double)drop;
    acc.s2[prim] += (double)drop * (double)drop;
    if (p.moments) {  // hit_accumulator::use (src/rti/trace/hit_accumulator.hpp:128-138)
      acc.s3[prim] += (double)drop * (double)drop * (double)drop;
      acc.s4[prim] += (double)drop * (double)drop * (double)drop * (double)drop;
    }
    acc.cnt[prim] += 1;
    if (p.adjoint) {
      acc.ds[prim] += dds;
      acc.dn[prim] += ddn;
    }
  };
  for (uint64_t ray = r0; ray < r1; ++ray) {
    const uint32_t clo = (uint32_t)ray, chi = (uint32_t)(ray >> 32);
    u32x4 u = philox({{clo, chi, 0u, 0u}}, p.seed, p.instance);
    uint32_t event = 1;
    v3<R> o = source_origin<R>(p, u, clo, chi);
    float lncos;
    v3<R> d = source_dir<R>(u01(u.v[2]), u01_open0(u.v[3]), nexp, lncos);
    const double score_n = 1.0 / (nexp + 1.0) + (double)lncos;
    R w = 1;
    int k = 0;           // front hits so far
    double rfac = 1.0;   // product of roulette boosts
    for (uint32_t seg = 0; seg < kMaxSegments; ++seg) {
      acc.segments++;
      R tb;
      int face = boundary_hit(p.box, o, d, tnear, tb);
      R tp = tb;
      uint32_t prim = 0;
      bool hitp = closest_hit(s, o, d, tnear, tp, prim);
      if (!hitp && face < 0) {  // tracer.hpp:184-190
        acc.misses++;
        break;
      }
      if (!hitp) {  // boundary (tracer.hpp:193-209, boundary_x_y.hpp:91-143)
        v3<R> h = add(o, mul(d, tb));
        if (face == 0 || face == 2) {
          if (p.bc_x == 0) { d.x = -d.x; o = h; }
          else { o = {face == 0 ? xmin : xmax, h.y, h.z}; }
        } else {
          if (p.bc_y == 0) { d.y = -d.y; o = h; }
          else { o = {h.x, face == 1 ? ymax : ymin, h.z}; }
        }
        continue;
      }
      v3<R> n = s.nrm[prim];
      if (dot(d, n) > 0) {  // back face: pass through (tracer.hpp:212-227)
        o = add(o, mul(d, tp));
        continue;
      }
      acc.hits++;
      // stick and deposit (tracer.hpp:233-237)
      R drop = w * (R)sk;
      double dds = 0, ddn = 0;
      if (p.adjoint) {
        // d/ds of s (1-s)^k times the roulette boosts of this path
        double sd = sk, om = 1.0 - sd;
        double g = std::pow(om, k) - (k > 0 ? k * sd * std::pow(om, k - 1) : 0.0);
        dds = g * rfac;
        ddn = (double)drop * score_n;
      }
      deposit(prim, drop, dds, ddn);
      if (s.kind == 0 && s.nb_ptr) {  // tracer.hpp:339-383
        for (uint32_t q = s.nb_ptr[prim]; q < s.nb_ptr[prim + 1]; ++q) {
          uint32_t j = s.nb_col[q];
          if (local_intersect<R>(o, d, s.c[j], s.rad[j], s.nrm[j])) deposit(j, drop, dds, ddn);
        }
      }
      w -= drop;
      if (w == 0) break;  // tracer.hpp:238-240
      k++;
      u = philox({{clo, chi, event, 0u}}, p.seed, p.instance);
      event++;
      // rejection_control::check_weight_reweight_or_kill (src/rti/mc/rejection_control.hpp:27-61)
      if (p.roulette && w < (R)0.1) {
        R kill = (R)1 - w / (R)0.3;
        if (u01(u.v[0]) < kill) break;
        rfac *= (double)((R)0.3 / w);
        w = (R)0.3;
      }
      o = add(o, mul(d, tp));  // meta_geometry::get_new_origin (src/rti/geo/meta_geometry.hpp:17-23)
      // cos^2 from (0, 1]: a draw of 0 would leave the ray in the tangent
      // plane of the surface it left (cos_hemi.hpp:25-26 draws 64 bits)
      if (u01(u.v[3]) < p.diffuse_prob) d = diffuse_dir<R>(n, u01(u.v[1]), u01_open0(u.v[2]));
      else d = specular_dir<R>(d, n);
    }
  }
}

template <typename R>
int trace_impl(const orc_scene* sc, const orc_params* p, orc_result* res) {
  scene_view<R> s;
  make_scene<R>(sc, s);
  uint64_t n = sc->n_prims;
  int nt = p->threads > 0 ? p->threads : omp_get_max_threads();
  std::vector<thread_acc> accs(nt);
  auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(nt)
  {
    int tid = omp_get_thread_num();
    thread_acc& a = accs[tid];
    a.s1.assign(n, 0.0);
    a.s2.assign(n, 0.0);
    a.cnt.assign(n, 0);
    if (p->adjoint) { a.ds.assign(n, 0.0); a.dn.assign(n, 0.0); }
    if (p->moments) { a.s3.assign(n, 0.0); a.s4.assign(n, 0.0); }
    const uint64_t chunk = 256;
    uint64_t nchunks = (p->n_rays + chunk - 1) / chunk;
#pragma omp for schedule(dynamic, 4)
    for (uint64_t c = 0; c < nchunks; ++c) {
      uint64_t b = p->ray_begin + c * chunk;
      uint64_t e = std::min(p->ray_begin + p->n_rays, b + chunk);
      trace_range<R>(s, *p, a, b, e);
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  res->seconds = std::chrono::duration<double>(t1 - t0).count();
  // combine in thread order (tracer.hpp:110-114 combines per-thread copies)
  res->segments = res->surface_hits = res->misses = 0;
  for (uint64_t i = 0; i < n; ++i) {
    double a1 = 0, a2 = 0, a3 = 0, a4 = 0, m3 = 0, m4 = 0;
    uint64_t c = 0;
    for (int t = 0; t < nt; ++t) {
      a1 += accs[t].s1[i];
      a2 += accs[t].s2[i];
      c += accs[t].cnt[i];
      if (p->adjoint) { a3 += accs[t].ds[i]; a4 += accs[t].dn[i]; }
      if (p->moments) { m3 += accs[t].s3[i]; m4 += accs[t].s4[i]; }
    }
    if (p->moments && res->s3) res->s3[i] = m3;
    if (p->moments && res->s4) res->s4[i] = m4;
    if (res->s1) res->s1[i] = a1;
    if (res->s2) res->s2[i] = a2;
    if (res->cnt) res->cnt[i] = c;
    if (p->adjoint && res->dsum_ds) res->dsum_ds[i] = a3;
    if (p->adjoint && res->dsum_dn) res->dsum_dn[i] = a4;
  }
  for (int t = 0; t < nt; ++t) {
    res->segments += accs[t].segments;
    res->surface_hits += accs[t].hits;
    res->misses += accs[t].misses;
  }
  return 0;
}

}  // namespace

extern "C" {

// Samplers, exposed for the distribution and CRN tests.
// Primary-ray origins of rays [r0, r0 + n) (xyz per ray) for params p.
void orc_source_origins(const orc_params* p, uint64_t r0, uint64_t n, float* out) {
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t ray = r0 + i;
    const uint32_t clo = (uint32_t)ray, chi = (uint32_t)(ray >> 32);
    const u32x4 u = philox({{clo, chi, 0u, 0u}}, p->seed, p->instance);
    const v3<float> o = source_origin<float>(*p, u, clo, chi);
    out[3 * i] = o.x; out[3 * i + 1] = o.y; out[3 * i + 2] = o.z;
  }
}

void orc_source_dir(float r1, float r2, float n, float out[3]) {
  float lncos;
  v3<float> d = source_dir<float>(r1, r2, n, lncos);
  out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

// Primary rays [r0, r0 + n): origins and directions exactly as trace_range
// starts them (the rays a run with the same params traces).
void orc_source_rays(const orc_params* p, uint64_t r0, uint64_t n, float* org, float* dir) {
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t ray = r0 + i;
    const uint32_t clo = (uint32_t)ray, chi = (uint32_t)(ray >> 32);
    const u32x4 u = philox({{clo, chi, 0u, 0u}}, p->seed, p->instance);
    const v3<float> o = source_origin<float>(*p, u, clo, chi);
    float lncos;
    const v3<float> d = source_dir<float>(u01(u.v[2]), u01_open0(u.v[3]), p->cos_exponent, lncos);
    org[3 * i] = o.x; org[3 * i + 1] = o.y; org[3 * i + 2] = o.z;
    dir[3 * i] = d.x; dir[3 * i + 1] = d.y; dir[3 * i + 2] = d.z;
  }
}

// Mirror direction (specular::use, src/rti/reflection/specular.hpp:16-40),
// exposed for the invariant and method-of-images tests.
void orc_specular_dir(const float d[3], const float nrm[3], float out[3]) {
  const v3<float> r = specular_dir<float>({d[0], d[1], d[2]}, {nrm[0], nrm[1], nrm[2]});
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// The closest-hit query of one segment, as trace_range makes it, for n
// rays: the boundary faces of p->box bound the primitive query (own BVH).
// Per ray: prim = the closest primitive hit in [tnear, boundary distance]
// (-1 none), face = the boundary face when no primitive is hit (-1 none),
// t = the distance of whichever was returned (FLT_MAX for a miss).
void orc_closest_hits(const orc_scene* sc, const orc_params* p, uint64_t n, const float* org, const float* dir,
                      int32_t* prim, int32_t* face, float* t) {
  scene_view<float> s;
  make_scene<float>(sc, s);
  for (uint64_t i = 0; i < n; ++i) {
    const v3<float> o = {org[3 * i], org[3 * i + 1], org[3 * i + 2]};
    const v3<float> d = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    float tb;
    const int f = boundary_hit(p->box, o, d, p->tnear, tb);
    float tp = tb;
    uint32_t pb = 0;
    const bool hit = closest_hit(s, o, d, p->tnear, tp, pb);
    prim[i] = hit ? (int32_t)pb : -1;
    face[i] = hit ? -1 : f;
    t[i] = hit ? tp : tb;
  }
}

void orc_diffuse_dir(const float nrm[3], float r1, float r2, float out[3]) {
  v3<float> d = diffuse_dir<float>({nrm[0], nrm[1], nrm[2]}, r1, r2);
  out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

// Rejection control decision for weight w and uniform u (returns 1 = continue).
int orc_roulette(float* w, float u) {
  if (*w >= 0.1f) return 1;
  if (u < 1.0f - *w / 0.3f) return 0;
  *w = 0.3f;
  return 1;
}

// Per-primitive relative error and variance of the variance from the power
// sums, with N = all deposits of the run (hit_accumulator::get_relative_error,
// src/rti/trace/hit_accumulator.hpp:152-186, and get_vov, :188-249):
//   rel = sqrt(S2 / S1^2 - 1/N)                     (DBL_MAX where S1 = 0)
//   vov = (S4 - 4 S1 S3/N + 8 S2 S1^2/N^2 - 4 S1^4/N^3 - S2^2/N) N^2 / (N S2 - S1^2)^2
//                                                   (DBL_MAX where the count is 0
//                                                    or the denominator is 0)
// The powers of N are formed in double (the reference forms N^2 and N^3 in
// size_t, which wraps for N > 2.6e6 deposits).
void orc_error_stats(const double* s1, const double* s2, const double* s3, const double* s4, const uint64_t* cnt,
                     uint64_t n, uint64_t total, double* rel, double* vov) {
  const double big = std::numeric_limits<double>::max();
  const double nn = (double)total;
  for (uint64_t i = 0; i < n; ++i) {
    if (rel) {
      const double sq = s1[i] * s1[i];
      rel[i] = (total == 0 || sq == 0) ? big : std::sqrt(s2[i] / sq - 1.0 / nn);
    }
    if (vov) {
      if (cnt[i] == 0) { vov[i] = big; continue; }
      const double a1 = s1[i], a2 = s2[i], a3 = s3[i], a4 = s4[i];
      const double sq = a1 * a1, q4 = sq * sq, a2sq = a2 * a2;
      const double num = a4 - (4 * a1 * a3 / nn) + (8 * a2 * sq / (nn * nn)) - (4 * q4 / (nn * nn * nn)) - (a2sq / nn);
      const double r2 = nn * a2 - sq;
      const double den2 = r2 * r2;
      vov[i] = den2 == 0 ? big : num / den2 * nn * nn;
    }
  }
}

int orc_trace(const orc_scene* sc, const orc_params* p, orc_result* res) {
  if (!sc || !p || !res) return 1;
  if (sc->n_prims == 0) return 2;
  if (sc->prim_kind == 0 && (!sc->xyzr || !sc->nrm)) return 3;
  if (sc->prim_kind == 1 && (!sc->verts || !sc->tris)) return 3;
  return p->fp64 ? trace_impl<double>(sc, p, res) : trace_impl<float>(sc, p, res);
}

// Sources of the fixed-point layout used by the product are not needed here:
// the oracle accumulates in FP64 (hit_accumulator.hpp:128-138 accumulates in
// double as well).

}  // extern "C"
