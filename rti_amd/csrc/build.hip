// Acceleration-structure build for gfx950: Morton codes, an LSD radix sort,
// the Karras (HPG 2012) binary radix tree, an atomic-flag bottom-up refit,
// the disc-overlap graph (CSR) queried through that tree, and the exposed
// disc areas.
//
// Replaces the Embree BVH build (reference src/rti/trace/tracer.hpp:74-100),
// disc_neighborhood::setup_neighborhood (src/rti/geo/disc_neighborhood.hpp:20-220)
// and tracer::compute_disc_areas (tracer.hpp:385-413).

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <utility>
#include <vector>

#include "internal.hpp"

using namespace rti_dev;

namespace rti_impl {
namespace {

constexpr int kScanBlock = 1024;

// ---------------------------------------------------------------------------
// Exclusive scan (uint32) -- three passes: per-block totals, scan of the
// totals by one block, per-block scan plus offset.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* sh, uint32_t& total) {
  // sh: blockDim.x / 64 words
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int w = 0; w < nw; ++w) {
      uint32_t t = sh[w];
      sh[w] = acc;
      acc += t;
    }
    sh[nw] = acc;
  }
  __syncthreads();
  uint32_t r = x - v + sh[wid];
  total = sh[nw];
  __syncthreads();
  return r;
}

__global__ void k_scan_totals(const uint32_t* in, uint32_t n, uint32_t* totals) {
  __shared__ uint32_t sh[kScanBlock / 64 + 1];
  uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
  uint32_t v = i < n ? in[i] : 0u;
  uint32_t tot;
  block_scan_excl(v, sh, tot);
  if (threadIdx.x == 0) totals[blockIdx.x] = tot;
}

__global__ void k_scan_single(uint32_t* data, uint32_t n) {
  __shared__ uint32_t sh[kScanBlock / 64 + 1];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < n; base += kScanBlock) {
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < n ? data[i] : 0u;
    uint32_t tot;
    uint32_t r = block_scan_excl(v, sh, tot);
    if (i < n) data[i] = r + carry;
    carry += tot;
  }
}

__global__ void k_scan_apply(const uint32_t* in, uint32_t* out, uint32_t n, const uint32_t* offsets) {
  __shared__ uint32_t sh[kScanBlock / 64 + 1];
  uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
  uint32_t v = i < n ? in[i] : 0u;
  uint32_t tot;
  uint32_t r = block_scan_excl(v, sh, tot);
  if (i < n) out[i] = r + offsets[blockIdx.x];
}

// ---------------------------------------------------------------------------
// LSD radix sort of (30-bit key, uint32 value), 8 bits per pass, stable.
// ---------------------------------------------------------------------------
constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr uint32_t kRsTile = kRsThreads * kRsItems;

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t* keys, uint32_t n, int shift,
                                                        uint32_t* hist, uint32_t nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint32_t base = blockIdx.x * kRsTile;
  for (int it = 0; it < kRsItems; ++it) {
    uint32_t i = base + it * kRsThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const uint32_t* kin, const uint32_t* vin,
                                                           uint32_t* kout, uint32_t* vout, uint32_t n, int shift,
                                                           const uint32_t* offs, uint32_t nblocks) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcnt[kRsThreads / 64][257];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  base[threadIdx.x] = offs[threadIdx.x * nblocks + blockIdx.x];
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t tile = blockIdx.x * kRsTile;
  for (int it = 0; it < kRsItems; ++it) {
    for (int q = threadIdx.x; q < (kRsThreads / 64) * 257; q += kRsThreads) (&wcnt[0][0])[q] = 0;
    __syncthreads();
    uint32_t i = tile + it * kRsThreads + threadIdx.x;
    bool valid = i < n;
    uint32_t key = valid ? kin[i] : 0u;
    uint32_t val = valid ? vin[i] : 0u;
    uint32_t dig = valid ? ((key >> shift) & 255u) : 256u;
    // lanes of this wave holding the same digit (invalid lanes form their own group)
    unsigned long long same = __ballot(valid) ;
    if (!valid) same = ~same;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      unsigned long long bal = __ballot((dig >> b) & 1u);
      same &= ((dig >> b) & 1u) ? bal : ~bal;
    }
    uint32_t rank = __popcll(same & lt);
    if (rank == 0) wcnt[wid][dig] = __popcll(same);
    __syncthreads();
    if (valid) {
      uint32_t off = base[dig] + rank;
      for (int w = 0; w < wid; ++w) off += wcnt[w][dig];
      kout[off] = key;
      vout[off] = val;
    }
    __syncthreads();
    {
      uint32_t add = 0;
      for (int w = 0; w < kRsThreads / 64; ++w) add += wcnt[w][threadIdx.x];
      base[threadIdx.x] += add;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Primitive setup: Morton code of the centroid.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

struct setup_args {
  int kind;
  uint32_t n;
  const float* xyzr;
  const float* nrm;
  const float* verts;
  const uint32_t* tris;
  float cmin0, cmin1, cmin2, cinv0, cinv1, cinv2;
  float pad;  // absolute padding of every leaf box
};

__device__ __forceinline__ void prim_centroid(const setup_args& a, uint32_t i, float& x, float& y, float& z) {
  if (a.kind == 0) {
    x = a.xyzr[4 * i];
    y = a.xyzr[4 * i + 1];
    z = a.xyzr[4 * i + 2];
  } else {
    const uint32_t* t = a.tris + 3 * (size_t)i;
    const float* v0 = a.verts + 3 * (size_t)t[0];
    const float* v1 = a.verts + 3 * (size_t)t[1];
    const float* v2 = a.verts + 3 * (size_t)t[2];
    x = (v0[0] + v1[0] + v2[0]) * (1.0f / 3.0f);
    y = (v0[1] + v1[1] + v2[1]) * (1.0f / 3.0f);
    z = (v0[2] + v1[2] + v2[2]) * (1.0f / 3.0f);
  }
}

__global__ void k_morton(setup_args a, uint32_t* keys, uint32_t* vals) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float x, y, z;
  prim_centroid(a, i, x, y, z);
  auto q = [](float v, float lo, float inv) {
    float t = (v - lo) * inv;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return (uint32_t)fminf(t * 1024.0f, 1023.0f);
  };
  uint32_t code = (spread10(q(x, a.cmin0, a.cinv0)) << 2) | (spread10(q(y, a.cmin1, a.cinv1)) << 1) |
                  spread10(q(z, a.cmin2, a.cinv2));
  keys[i] = code;
  vals[i] = i;
}

// Stored primitive records in Morton order, the permutation, and leaf boxes.
__global__ void k_scatter_prims(setup_args a, const uint32_t* order, uint32_t prim_begin, float4* prims,
                                uint32_t shift, uint32_t* perm, float4* leaf_lo, float4* leaf_hi) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.n) return;
  uint32_t i = order[j];
  size_t g = (size_t)prim_begin + j;
  perm[g] = i;
  float4* rec = prims + (g << shift);
  float lo[3], hi[3];
  if (a.kind == 0) {
    float4 c = make_float4(a.xyzr[4 * (size_t)i], a.xyzr[4 * (size_t)i + 1], a.xyzr[4 * (size_t)i + 2],
                           a.xyzr[4 * (size_t)i + 3]);
    float n0 = a.nrm[3 * (size_t)i], n1 = a.nrm[3 * (size_t)i + 1], n2 = a.nrm[3 * (size_t)i + 2];
    rec[0] = c;
    rec[1] = make_float4(n0, n1, n2, 0.0f);
    if (shift == kPrimShiftMixed) {
      rec[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      rec[3] = make_float4(n0, n1, n2, 0.0f);
    }
    float nn[3] = {n0, n1, n2};
    float cc[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      // oriented disc extent along axis k: r * sqrt(1 - n_k^2), padded
      float ext = c.w * sqrtf(fmaxf(0.0f, 1.0f - nn[k] * nn[k])) + c.w * 1e-5f + a.pad;
      lo[k] = cc[k] - ext;
      hi[k] = cc[k] + ext;
    }
  } else {
    const uint32_t* t = a.tris + 3 * (size_t)i;
    float v[3][3];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int k = 0; k < 3; ++k) v[q][k] = a.verts[3 * (size_t)t[q] + k];
    rec[0] = make_float4(v[0][0], v[0][1], v[0][2], -1.0f);  // w < 0 marks a triangle
    rec[1] = make_float4(v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2], 0.0f);
    rec[2] = make_float4(v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2], 0.0f);
    {
      // unit normal from the winding, normalize(e1 x e2), in the plain
      // (unfused) single-precision arithmetic of this translation unit
      const float e1x = v[1][0] - v[0][0], e1y = v[1][1] - v[0][1], e1z = v[1][2] - v[0][2];
      const float e2x = v[2][0] - v[0][0], e2y = v[2][1] - v[0][1], e2z = v[2][2] - v[0][2];
      const float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
      const float l = sqrtf(cx * cx + cy * cy + cz * cz);
      rec[3] = make_float4(cx / l, cy / l, cz / l, 0.0f);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float l = fminf(v[0][k], fminf(v[1][k], v[2][k]));
      float h = fmaxf(v[0][k], fmaxf(v[1][k], v[2][k]));
      float pad = fmaxf(a.pad, 1e-6f * fmaxf(1.0f, fmaxf(fabsf(l), fabsf(h))));
      lo[k] = l - pad;
      hi[k] = h + pad;
    }
  }
  leaf_lo[j] = make_float4(lo[0], lo[1], lo[2], 0.0f);
  leaf_hi[j] = make_float4(hi[0], hi[1], hi[2], 0.0f);
}

// ---------------------------------------------------------------------------
// Karras radix tree over the sorted codes (ties broken by sorted position).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int delta(const uint32_t* codes, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = codes[i], b = codes[j];
  if (a == b) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(a ^ b);
}

__global__ void k_karras(const uint32_t* codes, int n, uint32_t prim_begin, uint32_t node_begin, bvh_node* nodes,
                         int32_t* parent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(codes, n, i, i - d);
  int lmax = 2;
  while (delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta(codes, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int gamma = i + s * d + min(d, 0);
  int lo = min(i, j), hi = max(i, j);
  int c0, c1;
  // parent[] local ids: internal nodes 0..n-2, leaves n-1 .. 2n-2
  if (lo == gamma) {
    c0 = ~(int)(prim_begin + gamma);
    parent[(n - 1) + gamma] = i;
  } else {
    c0 = (int)(node_begin + gamma);
    parent[gamma] = i;
  }
  if (hi == gamma + 1) {
    c1 = ~(int)(prim_begin + gamma + 1);
    parent[(n - 1) + gamma + 1] = i;
  } else {
    c1 = (int)(node_begin + gamma + 1);
    parent[gamma + 1] = i;
  }
  nodes[node_begin + i].child = make_int4(c0, c1, 0, 0);
  if (i == 0) parent[0] = -1;
}

__device__ __forceinline__ void write_slot(bvh_node* nd, int slot, float4 lo, float4 hi) {
  if (slot == 0) {
    nd->xy0 = make_float4(lo.x, hi.x, lo.y, hi.y);
    nd->z01.x = lo.z;
    nd->z01.y = hi.z;
  } else {
    nd->xy1 = make_float4(lo.x, hi.x, lo.y, hi.y);
    nd->z01.z = lo.z;
    nd->z01.w = hi.z;
  }
}

// Bottom-up refit: each leaf writes its box into its parent's slot; the
// second arrival at a node (acq_rel counter) merges both slots and climbs.
// Also records the deepest leaf (number of internal ancestors + 1).
__global__ void k_refit(int n, uint32_t prim_begin, uint32_t node_begin, bvh_node* nodes, const int32_t* parent,
                        uint32_t* flags, const float4* leaf_lo, const float4* leaf_hi, uint32_t* max_depth) {
  int leaf = blockIdx.x * blockDim.x + threadIdx.x;
  if (leaf >= n) return;
  float4 lo = leaf_lo[leaf], hi = leaf_hi[leaf];
  int node = parent[(n - 1) + leaf];
  uint32_t depth = 1;
  for (int a = node; a >= 0; a = parent[a]) ++depth;
  atomicMax(max_depth, depth);
  int code = ~(int)(prim_begin + leaf);  // how the parent refers to where we come from
  while (node >= 0) {
    bvh_node* nd = nodes + node_begin + node;
    int4 ch = nd->child;
    write_slot(nd, ch.x == code ? 0 : 1, lo, hi);
    unsigned prev = __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == 0) return;  // the sibling subtree is not finished; it will continue from here
    float4 a = nd->xy0, b = nd->xy1, z = nd->z01;
    lo = make_float4(fminf(a.x, b.x), fminf(a.z, b.z), fminf(z.x, z.z), 0.0f);
    hi = make_float4(fmaxf(a.y, b.y), fmaxf(a.w, b.w), fmaxf(z.y, z.w), 0.0f);
    code = (int)(node_begin + node);
    node = parent[node];
  }
}

// ---------------------------------------------------------------------------
// PLOC (Meister and Bittner, "Parallel Locally-Ordered Clustering for
// Bounding Volume Hierarchy Construction", IEEE TVCG 2018): the clusters
// (initially the leaves in Morton order) each find the cluster within
// kPlocRadius positions whose merged box has the smallest surface area;
// mutual nearest neighbours merge into a new binary node, and the survivors
// are compacted in order. Every iteration merges at least the pair that is
// smallest in the total order (area, lower index, higher index), so the
// loop ends after at most n - 1 iterations. The merges of an iteration with
// c clusters and m merges take node slots c - m - 1 .. c - 2 (in index
// order), so the last merge is node 0, the root; the node boxes are then
// written by the same bottom-up refit as the Karras tree. A radius of 12
// traced the 199k-triangle hole fastest (against 16: 4 and 6 -3 %, 8 +0.5 %,
// 10 +1.0 %, 12 +1.2 %, 24 -0.5 %, 32 -1.3 %).
// ---------------------------------------------------------------------------
constexpr int kPlocRadius = 12;

__global__ void k_ploc_init(uint32_t n, uint32_t prim_begin, const float4* leaf_lo, const float4* leaf_hi, int* cid,
                            float4* clo, float4* chi) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  cid[j] = ~(int)(prim_begin + j);
  clo[j] = leaf_lo[j];
  chi[j] = leaf_hi[j];
}

// Half the surface area of the union of two boxes; symmetric in its
// arguments bit for bit (min / max commute), so both clusters of a pair
// see the same value.
__device__ __forceinline__ float merged_area(float4 alo, float4 ahi, float4 blo, float4 bhi) {
  const float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
  const float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
  const float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
  return dx * dy + dy * dz + dz * dx;
}

// Nearest neighbour of each cluster. Candidates are scanned in increasing
// index, for which the pair keys (lower index, higher index) increase, so
// keeping the first of equal areas realises the total order on pairs.
__global__ void k_ploc_nn(uint32_t c, const float4* clo, const float4* chi, uint32_t* nn) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= (int)c) return;
  const float4 alo = clo[i], ahi = chi[i];
  const int lo = max(0, i - kPlocRadius), hi = min((int)c - 1, i + kPlocRadius);
  float best = FLT_MAX;
  int bj = i == lo ? hi : lo;
  for (int j = lo; j <= hi; ++j) {
    if (j == i) continue;
    const float a = merged_area(alo, ahi, clo[j], chi[j]);
    if (a < best) {
      best = a;
      bj = j;
    }
  }
  nn[i] = (uint32_t)bj;
}

// keep[i]: cluster i survives (it is not the higher-index half of a mutual
// pair); lead[i]: cluster i is the lower-index half of a mutual pair.
__global__ void k_ploc_flags(uint32_t c, const uint32_t* nn, uint32_t* keep, uint32_t* lead) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    keep[c] = 0;
    lead[c] = 0;
  }
  if (i >= c) return;
  const uint32_t j = nn[i];
  const bool mutual = nn[j] == i;
  keep[i] = (mutual && j < i) ? 0u : 1u;
  lead[i] = (mutual && i < j) ? 1u : 0u;
}

// keep_pos / lead_pos: exclusive scans of keep / lead over c + 1 entries
// (entry c holds the totals).
__global__ void k_ploc_merge(uint32_t c, uint32_t n, uint32_t prim_begin, uint32_t node_begin, const uint32_t* nn,
                             const uint32_t* keep_pos, const uint32_t* lead_pos, const int* cid, const float4* clo,
                             const float4* chi, int* cid_out, float4* clo_out, float4* chi_out, bvh_node* nodes,
                             int32_t* parent) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c) return;
  const uint32_t j = nn[i];
  const bool mutual = nn[j] == i;
  if (mutual && j < i) return;
  const uint32_t p = keep_pos[i];
  if (!mutual) {
    cid_out[p] = cid[i];
    clo_out[p] = clo[i];
    chi_out[p] = chi[i];
    return;
  }
  const uint32_t id = c - lead_pos[c] - 1 + lead_pos[i];
  const int a = cid[i], b = cid[j];
  nodes[node_begin + id].child = make_int4(a, b, 0, 0);
  // parent[] local ids: internal nodes 0..n-2, leaves n-1 .. 2n-2
  parent[a < 0 ? (n - 1) + ((uint32_t)~a - prim_begin) : (uint32_t)a - node_begin] = (int32_t)id;
  parent[b < 0 ? (n - 1) + ((uint32_t)~b - prim_begin) : (uint32_t)b - node_begin] = (int32_t)id;
  if (id == 0) parent[0] = -1;
  const float4 alo = clo[i], ahi = chi[i], blo = clo[j], bhi = chi[j];
  cid_out[p] = (int)(node_begin + id);
  clo_out[p] = make_float4(fminf(alo.x, blo.x), fminf(alo.y, blo.y), fminf(alo.z, blo.z), 0.0f);
  chi_out[p] = make_float4(fmaxf(ahi.x, bhi.x), fmaxf(ahi.y, bhi.y), fmaxf(ahi.z, bhi.z), 0.0f);
}

// One primitive: a root whose two children are that leaf.
__global__ void k_single_root(uint32_t prim_begin, uint32_t node_begin, bvh_node* nodes, const float4* leaf_lo,
                              const float4* leaf_hi, uint32_t* max_depth) {
  bvh_node* nd = nodes + node_begin;
  write_slot(nd, 0, leaf_lo[0], leaf_hi[0]);
  write_slot(nd, 1, leaf_lo[0], leaf_hi[0]);
  nd->child = make_int4(~(int)prim_begin, ~(int)prim_begin, 0, 0);
  *max_depth = 2;
}

// ---------------------------------------------------------------------------
// Four-wide traversal tree. Each four-wide node is a binary node whose
// children are expanded greedily (largest surface area first) until it has
// four children or only leaves, so one node fetch replaces up to three binary
// fetches. Built level by level: a count pass gives each node's number of
// internal children, a scan turns the counts into contiguous child slots
// (siblings adjacent, breadth-first, deterministic), and a write pass stores
// the quantised node and the next level's work items. Every family of
// siblings starts on an even slot, i.e. on a 128-byte line (a family of odd
// size is followed by one padding slot, and so is the root): a node and its
// next sibling share the line a fetch brings into the cache. Padding slots
// hold an inverted box in every slot and are never referenced.
// ---------------------------------------------------------------------------
struct c4_box {
  float lo[3], hi[3];
};

__device__ __forceinline__ void binary_child(const bvh_node& nd, int c, int& ref, c4_box& b) {
  if (c == 0) {
    ref = nd.child.x;
    b.lo[0] = nd.xy0.x; b.hi[0] = nd.xy0.y; b.lo[1] = nd.xy0.z; b.hi[1] = nd.xy0.w;
    b.lo[2] = nd.z01.x; b.hi[2] = nd.z01.y;
  } else {
    ref = nd.child.y;
    b.lo[0] = nd.xy1.x; b.hi[0] = nd.xy1.y; b.lo[1] = nd.xy1.z; b.hi[1] = nd.xy1.w;
    b.lo[2] = nd.z01.z; b.hi[2] = nd.z01.w;
  }
}

__device__ __forceinline__ float half_area(const c4_box& b) {
  const float ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
  return ex * ey + ey * ez + ez * ex;
}

// Children of the four-wide node made from binary node `bin`, in the binary
// tree's left-to-right order.
__device__ int c4_members(const bvh_node* nodes, int bin, int ref[4], c4_box box[4]) {
  const bvh_node nd = nodes[bin];
  binary_child(nd, 0, ref[0], box[0]);
  binary_child(nd, 1, ref[1], box[1]);
  int m = (ref[0] == ref[1]) ? 1 : 2;  // single-primitive root
  while (m < 4) {
    int best = -1;
    float ba = -1.0f;
    for (int k = 0; k < m; ++k) {
      if (ref[k] < 0) continue;
      const float a = half_area(box[k]);
      if (a > ba) { ba = a; best = k; }
    }
    if (best < 0) break;
    const bvh_node sub = nodes[ref[best]];
    for (int k = m - 1; k > best; --k) { ref[k + 1] = ref[k]; box[k + 1] = box[k]; }
    binary_child(sub, 0, ref[best], box[best]);
    binary_child(sub, 1, ref[best + 1], box[best + 1]);
    ++m;
  }
  return m;
}

// Smallest biased exponent e with fmaf(255, 2^(e-127), lo) >= hi.
__device__ uint32_t c4_exponent(float lo, float hi) {
  int e = 1;
  const float ext = hi - lo;
  if (ext > 0.0f) {
    int x;
    frexpf(ext * (1.0f / 255.0f), &x);
    e = min(254, max(1, x + 127 - 1));
  }
  while (e < 254 && fmaf(255.0f, __uint_as_float((uint32_t)e << 23), lo) < hi) ++e;
  return (uint32_t)e;
}

__device__ uint32_t c4_qlo(float v, float o, float s) {
  int q = (int)fminf(fmaxf(floorf((v - o) / s), 0.0f), 255.0f);
  while (q > 0 && fmaf((float)q, s, o) > v) --q;
  return (uint32_t)q;
}

__device__ uint32_t c4_qhi(float v, float o, float s) {
  int q = (int)fminf(fmaxf(ceilf((v - o) / s), 0.0f), 255.0f);
  while (q < 255 && fmaf((float)q, s, o) < v) ++q;
  return (uint32_t)q;
}

// Internal children of each work item, rounded up to even (a padding slot
// after a family of odd size); 0 for a padding item (< 0).
__global__ void k_c4_count(const int* items, uint32_t width, const bvh_node* nodes, uint32_t* counts) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) counts[width] = 0;
  if (i >= width) return;
  if (items[i] < 0) {
    counts[i] = 0;
    return;
  }
  int ref[4];
  c4_box box[4];
  const int m = c4_members(nodes, items[i], ref, box);
  uint32_t k = 0;
  for (int c = 0; c < m; ++c) k += ref[c] >= 0 ? 1u : 0u;
  counts[i] = k + (k & 1u);
}

// items[i] (binary node) becomes four-wide slot node4_begin + base + i; its
// internal children get slots node4_begin + base + width + offsets[i] + j.
// A padding item (< 0) becomes a node whose four slots all hold an inverted
// box. pad[base + i] marks the padding slots for the stack bound.
__global__ void k_c4_write(const int* items, uint32_t width, uint32_t base, const bvh_node* nodes,
                           const uint32_t* offsets, uint32_t node4_begin, bvh4_node* nodes4, int* next,
                           uint32_t* pad_flag) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= width) return;
  pad_flag[base + i] = items[i] < 0 ? 1u : 0u;
  if (items[i] < 0) {
    bvh4_node pad;
    pad.ox = pad.oy = pad.oz = 0.0f;
    pad.meta = 127u | (127u << 8) | (127u << 16);  // unit scales, one child slot, no internal child
    pad.lo_x = pad.lo_y = pad.lo_z = 0xffffffffu;
    pad.hi_x = pad.hi_y = pad.hi_z = 0u;
    pad.int_base = 0u;
    pad.leaf_info = 0u;
    pad.leaf[0] = pad.leaf[1] = pad.leaf[2] = pad.leaf[3] = 0u;
    nodes4[node4_begin + base + i] = pad;
    return;
  }
  int ref0[4];
  c4_box box0[4];
  const int m = c4_members(nodes, items[i], ref0, box0);
  // internal children first, then the leaves, each in left-to-right order
  int ref[4];
  c4_box box[4];
  int nint = 0;
  for (int c = 0; c < m; ++c)
    if (ref0[c] >= 0) { ref[nint] = ref0[c]; box[nint] = box0[c]; ++nint; }
  for (int c = 0, k = nint; c < m; ++c)
    if (ref0[c] < 0) { ref[k] = ref0[c]; box[k] = box0[c]; ++k; }
  float lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = box[0].lo[k];
    hi[k] = box[0].hi[k];
    for (int c = 1; c < m; ++c) {
      lo[k] = fminf(lo[k], box[c].lo[k]);
      hi[k] = fmaxf(hi[k], box[c].hi[k]);
    }
  }
  uint32_t e[3];
  float sc[3];
  for (int k = 0; k < 3; ++k) {
    e[k] = c4_exponent(lo[k], hi[k]);
    sc[k] = __uint_as_float(e[k] << 23);
  }
  // Slots beyond the m children hold an inverted box (low planes 255, high
  // planes 0), which no slab test enters.
  uint32_t ql[3] = {0, 0, 0}, qh[3] = {0, 0, 0};
  for (int c = m; c < 4; ++c)
    for (int k = 0; k < 3; ++k) ql[k] |= 0xffu << (8 * c);
  const uint32_t first = base + width + offsets[i];
  for (int c = 0; c < m; ++c) {
    for (int k = 0; k < 3; ++k) {
      ql[k] |= c4_qlo(box[c].lo[k], lo[k], sc[k]) << (8 * c);
      qh[k] |= c4_qhi(box[c].hi[k], lo[k], sc[k]) << (8 * c);
    }
    if (c < nint) next[offsets[i] + c] = ref[c];
  }
  if (nint & 1) next[offsets[i] + nint] = -1;  // the family's padding slot
  bvh4_node out;
  out.ox = lo[0];
  out.oy = lo[1];
  out.oz = lo[2];
  out.meta = e[0] | (e[1] << 8) | (e[2] << 16) | ((uint32_t)(m - 1) << 24) | ((uint32_t)nint << 26);
  out.lo_x = ql[0];
  out.lo_y = ql[1];
  out.lo_z = ql[2];
  out.hi_x = qh[0];
  out.hi_y = qh[1];
  out.hi_z = qh[2];
  out.int_base = node4_begin + first;
  out.leaf_info = 0u;  // set by k_c4_stack (discs) or k_c4_place (triangles)
  for (int c = 0; c < 4; ++c) out.leaf[c] = (c >= nint && c < m) ? (uint32_t)(~ref[c]) : 0u;
  nodes4[node4_begin + base + i] = out;
}

// Stack bound and primitive counts, deepest level first: need(v) = k - 1 +
// max need(internal child) for a node with k >= 1 internal children, 0
// otherwise (a traversal that pushes all hit internal children but the one
// it descends into never holds more than need(root) entries); tot(v) = the
// primitives below v. Padding slots count nothing. Also marks the internal
// children that have leaf children (leaf_info; a triangle mesh replaces it
// with the node's first leaf in k_c4_place).
__global__ void k_c4_stack(bvh4_node* nodes4, const uint32_t* pad_flag, uint32_t node4_begin, uint32_t base,
                           uint32_t width, uint32_t* need, uint32_t* tot) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= width) return;
  if (pad_flag[base + i]) {
    need[base + i] = 0u;
    tot[base + i] = 0u;
    return;
  }
  bvh4_node* nd = nodes4 + node4_begin + base + i;
  const int nint = bvh4_nint(*nd);
  uint32_t deepest = 0, marks = 0, t = (uint32_t)(bvh4_nchild(*nd) - nint);
  for (int c = 0; c < nint; ++c) {
    const uint32_t v = nd->int_base + (uint32_t)c - node4_begin;
    const bvh4_node& ch = nodes4[nd->int_base + (uint32_t)c];
    deepest = max(deepest, need[v]);
    t += tot[v];
    if (bvh4_nchild(ch) > bvh4_nint(ch)) marks |= 1u << c;
  }
  nd->leaf_info = marks;
  need[base + i] = nint ? ((uint32_t)nint - 1u + deepest) : 0u;
  tot[base + i] = t;
}

// Leaf placement of a triangle mesh, root level first: node v's
// leaves take the stored positions [off(v), off(v) + leaves of v), its
// internal children's subtrees follow in slot order (a depth-first order of
// the tree, so each node's leaves are contiguous). newpos[old local
// position] = new local position; the node's leaf_info and leaf[] get the
// new positions. The binary tree still holds the old positions here.
__global__ void k_c4_place(bvh4_node* nodes4, const uint32_t* pad_flag, uint32_t node4_begin, uint32_t prim_begin,
                           uint32_t base, uint32_t width, const uint32_t* tot, uint32_t* off, uint32_t* newpos) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= width) return;
  if (pad_flag[base + i]) return;
  bvh4_node* nd = nodes4 + node4_begin + base + i;
  const int m = bvh4_nchild(*nd), nint = bvh4_nint(*nd);
  const uint32_t o = off[base + i];
  nd->leaf_info = prim_begin + o;
  nd->meta |= rti_dev::kNodeLeavesContiguous;
  uint32_t run = o + (uint32_t)(m - nint);
  for (int c = 0; c < nint; ++c) {
    const uint32_t v = nd->int_base + (uint32_t)c - node4_begin;
    off[v] = run;
    run += tot[v];
  }
  for (int c = nint; c < m; ++c) {
    newpos[nd->leaf[c] - prim_begin] = o + (uint32_t)(c - nint);
    nd->leaf[c] = prim_begin + o + (uint32_t)(c - nint);
  }
}

// Records, input ids and areas of one instance to their new stored
// positions: gather into scratch (old order), then scatter.
__global__ void k_perm_gather(const float4* prims, uint32_t shift, const uint32_t* perm, const float* areas,
                              uint32_t prim_begin, uint32_t n, float4* t0, float4* t1, float4* t2, float4* t3,
                              uint32_t* tperm, float* tarea) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float4* rec = prims + ((size_t)(prim_begin + p) << shift);
  t0[p] = rec[0];
  t1[p] = rec[1];
  if (shift == 2) {
    t2[p] = rec[2];
    t3[p] = rec[3];
  }
  tperm[p] = perm[prim_begin + p];
  tarea[p] = areas[prim_begin + p];
}

__global__ void k_perm_scatter(float4* prims, uint32_t shift, uint32_t* perm, float* areas, uint32_t prim_begin,
                               uint32_t n, const uint32_t* newpos, const float4* t0, const float4* t1,
                               const float4* t2, const float4* t3, const uint32_t* tperm, const float* tarea) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (newpos[p] >= n) return;  // cannot happen: newpos is a permutation of [0, n)
  const uint32_t q = prim_begin + newpos[p];
  float4* rec = prims + ((size_t)q << shift);
  rec[0] = t0[p];
  rec[1] = t1[p];
  if (shift == 2) {
    rec[2] = t2[p];
    rec[3] = t3[p];
  }
  perm[q] = tperm[p];
  areas[q] = tarea[p];
}

// Leaf references of the binary tree to the new stored positions.
__global__ void k_remap_binary(bvh_node* nodes, uint32_t node_begin, uint32_t count, uint32_t prim_begin,
                               const uint32_t* newpos) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  int4 ch = nodes[node_begin + i].child;
  if (ch.x < 0) ch.x = ~(int)(prim_begin + newpos[(uint32_t)(~ch.x) - prim_begin]);
  if (ch.y < 0) ch.y = ~(int)(prim_begin + newpos[(uint32_t)(~ch.y) - prim_begin]);
  nodes[node_begin + i].child = ch;
}

__global__ void k_set_first_u32(uint32_t* p, uint32_t v) { p[0] = v; }

// The root level: the binary root and one padding slot.
__global__ void k_c4_root(int* items, uint32_t node_begin) {
  items[0] = (int)node_begin;
  items[1] = -1;
}

// ---------------------------------------------------------------------------
// Disc-overlap graph. The pair predicate restates disc_neighborhood::check_dist
// (src/rti/geo/disc_neighborhood.hpp:190-220) in FP32 without contraction:
// per-axis rejection at >= r_i + r_j, then sqrt(dx^2 + dy^2 + dz^2) < r_i + r_j.
// Candidates come from the instance's tree with the query box
// c_i +- (r_i + r_max), which contains every disc that can overlap disc i.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool overlap_pred(float4 p1, float4 p2) {
#pragma clang fp contract(off)
  float sr = p1.w + p2.w;
  if (fabsf(p2.x - p1.x) >= sr) return false;
  if (fabsf(p1.y - p2.y) >= sr) return false;
  if (fabsf(p1.z - p2.z) >= sr) return false;
  float d1 = p1.x - p2.x, d2 = p1.y - p2.y, d3 = p1.z - p2.z;
  float s = __fadd_rn(__fadd_rn(__fmul_rn(d1, d1), __fmul_rn(d2, d2)), __fmul_rn(d3, d3));
  // The reference compares the correctly rounded FP32 sqrt(s) with sr. On the
  // grid inputs many pairs sit exactly at distance sr, so the decision must
  // be bit-exact without relying on the device sqrt's rounding:
  //   sqrt_rn(s) < sr  <=>  sqrt(s) < mid  <=>  s < mid^2,
  // with mid the midpoint of sr and its float predecessor. mid has at most 25
  // significant bits, so mid^2 is exact in FP64; s == mid^2 is the tie, which
  // rounds to the neighbour with an even significand.
  const double hi = (double)sr;
  const double lo = (double)__int_as_float(__float_as_int(sr) - 1);
  const double mid = 0.5 * (hi + lo);
  const double m2 = mid * mid;
  const double sd = (double)s;
  if (sd < m2) return true;
  if (sd > m2) return false;
  return (__float_as_int(sr) & 1) != 0;
}

__device__ __forceinline__ bool box_overlap(float lox, float hix, float loy, float hiy, float loz, float hiz,
                                            float3 qlo, float3 qhi) {
  return !(hix < qlo.x || lox > qhi.x || hiy < qlo.y || loy > qhi.y || hiz < qlo.z || loz > qhi.z);
}

template <bool kWrite>
__global__ void k_neighbors(const float4* prims, uint32_t shift, uint32_t prim_begin, uint32_t n, uint32_t root, const bvh_node* nodes,
                            float rmax, uint32_t* counts, const uint32_t* row_ptr, uint32_t* col) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t self = prim_begin + j;
  float4 p = prims[(size_t)self << shift];
  float reach = p.w + rmax;
  float3 qlo = make_float3(p.x - reach, p.y - reach, p.z - reach);
  float3 qhi = make_float3(p.x + reach, p.y + reach, p.z + reach);
  uint32_t cnt = 0;
  uint32_t w = kWrite ? row_ptr[self] : 0u;
  int stack[64];
  int sp = 0;
  int cur = (int)root;
  for (;;) {
    const bvh_node& nd = nodes[cur];
    int4 ch = nd.child;
    float4 a = nd.xy0, b = nd.xy1, z = nd.z01;
    bool h0 = box_overlap(a.x, a.y, a.z, a.w, z.x, z.y, qlo, qhi);
    bool h1 = box_overlap(b.x, b.y, b.z, b.w, z.z, z.w, qlo, qhi);
    if (ch.x == ch.y) h1 = false;  // single-primitive root
    int next = -1;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      bool h = c == 0 ? h0 : h1;
      int cc = c == 0 ? ch.x : ch.y;
      if (!h) continue;
      if (cc < 0) {
        uint32_t q = (uint32_t)(~cc);
        if (q != self && overlap_pred(p, prims[(size_t)q << shift])) {
          if (kWrite) col[w] = q;
          ++w;
          ++cnt;
        }
      } else if (next < 0) {
        next = cc;
      } else {
        stack[sp++] = cc;
      }
    }
    if (next >= 0) {
      cur = next;
    } else if (sp > 0) {
      cur = stack[--sp];
    } else {
      break;
    }
  }
  if (!kWrite) counts[j] = cnt;
}

// ---------------------------------------------------------------------------
// Exposed disc area inside the x-y box. Restates
// disc_bounding_box_intersector::area_inside
// (src/rti/geo/disc_bounding_box_intersector.hpp:42-76; closest approach
// :372-441; face segments :116-125; adjacent-face corner overlap :127-261).
// Faces in the reference's order: right, bottom, left, top (:349-356).
// Triangles: full triangle area.
// ---------------------------------------------------------------------------
__device__ float face_approach(int f, float xmin, float ymin, float xmax, float ymax, float cx, float cy, float r,
                               float nx, float ny, float nz) {
  float pos, plane, nperp;
  if (f == 0) { pos = cx; plane = xmax; nperp = ny; }
  else if (f == 1) { pos = -cy; plane = -ymin; nperp = nx; }
  else if (f == 2) { pos = -cx; plane = -xmin; nperp = ny; }
  else { pos = cy; plane = ymax; nperp = nx; }
  float xterm = r * sqrtf(nz * nz + nperp * nperp);
  if (pos + xterm <= plane) return FLT_MAX;
  if (pos - xterm >= plane) return -FLT_MAX;
  if (xterm <= 1e-9f) return FLT_MAX;
  return (plane - pos) * r / xterm;
}

__device__ float disc_area_inside(float xmin, float ymin, float xmax, float ymax, float4 disc, float4 nin) {
#pragma clang fp contract(off)
  float r = disc.w;
  float nl = sqrtf(nin.x * nin.x + nin.y * nin.y + nin.z * nin.z);
  f3 n = make(nin.x / nl, nin.y / nl, nin.z / nl);
  if (r <= 0.0f) return 0.0f;
  float full = r * r * 3.14159265358979323846f;
  float cx = disc.x, cy = disc.y;
  if ((xmin <= cx - r && cx + r <= xmax) && (ymin <= cy - r && cy + r <= ymax)) return full;
  if ((cx + r <= xmin || xmax <= cx - r) || (cy + r <= ymin || ymax <= cy - r)) return 0.0f;
  float a[4];
  for (int f = 0; f < 4; ++f) {
    a[f] = face_approach(f, xmin, ymin, xmax, ymax, cx, cy, r, n.x, n.y, n.z);
    if (a[f] < -r) return 0.0f;
  }
  const float fnx[4] = {-1.0f, 0.0f, 1.0f, 0.0f}, fny[4] = {0.0f, 1.0f, 0.0f, -1.0f};
  const float fcx[4] = {xmax, xmax, xmin, xmin}, fcy[4] = {ymax, ymin, ymin, ymax};
  float area = 0.0f;
  for (int f = 0; f < 4; ++f) {
    if (-r < a[f] && a[f] < r) {
      double ang = 2.0 * acos((double)a[f] / r);
      double seg = (double)(r * r / 2.0f) * (ang - sin(ang));
      area = (float)((double)area + seg);
    }
  }
  f3 cp = make(disc.x, disc.y, disc.z);
  for (int f = 0; f < 4; ++f) {
    int g = (f + 1) & 3;
    float d1 = a[f], d2 = a[g];
    if (!(-r < d1 && d1 < r && -r < d2 && d2 < r)) continue;
    f3 n1 = make(fnx[f], fny[f], 0.0f), n2 = make(fnx[g], fny[g], 0.0f);
    f3 i1 = normalize(cross(n, n1));
    f3 i2 = normalize(cross(n, n2));
    if (dot(i1, n2) >= 0.0f) i1 = i1 * -1.0f;
    if (dot(i2, n1) >= 0.0f) i2 = i2 * -1.0f;
    float xi = fcx[g], yi = fcy[g];
    float zi = (n.x * cp.x + n.y * cp.y + n.z * cp.z - n.x * xi - n.y * yi) / n.z;
    f3 ip = make(xi, yi, zi);
    f3 dc = cp - ip;
    if (sqrtf(dot(dc, dc)) >= r) continue;
    f3 q1 = ip + i1 * dot(dc, i1) + i1 * sqrtf(r * r - d1 * d1);
    f3 q2 = ip + i2 * dot(dc, i2) + i2 * sqrtf(r * r - d2 * d2);
    f3 e1 = q1 - cp, e2 = q2 - cp;
    float ang = acosf(dot(e1, e2) / sqrtf(dot(e1, e1)) / sqrtf(dot(e2, e2)));
    float seg = r * r / 2.0f * (ang - sinf(ang));
    f3 cr = cross(q1 - ip, q2 - ip);
    float tri = 0.5f * sqrtf(dot(cr, cr));
    area -= seg + tri;
  }
  return full - area;
}

__global__ void k_areas(int kind, const float4* prims, uint32_t shift, uint32_t prim_begin,
                        uint32_t n, float xmin, float ymin, float xmax, float ymax, float* areas) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  size_t g = (size_t)prim_begin + j;
  float A;
  if (kind == 0) {
    A = disc_area_inside(xmin, ymin, xmax, ymax, prims[g << shift], prims[(g << shift) + 1]);
  } else {
    float4 e1 = prims[(g << shift) + 1], e2 = prims[(g << shift) + 2];
    f3 c = cross(make(e1.x, e1.y, e1.z), make(e2.x, e2.y, e2.z));
    A = 0.5f * sqrtf(dot(c, c));
  }
  areas[g] = A;
}

inline unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

size_t scan_tmp_size(uint32_t n) { return (size_t)(n + kScanBlock - 1) / kScanBlock + 1; }

hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
  k_scan_totals<<<nb, kScanBlock, 0, stream>>>(in, n, tmp);
  k_scan_single<<<1, kScanBlock, 0, stream>>>(tmp, nb);
  k_scan_apply<<<nb, kScanBlock, 0, stream>>>(in, out, n, tmp);
  return hipGetLastError();
}

hipError_t scratch_reserve(build_scratch& s, uint32_t n) {
  if (n <= s.cap && s.keys[0]) return hipSuccess;
  scratch_free(s);
  uint32_t cap = n < 1024 ? 1024 : n;
  size_t nblocks = (cap + kRsTile - 1) / kRsTile;
  hipError_t e;
#define RTI_ALLOC(ptr, bytes)                              \
  e = hipMalloc((void**)&(ptr), (bytes));                 \
  if (e != hipSuccess) { scratch_free(s); return e; }
  // + 2: the four-wide collapse's work lists hold up to n items and counts
  // one more (its levels include padding slots)
  RTI_ALLOC(s.keys[0], (cap + 2ull) * 4ull);
  RTI_ALLOC(s.keys[1], (cap + 2ull) * 4ull);
  RTI_ALLOC(s.vals[0], (cap + 2ull) * 4ull);
  RTI_ALLOC(s.vals[1], (cap + 2ull) * 4ull);
  RTI_ALLOC(s.hist, 256ull * nblocks * 4ull);
  size_t st = scan_tmp_size((uint32_t)(256 * nblocks));
  size_t st2 = scan_tmp_size(cap + 1);
  RTI_ALLOC(s.scan_tmp, (st > st2 ? st : st2) * 4ull);
  RTI_ALLOC(s.parent, (2ull * cap) * 4ull);
  RTI_ALLOC(s.flags, (2ull * cap + 2ull) * 4ull);  // binary refit flags; four-wide stack needs
  RTI_ALLOC(s.leaf_lo, cap * 16ull);
  RTI_ALLOC(s.leaf_hi, cap * 16ull);
  RTI_ALLOC(s.counts, (cap + 1ull) * 4ull);
  RTI_ALLOC(s.depth, 4ull);
  RTI_ALLOC(s.clo[0], cap * 16ull);
  RTI_ALLOC(s.clo[1], cap * 16ull);
  RTI_ALLOC(s.chi[0], cap * 16ull);
  RTI_ALLOC(s.chi[1], cap * 16ull);
  RTI_ALLOC(s.lead, (cap + 1ull) * 4ull);
#undef RTI_ALLOC
  s.cap = cap;
  return hipSuccess;
}

void scratch_free(build_scratch& s) {
  void* ptrs[] = {s.keys[0], s.keys[1], s.vals[0], s.vals[1], s.hist, s.scan_tmp, s.parent, s.flags, s.leaf_lo,
                  s.leaf_hi, s.counts, s.depth, s.clo[0], s.clo[1], s.chi[0], s.chi[1], s.lead};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  s = build_scratch{};
}

// PLOC topology (child ids and parent links) over the instance's leaves,
// which k_scatter_prims has stored in Morton order with their boxes. The
// sort buffers are free here: keys[] hold the cluster ids, vals[0] the
// nearest neighbours, counts[] the survivor flags.
hipError_t build_ploc(const build_job& job, const store_ptrs& st, build_scratch& s, hipStream_t stream) {
  const uint32_t n = job.n;
  const unsigned T = 256;
  int* cid[2] = {reinterpret_cast<int*>(s.keys[0]), reinterpret_cast<int*>(s.keys[1])};
  k_ploc_init<<<blocks_for(n, T), T, 0, stream>>>(n, job.prim_begin, s.leaf_lo, s.leaf_hi, cid[0], s.clo[0],
                                                  s.chi[0]);
  uint32_t c = n;
  int cur = 0;
  hipError_t e;
  while (c > 1) {
    k_ploc_nn<<<blocks_for(c, T), T, 0, stream>>>(c, s.clo[cur], s.chi[cur], s.vals[0]);
    k_ploc_flags<<<blocks_for(c, T), T, 0, stream>>>(c, s.vals[0], s.counts, s.lead);
    e = exclusive_scan_u32(s.counts, s.counts, c + 1, s.scan_tmp, stream);
    if (e != hipSuccess) return e;
    e = exclusive_scan_u32(s.lead, s.lead, c + 1, s.scan_tmp, stream);
    if (e != hipSuccess) return e;
    k_ploc_merge<<<blocks_for(c, T), T, 0, stream>>>(c, n, job.prim_begin, job.node_begin, s.vals[0], s.counts,
                                                     s.lead, cid[cur], s.clo[cur], s.chi[cur], cid[cur ^ 1],
                                                     s.clo[cur ^ 1], s.chi[cur ^ 1], st.nodes, s.parent);
    uint32_t next = 0;
    e = hipMemcpyAsync(&next, s.counts + c, 4, hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    if (next >= c || next == 0) return hipErrorUnknown;  // no merge: cannot happen (see above)
    c = next;
    cur ^= 1;
  }
  return hipGetLastError();
}

hipError_t build_instance(const build_job& job, const store_ptrs& st, build_scratch& s, uint32_t* depth_out,
                          hipStream_t stream) {
  const uint32_t n = job.n;
  if (n == 0) return hipErrorInvalidValue;
  hipError_t e = scratch_reserve(s, n);
  if (e != hipSuccess) return e;
  setup_args a;
  a.kind = job.prim_kind;
  a.n = n;
  a.xyzr = job.d_xyzr;
  a.nrm = job.d_nrm;
  a.verts = job.d_verts;
  a.tris = job.d_tris;
  // One scale for all three axes (the largest centroid extent): Morton
  // splits then halve cubes, not slabs stretched along the long axis, which
  // measured 9-14 % fewer node visits per segment on trench and hole scenes.
  const float ext = std::max(job.cmax[0] - job.cmin[0], std::max(job.cmax[1] - job.cmin[1], job.cmax[2] - job.cmin[2]));
  float inv[3];
  for (int k = 0; k < 3; ++k) inv[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
  a.cmin0 = job.cmin[0]; a.cmin1 = job.cmin[1]; a.cmin2 = job.cmin[2];
  a.cinv0 = inv[0]; a.cinv1 = inv[1]; a.cinv2 = inv[2];
  a.pad = job.pad;
  const unsigned T = 256;
  k_morton<<<blocks_for(n, T), T, 0, stream>>>(a, s.keys[0], s.vals[0]);
  // 30-bit keys: four 8-bit passes, result back in buffer 0
  uint32_t nblocks = (n + kRsTile - 1) / kRsTile;
  int cur = 0;
  for (int shift = 0; shift < 32; shift += 8) {
    k_rs_hist<<<nblocks, kRsThreads, 0, stream>>>(s.keys[cur], n, shift, s.hist, nblocks);
    e = exclusive_scan_u32(s.hist, s.hist, 256u * nblocks, s.scan_tmp, stream);
    if (e != hipSuccess) return e;
    k_rs_scatter<<<nblocks, kRsThreads, 0, stream>>>(s.keys[cur], s.vals[cur], s.keys[cur ^ 1], s.vals[cur ^ 1], n,
                                                     shift, s.hist, nblocks);
    cur ^= 1;
  }
  k_scatter_prims<<<blocks_for(n, T), T, 0, stream>>>(a, s.vals[cur], job.prim_begin, st.prims, st.prim_shift, st.perm, s.leaf_lo,
                                                       s.leaf_hi);
  e = hipMemsetAsync(s.depth, 0, 4, stream);
  if (e != hipSuccess) return e;
  if (n == 1) {
    k_single_root<<<1, 1, 0, stream>>>(job.prim_begin, job.node_begin, st.nodes, s.leaf_lo, s.leaf_hi, s.depth);
  } else {
    if (job.builder == kBuildPloc) {
      e = build_ploc(job, st, s, stream);
      if (e != hipSuccess) return e;
    } else {
      k_karras<<<blocks_for(n - 1, T), T, 0, stream>>>(s.keys[cur], (int)n, job.prim_begin, job.node_begin,
                                                       st.nodes, s.parent);
    }
    e = hipMemsetAsync(s.flags, 0, (n - 1) * 4ull, stream);
    if (e != hipSuccess) return e;
    k_refit<<<blocks_for(n, T), T, 0, stream>>>((int)n, job.prim_begin, job.node_begin, st.nodes, s.parent, s.flags,
                                                s.leaf_lo, s.leaf_hi, s.depth);
  }
  k_areas<<<blocks_for(n, T), T, 0, stream>>>(job.prim_kind, st.prims, st.prim_shift, job.prim_begin, n,
                                              job.area_box[0], job.area_box[1], job.area_box[2], job.area_box[3],
                                              st.areas);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(depth_out, s.depth, 4, hipMemcpyDeviceToHost, stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(stream);
}

hipError_t build_wide(const build_job& job, const store_ptrs& st, build_scratch& s, uint32_t* n_wide,
                      uint32_t* stack_bound, hipStream_t stream) {
  // work items ping-pong in keys[0] / keys[1], counts in vals[0], child
  // offsets in vals[1], stack needs in flags, padding marks in parent (as
  // words; all at least wide_capacity(n) entries, the binary build is done
  // with parent)
  const unsigned T = 256;
  int* items[2] = {reinterpret_cast<int*>(s.keys[0]), reinterpret_cast<int*>(s.keys[1])};
  uint32_t* pad_flag = reinterpret_cast<uint32_t*>(s.parent);
  k_c4_root<<<1, 1, 0, stream>>>(items[0], job.node_begin);
  std::vector<std::pair<uint32_t, uint32_t>> levels;  // (base, width)
  uint32_t base = 0, width = 2;
  int cur = 0;
  hipError_t e;
  while (width > 0) {
    levels.emplace_back(base, width);
    k_c4_count<<<blocks_for(width, T), T, 0, stream>>>(items[cur], width, st.nodes, s.vals[0]);
    e = exclusive_scan_u32(s.vals[0], s.vals[1], width + 1, s.scan_tmp, stream);
    if (e != hipSuccess) return e;
    k_c4_write<<<blocks_for(width, T), T, 0, stream>>>(items[cur], width, base, st.nodes, s.vals[1],
                                                       job.node4_begin, st.nodes4, items[cur ^ 1], pad_flag);
    uint32_t next = 0;
    e = hipMemcpyAsync(&next, s.vals[1] + width, 4, hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    base += width;
    width = next;
    cur ^= 1;
    if (base + width > wide_capacity(job.n)) return hipErrorInvalidValue;  // cannot happen
  }
  uint32_t* tot = reinterpret_cast<uint32_t*>(s.leaf_lo);
  for (size_t l = levels.size(); l-- > 0;)
    k_c4_stack<<<blocks_for(levels[l].second, T), T, 0, stream>>>(st.nodes4, pad_flag, job.node4_begin,
                                                                   levels[l].first, levels[l].second, s.flags, tot);
  if (job.prim_kind == 1) {
    // a triangle mesh: primitives in the tree's depth-first order
    // (leaf totals and offsets in leaf_lo / leaf_hi as words, new positions in
    // counts, the records' staging in clo / chi / keys; the binary build is
    // done with all of them)
    uint32_t* off = reinterpret_cast<uint32_t*>(s.leaf_hi);
    uint32_t* newpos = s.counts;
    // every primitive is the leaf of exactly one node: the root's subtree
    // holds all n (the placement then writes a permutation of [0, n))
    uint32_t total = 0;
    e = hipMemcpyAsync(&total, tot, 4, hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    if (total != job.n) return hipErrorInvalidValue;  // cannot happen
    k_set_first_u32<<<1, 1, 0, stream>>>(off, 0u);
    for (size_t l = 0; l < levels.size(); ++l)
      k_c4_place<<<blocks_for(levels[l].second, T), T, 0, stream>>>(st.nodes4, pad_flag, job.node4_begin,
                                                                     job.prim_begin, levels[l].first,
                                                                     levels[l].second, tot, off, newpos);
    const uint32_t n = job.n;
    k_perm_gather<<<blocks_for(n, T), T, 0, stream>>>(st.prims, st.prim_shift, st.perm, st.areas, job.prim_begin,
                                                      n, s.clo[0], s.clo[1], s.chi[0], s.chi[1], s.keys[0],
                                                      reinterpret_cast<float*>(s.keys[1]));
    k_perm_scatter<<<blocks_for(n, T), T, 0, stream>>>(st.prims, st.prim_shift, st.perm, st.areas, job.prim_begin,
                                                       n, newpos, s.clo[0], s.clo[1], s.chi[0], s.chi[1], s.keys[0],
                                                       reinterpret_cast<const float*>(s.keys[1]));
    const uint32_t nbin = n > 1 ? n - 1 : 1;
    k_remap_binary<<<blocks_for(nbin, T), T, 0, stream>>>(st.nodes, job.node_begin, nbin, job.prim_begin, newpos);
  }
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(stack_bound, s.flags, 4, hipMemcpyDeviceToHost, stream);
  if (e != hipSuccess) return e;
  *n_wide = base;
  return hipStreamSynchronize(stream);
}

hipError_t neighbor_count(const build_job& job, const store_ptrs& st, build_scratch& s, hipStream_t stream) {
  const unsigned T = 128;
  k_neighbors<false><<<blocks_for(job.n, T), T, 0, stream>>>(st.prims, st.prim_shift, job.prim_begin, job.n, job.node_begin,
                                                              st.nodes, job.rmax, s.counts, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t neighbor_fill(const build_job& job, const store_ptrs& st, hipStream_t stream) {
  const unsigned T = 128;
  k_neighbors<true><<<blocks_for(job.n, T), T, 0, stream>>>(st.prims, st.prim_shift, job.prim_begin, job.n, job.node_begin,
                                                             st.nodes, job.rmax, nullptr, st.nb_ptr, st.nb_col);
  return hipGetLastError();
}

}  // namespace rti_impl
