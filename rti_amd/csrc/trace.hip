// Monte Carlo flux trace for gfx950.
//
// Replaces the per-ray loop of trace::tracer::run (reference
// src/rti/trace/tracer.hpp:124-283): source sample, closest hit, boundary
// (periodic / reflective), back-face pass-through, stick + deposit on the hit
// disc and on the overlapping neighbours the ray crosses, Russian roulette,
// diffuse / specular reflection.
//
// Two launch shapes share every device function:
//   * persistent: a grid sized to the resident capacity; each wave keeps 64
//     live rays and, after its lanes' segments, refills idle lanes from a
//     global ray queue, reserved in blocks by one atomic per wave (ballot +
//     popcount give each idle lane its id), so lanes never wait for the
//     longest path of a batch;
//   * plain: one thread per ray, traced to termination (the baseline).
// Traversal (closest hit), on one of two trees built from the same binary tree:
//   four-wide (default): one 64-byte node holds four children's boxes,
//     quantised to 8 bits per plane; hit children are visited nearest first
//     and the others pushed with their entry distance, so a popped entry that
//     lies beyond the closest hit found meanwhile is dropped without a fetch;
//   binary: both child boxes in the parent (one 64-byte node per visit).
// The per-lane stack keeps its top LS entries in LDS and the rest in private
// memory. In the persistent kernel on the four-wide tree, hit leaves can be
// put on a wave-shared list in LDS and tested in leaf phases, one entry per
// lane whichever lane's ray it belongs to (leaf_batch).
// Accumulation: per-primitive 64-bit integer atomics of round(v * 2^shift),
// so totals are independent of the order of deposits (bit-exact run to run).

#include <hip/hip_runtime.h>

#include <cfloat>

#include "internal.hpp"

using namespace rti_dev;

namespace rti_impl {
namespace {

constexpr uint32_t kMaxSegmentsPerRay = 1u << 20;
constexpr int kMaxBlock = 256;
// Stack entries of the four-wide traversal (LDS + private). The build
// reports the exact bound for each tree; a launch checks it.
constexpr int kWideStack = 64;

struct tparams {
  const inst_desc* desc;
  uint32_t n_inst;
  uint64_t ray_lo, ray_hi;  // global ray ids traced by this launch
  const float4* __restrict__ prims;  // 1 << prim_shift float4 per primitive
  uint32_t prim_shift;
  const bvh_node* __restrict__ nodes;
  const bvh4_node* __restrict__ nodes4;
  const uint32_t* __restrict__ nb_col;
  unsigned long long* acc;  // kAccStride per accumulator slot: sum, sum of squares, count
  unsigned long long* adj;  // 2 per slot: d/ds, d/dn (adjoint only)
  double* mom;              // 2 per slot: sums of deposit^3, ^4 (or null)
  unsigned long long* queue;
  unsigned long long* stats;
  int compact;  // lanes without an open query before a wave shades and refills
  int leaf_min;  // postponed leaves: listed leaves, in 64ths of the lanes with a node left, that start a leaf phase
  int refill_min;  // lanes waiting for a new ray before a wave refills (1..64)
  uint32_t block;  // threads per block (the LDS slot stride), uniform
};

struct counters {
  uint32_t seg = 0, hit = 0, miss = 0, trunc = 0;
  uint32_t nodes = 0, prims = 0, nbt = 0, deps = 0;
  uint32_t longest = 0;  // segments of the longest path this lane ended
};

// Resumable closest-hit traversal (one node per trav_step), so that a wave
// can stop traversing once enough of its lanes have finished their query.
// Closest primitive hit in [tnear, tbest]; ties go to the lower primitive id.
struct trav_state {
  f3 inv, oi;       // 1 / direction, origin / direction
  float tnear;
  float tbest;      // starts at the boundary-face distance
  float tb;         // boundary-face distance
  int face;         // boundary face hit, -1 none
  int cur;          // node being visited (postponed leaves: -1 = no node left)
  float curd;       // entry distance of cur (postponed leaves: skipped beyond tbest)
  int sp;           // stack depth
  uint32_t lastq;   // postponed leaves: list sequence number after this lane's last entry
  uint32_t pbest;   // closest primitive so far
  bool found;
};

// Hit distance th >= tnear of the ray o + t d on primitive p (disc or
// triangle); false when the ray misses it. No upper bound on th: the callers
// compare it with their closest hit.
template <bool COUNT>
__device__ __forceinline__ bool prim_hit(const tparams& P, const f3 o, const f3 d, float tnear, uint32_t p,
                                         float& th, counters& C) {
  if (COUNT) C.prims++;
  const float4* rec = P.prims + ((size_t)p << P.prim_shift);
  const float4 A = rec[0];
  const float4 B = rec[1];
  bool ok;
  if (A.w > 0.0f) {
    // oriented disc (Embree ORIENTED_DISC_POINT semantics: closest t in
    // [tnear, tfar], no back-face culling, |hit - centre| < radius)
    const f3 n = make(B.x, B.y, B.z);
    const float dn = dot(d, n);
    const f3 co = make(A.x - o.x, A.y - o.y, A.z - o.z);
    th = dot(co, n) * __builtin_amdgcn_rcpf(dn);
    const f3 hv = make(fmaf(d.x, th, -co.x), fmaf(d.y, th, -co.y), fmaf(d.z, th, -co.z));
    ok = (dn != 0.0f) && th >= tnear && dot(hv, hv) < A.w * A.w;
  } else {
    // triangle: Moller-Trumbore
    const float4 C2 = rec[2];
    const f3 e1 = make(B.x, B.y, B.z), e2 = make(C2.x, C2.y, C2.z);
    const f3 pv = cross(d, e2);
    const float det = dot(e1, pv);
    const float idet = __builtin_amdgcn_rcpf(det);
    const f3 tv = make(o.x - A.x, o.y - A.y, o.z - A.z);
    const float u = dot(tv, pv) * idet;
    const f3 qv = cross(tv, e1);
    const float v = dot(d, qv) * idet;
    th = dot(e2, qv) * idet;
    ok = (det != 0.0f) && u >= 0.0f && u <= 1.0f && v >= 0.0f && u + v <= 1.0f && th >= tnear;
  }
  return ok;
}

// Closest-hit update for one primitive in [tnear, tbest]; equal distances go
// to the lower primitive id.
template <bool COUNT>
__device__ __forceinline__ void test_prim(const tparams& P, const f3 o, const f3 d, uint32_t p, trav_state& T,
                                          counters& C) {
  float th;
  if (prim_hit<COUNT>(P, o, d, T.tnear, p, th, C) && th <= T.tbest && (th < T.tbest || !T.found || p < T.pbest)) {
    T.tbest = th;
    T.pbest = p;
    T.found = true;
  }
}

template <int LS, bool COUNT>
__device__ __forceinline__ bool trav_step(const tparams& P, const f3 o, const f3 d, trav_state& T, uint32_t* lds,
                                          uint32_t* ovf, counters& C) {
  const uint32_t stride = P.block;
  const bvh_node* nd = P.nodes + T.cur;
  const float4 a = nd->xy0, b = nd->xy1, z = nd->z01;
  const int4 ch = nd->child;
  if (COUNT) C.nodes++;
  const f3 inv = T.inv, oi = T.oi;
  float ax0 = fmaf(a.x, inv.x, -oi.x), ax1 = fmaf(a.y, inv.x, -oi.x);
  float ay0 = fmaf(a.z, inv.y, -oi.y), ay1 = fmaf(a.w, inv.y, -oi.y);
  float az0 = fmaf(z.x, inv.z, -oi.z), az1 = fmaf(z.y, inv.z, -oi.z);
  float bx0 = fmaf(b.x, inv.x, -oi.x), bx1 = fmaf(b.y, inv.x, -oi.x);
  float by0 = fmaf(b.z, inv.y, -oi.y), by1 = fmaf(b.w, inv.y, -oi.y);
  float bz0 = fmaf(z.z, inv.z, -oi.z), bz1 = fmaf(z.w, inv.z, -oi.z);
  float tmin0 = fmaxf(fmaxf(fminf(ax0, ax1), fminf(ay0, ay1)), fmaxf(fminf(az0, az1), T.tnear));
  float tmax0 = fminf(fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1)), fminf(fmaxf(az0, az1), T.tbest));
  float tmin1 = fmaxf(fmaxf(fminf(bx0, bx1), fminf(by0, by1)), fmaxf(fminf(bz0, bz1), T.tnear));
  float tmax1 = fminf(fminf(fmaxf(bx0, bx1), fmaxf(by0, by1)), fminf(fmaxf(bz0, bz1), T.tbest));
  bool h0 = tmin0 <= tmax0;
  bool h1 = (tmin1 <= tmax1) && (ch.y != ch.x);
  // leaf children are tested right away
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    bool h = c == 0 ? h0 : h1;
    int cc = c == 0 ? ch.x : ch.y;
    if (!(h && cc < 0)) continue;
    if (c == 0) h0 = false; else h1 = false;
    test_prim<COUNT>(P, o, d, (uint32_t)(~cc), T, C);
  }
  if (h0 && h1) {
    const bool near0 = tmin0 <= tmin1;
    const int far = near0 ? ch.y : ch.x;
    T.cur = near0 ? ch.x : ch.y;
    if (T.sp < LS) lds[__umul24((uint32_t)T.sp, stride)] = (uint32_t)far;
    else ovf[T.sp - LS] = (uint32_t)far;
    ++T.sp;
    return false;
  }
  if (h0) {
    T.cur = ch.x;
    return false;
  }
  if (h1) {
    T.cur = ch.y;
    return false;
  }
  if (T.sp == 0) {
    T.cur = -1;  // finished: the persistent loop ballots T.cur >= 0
    return true;
  }
  --T.sp;
  T.cur = (int)(T.sp < LS ? lds[__umul24((uint32_t)T.sp, stride)] : ovf[T.sp - LS]);
  return false;
}

// Dynamic LDS of the trace kernels: the per-lane stacks, entry j of thread t
// at word/pair j * blockDim.x + t (a wave's lanes hit distinct banks).
// Addresses are formed from this symbol at each use so that the accesses
// stay ds_read / ds_write.
extern __shared__ uint32_t trace_lds[];
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef __attribute__((address_space(3))) uint32_t lds_u32;

// Stack slot j of this thread, typed as LDS so that a conditional access to
// either the LDS part or the private part cannot be merged into one generic
// (flat) access.
// The stride is the launch's block size passed as a kernel argument (a
// uniform register; blockDim.x would be re-read from the dispatch packet with
// a vector load at every access), j * stride fits the 24-bit multiply.
__device__ __forceinline__ lds_u64* lds_slot(uint32_t j, uint32_t stride) {
  return (lds_u64*)(trace_lds) + (__umul24(j, stride) + threadIdx.x);
}

// tmin >= tnear > 0, so the float's bits order like the float; the low two
// bits carry the child slot (ties go to the lower slot).
__device__ __forceinline__ uint32_t sort_key(bool hit, float tmin, uint32_t c) {
  return hit ? ((__float_as_uint(tmin) & ~3u) | c) : 0xffffffffu;
}

// Child boxes of the four-wide node T.cur: entry distance and hit (slab
// interval non-empty within [tnear, tbest]) per child, and where its
// children are. Empty slots hold an inverted box, which no ray enters.
//
// Slab distances: the plane of byte q is org + q * 2^e, and
//   t = (org + q * 2^e) * inv - o * inv = q * (2^e * inv) + (org * inv - o * inv),
// one fma per plane with the two per-node factors.
struct node4_hits {
  float tmin[4];
  bool hit[4];
  uint32_t int_base;   // internal child c (c < nint): node int_base + c
  uint32_t leaf_info;  // triangle stores: first leaf; disc stores: internal children with leaves
  uint32_t nint, nchild;
  bool contiguous;     // leaves contiguous from leaf_info (a tree of triangles)
  uint4 lp;            // the leaf slots' primitives (a node of a disc cloud's tree)
};

// Primitive of leaf slot c (nint <= c < nchild).
__device__ __forceinline__ uint32_t leaf_prim(const node4_hits& H, uint32_t c) {
  if (H.contiguous) return H.leaf_info + c - H.nint;
  return c == 0 ? H.lp.x : c == 1 ? H.lp.y : c == 2 ? H.lp.z : H.lp.w;
}

// Reference to internal child `slot`: its index and, in a tree of discs, the
// parent's mark that it has leaf children.
__device__ __forceinline__ uint32_t child_ref(const node4_hits& H, bool contiguous, uint32_t slot) {
  const uint32_t mark = contiguous ? 0u : ((H.leaf_info >> slot) & 1u) << 30;
  return (H.int_base + slot) | mark;
}

// MESH: a launch over triangle meshes only, whose nodes all keep their
// leaves contiguous (no fourth word, no marks).
template <bool COUNT, bool MESH>
__device__ __forceinline__ void node4_boxes(const tparams& P, const trav_state& T, node4_hits& H, counters& C) {
  const uint4* q = reinterpret_cast<const uint4*>(
      P.nodes4 + (MESH ? (uint32_t)T.cur : ((uint32_t)T.cur & (kNodeHasLeaves - 1u))));
  const uint4 q0 = q[0], q1 = q[1], q2 = q[2];
  // the fourth word only for a node its parent marked (trees of discs)
  H.lp = make_uint4(0u, 0u, 0u, 0u);
  if (!MESH && ((uint32_t)T.cur & kNodeHasLeaves)) H.lp = q[3];
  if (COUNT) C.nodes++;
  const uint32_t meta = q0.w;
  H.int_base = q2.z;
  H.leaf_info = q2.w;
  H.nint = (meta >> 26) & 7u;
  H.nchild = 1u + ((meta >> 24) & 3u);
  H.contiguous = MESH || (meta & kNodeLeavesContiguous) != 0u;

  // 2^(e - 127) from each biased exponent byte (meta << 23 puts e_x in the
  // exponent field and one bit of e_y in the sign)
  const f3 inv = T.inv, oi = T.oi;
  const float ax = fabsf(__uint_as_float(meta << 23)) * inv.x;
  const float ay = __uint_as_float((meta << 15) & 0x7f800000u) * inv.y;
  const float az = __uint_as_float((meta << 7) & 0x7f800000u) * inv.z;
  const float bx = fmaf(__uint_as_float(q0.x), inv.x, -oi.x);
  const float by = fmaf(__uint_as_float(q0.y), inv.y, -oi.y);
  const float bz = fmaf(__uint_as_float(q0.z), inv.z, -oi.z);
  // entry / exit planes per axis follow the sign of the direction
  const uint32_t nx = inv.x >= 0.0f ? q1.x : q1.w, fx = inv.x >= 0.0f ? q1.w : q1.x;
  const uint32_t ny = inv.y >= 0.0f ? q1.y : q2.x, fy = inv.y >= 0.0f ? q2.x : q1.y;
  const uint32_t nz = inv.z >= 0.0f ? q1.z : q2.y, fz = inv.z >= 0.0f ? q2.y : q1.z;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int sh = 8 * c;
    const float tnx = fmaf((float)((nx >> sh) & 0xffu), ax, bx);
    const float tny = fmaf((float)((ny >> sh) & 0xffu), ay, by);
    const float tnz = fmaf((float)((nz >> sh) & 0xffu), az, bz);
    const float tfx = fmaf((float)((fx >> sh) & 0xffu), ax, bx);
    const float tfy = fmaf((float)((fy >> sh) & 0xffu), ay, by);
    const float tfz = fmaf((float)((fz >> sh) & 0xffu), az, bz);
    H.tmin[c] = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, T.tnear));
    const float tmax = fminf(fminf(tfx, tfy), fminf(tfz, T.tbest));
    H.hit[c] = H.tmin[c] <= tmax;
  }
}

// Slot mask of the hit leaf children.
__device__ __forceinline__ uint32_t leaf_mask(const node4_hits& H) {
  uint32_t leaves = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c)
    leaves |= (H.hit[c] && (uint32_t)c >= H.nint && (uint32_t)c < H.nchild) ? (1u << c) : 0u;
  return leaves;
}

// Hit internal children ordered by entry distance: the nearest becomes
// T.cur (with its distance in T.curd), the rest are pushed farthest first,
// each with its entry distance, so that a popped entry beyond the closest hit
// found meanwhile is dropped without fetching it. Children entered beyond
// tbest are not kept. Returns false when no internal child is kept.
// RECHECK: tbest may have dropped since the boxes were tested (leaves
// tested in between), so entry distances are compared with it again.
template <int LS, bool RECHECK>
__device__ __forceinline__ bool order_push(const tparams& P, trav_state& T, const node4_hits& H, uint64_t* ovf) {
  uint32_t k0 = sort_key(H.hit[0] && 0u < H.nint && (!RECHECK || H.tmin[0] <= T.tbest), H.tmin[0], 0);
  uint32_t k1 = sort_key(H.hit[1] && 1u < H.nint && (!RECHECK || H.tmin[1] <= T.tbest), H.tmin[1], 1);
  uint32_t k2 = sort_key(H.hit[2] && 2u < H.nint && (!RECHECK || H.tmin[2] <= T.tbest), H.tmin[2], 2);
  uint32_t k3 = sort_key(H.hit[3] && 3u < H.nint && (!RECHECK || H.tmin[3] <= T.tbest), H.tmin[3], 3);
  {
    uint32_t a = min(k0, k1), b = max(k0, k1);
    k0 = a; k1 = b;
    a = min(k2, k3); b = max(k2, k3);
    k2 = a; k3 = b;
    a = min(k0, k2); b = max(k0, k2);
    k0 = a; k2 = b;
    a = min(k1, k3); b = max(k1, k3);
    k1 = a; k3 = b;
    a = min(k1, k2); b = max(k1, k2);
    k1 = a; k2 = b;
  }
  if (k0 == 0xffffffffu) return false;
  const uint32_t ks[3] = {k3, k2, k1};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (ks[i] == 0xffffffffu) continue;
    // entry: node in the low word, entry distance bits in the high word
    const uint64_t e = ((uint64_t)(ks[i] & ~3u) << 32) | child_ref(H, H.contiguous, ks[i] & 3u);
    if (T.sp < LS) *lds_slot((uint32_t)T.sp, P.block) = e;
    else ovf[T.sp - LS] = e;
    ++T.sp;
  }
  T.cur = (int)child_ref(H, H.contiguous, k0 & 3u);
  T.curd = __uint_as_float(k0 & ~3u);
  return true;
}

// Pops entries until one lies within tbest (it becomes T.cur); false when
// the stack runs empty.
template <int LS>
__device__ __forceinline__ bool pop4(const tparams& P, trav_state& T, uint64_t* ovf) {
  while (T.sp > 0) {
    --T.sp;
    // the empty asm on the private read keeps the two reads from being merged
    // into one generic (flat) access through a selected pointer
    uint64_t e;
    if (T.sp < LS) {
      e = *lds_slot((uint32_t)T.sp, P.block);
    } else {
      e = ovf[T.sp - LS];
      __asm__ volatile("" : "+v"(e));
    }
    const float dist = __uint_as_float((uint32_t)(e >> 32));
    if (dist <= T.tbest) {
      T.cur = (int)(uint32_t)e;
      T.curd = dist;
      return true;
    }
  }
  return false;
}

// One four-wide node per call, leaf children tested right away (before the
// internal children are ordered, so they are culled with the updated closest
// hit). Returns true when the query is finished.
template <int LS, bool COUNT, bool MESH>
__device__ __forceinline__ bool trav4_step(const tparams& P, const f3 o, const f3 d, trav_state& T, uint64_t* ovf,
                                           counters& C) {
  node4_hits H;
  node4_boxes<COUNT, MESH>(P, T, H, C);
  uint32_t leaves = leaf_mask(H);
  if (leaves) {
    do {
      const uint32_t c = (uint32_t)(__ffs(leaves) - 1);
      leaves &= leaves - 1;
      test_prim<COUNT>(P, o, d, leaf_prim(H, c), T, C);
    } while (leaves);
  }
  if (order_push<LS, true>(P, T, H, ovf)) return false;
  if (pop4<LS>(P, T, ovf)) return false;
  T.cur = -1;  // finished: the persistent loop ballots T.cur >= 0
  return true;
}

// Postponed leaf tests. A wave tests leaves only in leaf phases instead of
// after each node visit, where a node visit's leaves keep the whole wave
// while a few lanes test them. Hit leaves go on a list shared by the wave's
// lanes, one 16-byte group per visited node with hit leaves: the node's four
// child words, a hit leaf child as a leaf word (owner lane and primitive id),
// any other slot 0. A node phase appends the groups with one 16-byte write
// per lane (a ballot ranks the writing lanes); a leaf phase takes the oldest
// groups whose leaves fit 64 lanes, compacts their leaf words in place (each
// group's lane writes its leaves at the exclusive scan of the groups' leaf
// counts) and gives one leaf word to each lane, whichever lane's ray it
// belongs to: the tester fetches the owner's origin, direction and tnear with
// ds_bpermute, tests the primitive, and resolves the owner's closest hit by a
// 64-bit LDS atomic minimum on keys (distance bits, primitive id): distances
// are >= 0, so their bits order like the distances, and equal distances go to
// the lower id, the order of test_prim. A lane's initial key is its closest
// hit so far, with id 0xffffffff while it has none (a primitive at the
// boundary-face distance then still wins), so the result does not depend on
// which lane tests which leaf, or when.
//
// LDS after the lane stacks: per wave a ring of kLeafGroups groups (512
// words), addressed from the wave's head group; then one 64-bit key per
// thread. A node phase adds at most one group per lane with a node, and a
// leaf phase is forced before one could overflow the ring.
constexpr uint32_t kLeafGroups = 128;
constexpr uint32_t kLeafWords = 4 * kLeafGroups;
// Leaf word: bit 31 set, owner lane in bits 25..30, primitive id below (so
// postponed leaves need primitive ids < 2^25; the host checks).
constexpr uint32_t kLeafPrimBits = 25;
constexpr uint32_t kLeafPrimMask = (1u << kLeafPrimBits) - 1u;

// Ray ids a persistent wave reserves beyond its lanes' needs.
constexpr unsigned long long kRayReserve = 256;

// Byte address of the wave's leaf ring (16 bytes per group).
template <int LS>
__device__ __forceinline__ uint32_t leaf_list_base(uint32_t stride) {
  return (uint32_t)(size_t)((lds_u64*)(trace_lds) + (LS * stride + (threadIdx.x >> 6) * (kLeafGroups * 2)));
}

// Byte address of the wave's first key (lane 0's).
template <int LS>
__device__ __forceinline__ uint32_t hit_key_base(uint32_t stride) {
  return (uint32_t)(size_t)((lds_u64*)(trace_lds) +
                            (LS * stride + (stride >> 6) * (kLeafGroups * 2) + (threadIdx.x & ~63u)));
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x4 lds_u128;

__device__ __forceinline__ lds_u64* lds_dword(uint32_t byte_addr) { return (lds_u64*)(size_t)byte_addr; }
__device__ __forceinline__ lds_u32* lds_word(uint32_t byte_addr) { return (lds_u32*)(size_t)byte_addr; }
__device__ __forceinline__ lds_u128* lds_quad(uint32_t byte_addr) { return (lds_u128*)(size_t)byte_addr; }

// Set bits of a wave mask in the lanes below this one (v_mbcnt: two
// instructions, against an and and a popcount per half).
__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ float lane_value(int src_byte, float v) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src_byte, __float_as_int(v)));
}

// Orders this wave's LDS accesses before the barrier ahead of those after it
// (one wave's LDS instructions execute in issue order; the fences keep the
// compiler from moving accesses across).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Closest-hit key of a distance (-0 taken as +0, so the bits order like
// the value) and primitive id.
__device__ __forceinline__ uint64_t hit_key(float t, uint32_t p) {
  return ((uint64_t)__float_as_uint(t + 0.0f) << 32) | p;
}

// Node phase: visits T.cur unless its entry distance lies beyond the closest
// hit found since it was chosen, descends into the nearest hit internal
// child or pops (T.cur = -1 when no node is left). Returns the visit's leaf
// group: slot c holds `owner | primitive` for a hit leaf child (owner: bit 31
// and the lane in bits 25..30), else 0.
template <int LS, bool COUNT, bool MESH>
__device__ __forceinline__ uint4 node4_visit(const tparams& P, trav_state& T, uint64_t* ovf, uint32_t owner,
                                             counters& C) {
  uint4 g = make_uint4(0u, 0u, 0u, 0u);
  if (T.curd <= T.tbest) {
    node4_hits H;
    node4_boxes<COUNT, MESH>(P, T, H, C);
    const uint32_t lm = leaf_mask(H);
    if (H.contiguous) {
      // leaf c is leaf_info + c - nint (no carry into the owner bits: ids < 2^25)
      const uint32_t lb = (owner | H.leaf_info) - H.nint;
      g.x = (lm & 1u) ? lb : 0u;
      g.y = (lm & 2u) ? lb + 1u : 0u;
      g.z = (lm & 4u) ? lb + 2u : 0u;
      g.w = (lm & 8u) ? lb + 3u : 0u;
    } else if (lm) {
      g.x = (lm & 1u) ? (owner | H.lp.x) : 0u;
      g.y = (lm & 2u) ? (owner | H.lp.y) : 0u;
      g.z = (lm & 4u) ? (owner | H.lp.z) : 0u;
      g.w = (lm & 8u) ? (owner | H.lp.w) : 0u;
    }
    if (order_push<LS, false>(P, T, H, ovf)) return g;
  }
  if (!pop4<LS>(P, T, ovf)) T.cur = -1;
  return g;
}

// Leaf phase over the ring's oldest groups (head group `head`, `avail` groups
// listed): takes the first groups whose leaves fit 64 lanes (at least one)
// and tests one leaf per lane; returns the groups taken, with their leaves in
// `took`. Runs with every lane of the wave active (ds_bpermute reads the
// owners' registers).
template <bool COUNT>
__device__ __forceinline__ uint32_t leaf_phase(const tparams& P, const f3 ro, const f3 rd, trav_state& T,
                                               uint32_t list, uint32_t keys, uint32_t head, uint32_t avail,
                                               uint32_t lane, uint32_t& took, counters& C) {
  lds_u64* mine = lds_dword(keys + lane * 8u);
  *mine = hit_key(T.tbest, T.found ? T.pbest : 0xffffffffu);
  uint4 g = make_uint4(0u, 0u, 0u, 0u);
  if (lane < avail && lane < 64u) {
    const u32x4 v = *lds_quad(list + ((head + lane) & (kLeafGroups - 1u)) * 16u);
    g = make_uint4(v.x, v.y, v.z, v.w);
  }
  const uint32_t n0 = g.x != 0u, n1 = g.y != 0u, n2 = g.z != 0u;
  const uint32_t n = n0 + n1 + n2 + (g.w != 0u ? 1u : 0u);
  // exclusive scan of the groups' leaf counts (0..4) from three ballots
  const unsigned long long b0 = __ballot(n & 1u), b1 = __ballot(n & 2u), b2 = __ballot(n & 4u);
  const uint32_t before = lanes_below(b0) + 2u * lanes_below(b1) + 4u * lanes_below(b2);
  // the scan is nondecreasing in the lane, so the groups that fit are a prefix
  const unsigned long long fit = __ballot(before + n <= 64u && lane < avail);
  const uint32_t ng = (uint32_t)__popcll(fit);
  const uint32_t take =
      (uint32_t)(__popcll(b0 & fit) + 2 * __popcll(b1 & fit) + 4 * __popcll(b2 & fit));
  // the taken groups' leaf words, compacted in place over the groups' own
  // words: leaf t goes to word t of the taken region, and t < 4 * (group + 1)
  // never reaches a word of a later group before that group's lane has read it
  // (all reads above completed before any write below)
  const uint32_t w0 = head * 4u;
  wave_lds_sync();
  if ((fit >> lane) & 1ull) {
    uint32_t t = w0 + before;
    if (g.x) { *lds_word(list + (t & (kLeafWords - 1u)) * 4u) = g.x; ++t; }
    if (g.y) { *lds_word(list + (t & (kLeafWords - 1u)) * 4u) = g.y; ++t; }
    if (g.z) { *lds_word(list + (t & (kLeafWords - 1u)) * 4u) = g.z; ++t; }
    if (g.w) { *lds_word(list + (t & (kLeafWords - 1u)) * 4u) = g.w; }
  }
  wave_lds_sync();
  const bool tester = lane < take;
  uint32_t e = 0u;
  if (tester) e = *lds_word(list + ((w0 + lane) & (kLeafWords - 1u)) * 4u);
  const uint32_t owner = (e >> kLeafPrimBits) & 63u;
  const int src = (int)(owner << 2);
  const f3 o = make(lane_value(src, ro.x), lane_value(src, ro.y), lane_value(src, ro.z));
  const f3 d = make(lane_value(src, rd.x), lane_value(src, rd.y), lane_value(src, rd.z));
  const float tn = lane_value(src, T.tnear);
  if (tester) {
    const uint32_t p = e & kLeafPrimMask;
    float th;
    if (prim_hit<COUNT>(P, o, d, tn, p, th, C))
      __hip_atomic_fetch_min(lds_dword(keys + owner * 8u), hit_key(th, p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WAVEFRONT);
  }
  wave_lds_sync();
  const uint64_t k = *mine;
  T.tbest = __uint_as_float((uint32_t)(k >> 32));
  T.pbest = (uint32_t)k;
  T.found = T.pbest != 0xffffffffu;
  took = take;
  return ng;
}

// Per-lane traversal stack of either tree: LS entries in LDS, the rest in a
// private array.
template <int LS, int W, bool MESH>
struct lane_stack;

template <int LS, bool MESH>
struct lane_stack<LS, 2, MESH> {
  uint32_t ovf[64 - LS];
  template <bool COUNT>
  __device__ __forceinline__ bool step(const tparams& P, const f3 o, const f3 d, trav_state& T, counters& C) {
    return trav_step<LS, COUNT>(P, o, d, T, trace_lds + threadIdx.x, ovf, C);
  }
};

template <int LS, bool MESH>
struct lane_stack<LS, 4, MESH> {
  uint64_t ovf[kWideStack - LS];
  template <bool COUNT>
  __device__ __forceinline__ bool step(const tparams& P, const f3 o, const f3 d, trav_state& T, counters& C) {
    return trav4_step<LS, COUNT, MESH>(P, o, d, T, ovf, C);
  }
  template <bool COUNT>
  __device__ __forceinline__ uint4 visit(const tparams& P, trav_state& T, uint32_t owner, counters& C) {
    return node4_visit<LS, COUNT, MESH>(P, T, ovf, owner, C);
  }
};

// Lateral boundary faces x = xmin/xmax, y = ymin/ymax over z in [zmin, zmax]
// (src/rti/geo/boundary_x_y.hpp:147-235; no top or bottom face).
// Face ids: 0 xmax, 1 ymin, 2 xmin, 3 ymax; -1 none.
// All four faces are tested, as a closest-hit query against the eight
// boundary triangles does: an origin may lie outside the box (a disc-source
// origin beyond a box narrower in y than in x, a reflection off a disc that
// sticks out of it, a reflective bounce rounded just past the face), where
// the face behind the direction's sign or the one beyond the nearer plane
// can be the first one at t >= tnear. Later faces win ties (t <= best).
__device__ __forceinline__ void face_test(int f, float t, float a, float alo, float ahi, float z,
                                          const inst_desc& D, float tnear, float& best, int& face) {
  if (t >= tnear && t <= best && a >= alo && a <= ahi && z >= D.zmin && z <= D.zmax) {
    best = t;
    face = f;
  }
}

// rx, ry: hardware reciprocals (1 ulp) of d.x and d.y, as for the slab tests.
__device__ __forceinline__ int boundary_hit(const inst_desc& D, f3 o, f3 d, float rx, float ry, float tnear,
                                            float& tb) {
  int face = -1;
  float best = FLT_MAX;
  if (d.x != 0.0f) {
    float t = (D.xmax - o.x) * rx;
    face_test(0, t, fmaf(d.y, t, o.y), D.ymin, D.ymax, fmaf(d.z, t, o.z), D, tnear, best, face);
    t = (D.xmin - o.x) * rx;
    face_test(2, t, fmaf(d.y, t, o.y), D.ymin, D.ymax, fmaf(d.z, t, o.z), D, tnear, best, face);
  }
  if (d.y != 0.0f) {
    float t = (D.ymin - o.y) * ry;
    face_test(1, t, fmaf(d.x, t, o.x), D.xmin, D.xmax, fmaf(d.z, t, o.z), D, tnear, best, face);
    t = (D.ymax - o.y) * ry;
    face_test(3, t, fmaf(d.x, t, o.x), D.xmin, D.xmax, fmaf(d.z, t, o.z), D, tnear, best, face);
  }
  tb = best;
  return face;
}

// Neighbour test: restates local_intersector::intersect
// (src/rti/trace/local_intersector.hpp:41-97).
__device__ __forceinline__ bool local_intersect(f3 o, f3 d, float4 A, float4 B) {
  const float pod = B.x * d.x + B.y * d.y + B.z * d.z;
  if (pod > 0.0f) return false;
  if (fabsf(pod) < 1e-9f) return false;
  const float ddneg = A.x * B.x + A.y * B.y + A.z * B.z;
  const float tt = (-(B.x * o.x + B.y * o.y + B.z * o.z) + ddneg) / pod;
  if (tt <= 0.0f) return false;
  const f3 e = make(o.x + d.x * tt - A.x, o.y + d.y * tt - A.y, o.z + d.z * tt - A.z);
  return A.w > sqrtf(dot(e, e));
}

// Orthonormal basis (Duff et al., JCGT 6(1), 2017) and the cosine lobe of
// cos_hemi::get (src/rti/ray/cos_hemi.hpp:10-45).
__device__ __forceinline__ f3 diffuse_dir(f3 n, float r1, float r2) {
  const float sign = copysignf(1.0f, n.z);
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  const f3 b1 = make(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  const f3 b2 = make(b, sign + n.y * n.y * a, -n.y);
  float s, c;
  sincospif(2.0f * r1, &s, &c);
  const float c1 = sqrtf(r2);
  const float st = sqrtf(1.0f - r2);
  return normalize(n * c1 + b1 * (c * st) + b2 * (s * st));
}

// One deposit v = drop on accumulator slot a. Fixed point: the products
// v * 2^shift and the adjoint weights * 2^(shift - 8) are exact in FP32
// (power-of-two scales) and are rounded once, to nearest, to 64-bit
// integers; v^2 is formed in FP32 (one rounding, 2^-24 relative) and put on
// its own grid 2^shift2, chosen from the sticking so that the smallest
// squared deposit still spans many grid units. No FP64 arithmetic unless
// FP64 accumulation or the higher moments are asked for.
template <bool ADJ, bool FP64ACC, bool COUNT>
__device__ __forceinline__ void deposit(const tparams& P, uint32_t a_slot, float drop, float scale, float scale2,
                                        float dds, float ddn, counters& C) {
  if (COUNT) C.deps++;
  unsigned long long* a = P.acc + (size_t)kAccStride * a_slot;  // one 32-byte record
  if (FP64ACC) {
    const double v = (double)drop;
    unsafeAtomicAdd(reinterpret_cast<double*>(a), v);
    unsafeAtomicAdd(reinterpret_cast<double*>(a + 1), v * v);
  } else {
    atomicAdd(a, (unsigned long long)__float2ll_rn(drop * scale));
    atomicAdd(a + 1, (unsigned long long)__float2ll_rn((drop * drop) * scale2));
  }
  atomicAdd(a + 2, 1ull);
  if (P.mom) {
    // higher moments span too many orders of magnitude for a fixed-point
    // grid; they use FP64 atomics (hit_accumulator.hpp:128-138)
    const double v = (double)drop;
    const double v2 = v * v;
    unsafeAtomicAdd(P.mom + 2 * (size_t)a_slot, v2 * v);
    unsafeAtomicAdd(P.mom + 2 * (size_t)a_slot + 1, v2 * v * v);
  }
  if (ADJ) {
    // derivative weights can exceed 1 (by about k s / (1 - s)); they use a
    // scale 2^8 smaller than the primal sums
    const float sa = scale * (1.0f / 256.0f);
    atomicAdd(P.adj + 2 * (size_t)a_slot, (unsigned long long)__float2ll_rn(dds * sa));
    atomicAdd(P.adj + 2 * (size_t)a_slot + 1, (unsigned long long)__float2ll_rn(ddn * sa));
  }
}

struct ray_state {
  f3 o, d;
  float w;
  float pk;       // (1 - s)^k, adjoint only
  float rfac;     // product of roulette boosts, adjoint only
  float score;    // d log p(direction) / dn, adjoint only
  uint64_t id;    // ray id local to the instance
  uint32_t inst;
  uint32_t event;
  uint32_t k;     // front-face hits so far
  uint32_t segs;
};

__device__ __forceinline__ uint32_t find_instance(const tparams& P, uint64_t gid) {
  uint32_t lo = 0, hi = P.n_inst - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (P.desc[mid].ray_begin <= gid) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Primary ray: rectangle_origin_z::get (src/rti/ray/rectangle_origin_z.hpp:23-37)
// or disc_origin_z::get (src/rti/ray/disc_origin_z.hpp:17-33), and
// power_cosine_direction_z::get (src/rti/ray/power_cosine_direction_z.hpp:17-32),
// cosine_direction_z for n = 1 (src/rti/ray/cosine_direction_z.hpp:13-27).
// Source disc of create_circular_source_from_bounding_box (src/rti/device.hpp:269-280):
// centre of the box, radius half its x extent. Offsets uniform in [-r, r]^2,
// redrawn until inside the disc; attempt a > 0 draws Philox block
// (id, 0, a), which no segment draw uses (those have word 2 >= 1, word 3 = 0).
// After kMaxSourceAttempts rejections (probability (1 - pi/4)^64 < 1e-42)
// the ray starts at the centre, so the loop is bounded.
constexpr uint32_t kMaxSourceAttempts = 64;

__device__ __forceinline__ f3 disc_origin(const inst_desc& D, uint64_t id, uint32_t ux, uint32_t uy) {
  const float cx = (D.xmin + D.xmax) * 0.5f, cy = (D.ymin + D.ymax) * 0.5f;
  const float rad = (D.xmax - D.xmin) * 0.5f;
  for (uint32_t a = 1;; ++a) {
    const float dx = fmaf(2.0f, u01(ux), -1.0f) * rad;
    const float dy = fmaf(2.0f, u01(uy), -1.0f) * rad;
    if (sqrtf(fmaf(dx, dx, dy * dy)) <= rad) return make(cx + dx, cy + dy, D.src_z);
    if (a == kMaxSourceAttempts) return make(cx, cy, D.src_z);
    const u4 v = philox({(uint32_t)id, (uint32_t)(id >> 32), 0u, a}, D.seed, D.key1);
    ux = v.x;
    uy = v.y;
  }
}

__device__ __forceinline__ void start_ray(const tparams& P, uint64_t gid, ray_state& R) {
  R.inst = find_instance(P, gid);
  const inst_desc& D = P.desc[R.inst];
  // the ray's id within its instance's run (the Philox counter), so that a
  // run of rays [b, b + n) traces exactly those rays of a larger run
  R.id = gid - D.ray_begin + D.ray_id0;
  const u4 u = philox({(uint32_t)R.id, (uint32_t)(R.id >> 32), 0u, 0u}, D.seed, D.key1);
  if (D.source == 0) {
    R.o = make(fmaf(D.xmax - D.xmin, u01(u.x), D.xmin), fmaf(D.ymax - D.ymin, u01(u.y), D.ymin), D.src_z);
  } else {
    R.o = disc_origin(D, R.id, u.x, u.y);
  }
  const float r1 = u01(u.z), r2 = u01_open0(u.w);
  const float tt = D.cos_exponent == 1.0f ? r2 : powf(r2, D.pow_e);
  float s, c;
  sincospif(2.0f * r1, &s, &c);
  const float st = sqrtf(1.0f - tt);
  R.d = normalize(make(c * st, s * st, -sqrtf(tt)));
  R.w = 1.0f;
  R.pk = 1.0f;
  R.rfac = 1.0f;
  R.score = D.score_c + logf(r2) * D.score_c;
  R.event = 1;
  R.k = 0;
  R.segs = 0;
}

// Start of a segment: the boundary-face distance bounds the closest-hit
// query, which starts at the instance's root.
// qseq: the wave's leaf-list head sequence number (postponed leaves; no
// entry of this lane is pending).
template <int W, bool MESH>
__device__ __forceinline__ void begin_segment(const tparams& P, ray_state& R, trav_state& T, counters& C,
                                              uint32_t qseq) {
  const inst_desc& D = P.desc[R.inst];
  C.seg++;
  R.segs++;
  const float tiny = 1e-20f;
  // slab factors only cull boxes (every box is padded beyond its
  // primitives), so the hardware reciprocal (1 ulp) suffices; the
  // boundary-face distances multiply by it too (Embree's own intersectors use
  // a refined reciprocal, not a correctly rounded division, so neither form
  // is the reference's bits)
  T.inv.x = __builtin_amdgcn_rcpf(fabsf(R.d.x) > tiny ? R.d.x : copysignf(tiny, R.d.x));
  T.inv.y = __builtin_amdgcn_rcpf(fabsf(R.d.y) > tiny ? R.d.y : copysignf(tiny, R.d.y));
  T.inv.z = __builtin_amdgcn_rcpf(fabsf(R.d.z) > tiny ? R.d.z : copysignf(tiny, R.d.z));
  T.tnear = D.tnear;
  T.face = boundary_hit(D, R.o, R.d, T.inv.x, T.inv.y, T.tnear, T.tb);
  T.tbest = T.tb;
  T.found = false;
  T.pbest = 0;
  // the root of a tree of discs loads its fourth word (it is nobody's child)
  T.cur = W == 4 ? (int)((uint32_t)D.root4 | (!MESH && D.prim_kind == 0 ? kNodeHasLeaves : 0u)) : D.root;
  T.curd = 0.0f;
  T.sp = 0;
  T.lastq = qseq;
  T.oi = make(R.o.x * T.inv.x, R.o.y * T.inv.y, R.o.z * T.inv.z);
}

// The consequences of a finished closest-hit query (tracer.hpp:184-253).
// Returns false when the ray terminates (the per-ray segment cap is applied
// by the caller).
template <bool ADJ, bool FP64ACC, bool COUNT>
__device__ __forceinline__ bool shade_segment(const tparams& P, ray_state& R, const trav_state& T, counters& C) {
  const inst_desc& D = P.desc[R.inst];
  // each counter is updated at one place, by straight-line arithmetic, so
  // the counters stay in registers
  C.miss += (!T.found && T.face < 0) ? 1u : 0u;
  if (!T.found) {
    if (T.face < 0) return false;  // tracer.hpp:184-190
    // boundary_x_y::process_hit (src/rti/geo/boundary_x_y.hpp:91-143)
    const f3 h = R.o + R.d * T.tb;
    if (T.face == 0 || T.face == 2) {
      if (D.bc_x == 0) { R.d.x = -R.d.x; R.o = h; }
      else R.o = make(T.face == 0 ? D.xmin : D.xmax, h.y, h.z);
    } else {
      if (D.bc_y == 0) { R.d.y = -R.d.y; R.o = h; }
      else R.o = make(h.x, T.face == 1 ? D.ymax : D.ymin, h.z);
    }
    return true;
  }
  const uint32_t prim = T.pbest;
  const float tp = T.tbest;
  // the record's last float4 is the unit normal
  const float4 N4 = P.prims[((size_t)(prim + 1) << P.prim_shift) - 1];
  const f3 n = make(N4.x, N4.y, N4.z);
  const bool disc = D.prim_kind == 0;
  if (dot(R.d, n) > 0.0f) {  // back face: pass through (tracer.hpp:212-227)
    R.o = R.o + R.d * tp;
    return true;
  }
  C.hit++;
  // stick and deposit (tracer.hpp:228-237)
  const float s = D.sticking;
  const float drop = R.w * s;
  float dds = 0.0f, ddn = 0.0f;
  if (ADJ) {
    // d/ds of s (1-s)^k, times the roulette boosts of this path
    const float g = (R.k == 0) ? 1.0f : R.pk * (1.0f - (float)R.k * s / (1.0f - s));
    dds = g * R.rfac;
    ddn = drop * R.score;
  }
  const float scale = D.fx_scale, scale2 = D.fx_scale2;
  const int32_t acc_off = D.acc_off;
  deposit<ADJ, FP64ACC, COUNT>(P, (uint32_t)((int32_t)prim + acc_off), drop, scale, scale2, dds, ddn, C);
  if (disc) {  // check_for_additional_intersections (tracer.hpp:339-383)
    // the neighbour row starts at the w word of this record's last float4
    // (loaded with the normal) and ends at that of the next record; the
    // column indices of up to four neighbours are loaded together, so the
    // records wait for one index latency per four neighbours
    const uint32_t q0 = __float_as_uint(N4.w);
    const uint32_t q1 = __float_as_uint(P.prims[((size_t)(prim + 2) << P.prim_shift) - 1].w);
    for (uint32_t q = q0; q < q1; q += 4) {
      uint32_t j[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) j[u] = (q + u < q1) ? P.nb_col[q + u] : prim;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (q + u >= q1) break;
        if (COUNT) C.nbt++;
        const float4* nr = P.prims + ((size_t)j[u] << P.prim_shift);
        if (local_intersect(R.o, R.d, nr[0], nr[1]))
          deposit<ADJ, FP64ACC, COUNT>(P, (uint32_t)((int32_t)j[u] + acc_off), drop, scale, scale2, dds, ddn, C);
      }
    }
  }
  R.w -= drop;
  if (R.w == 0.0f) return false;  // tracer.hpp:238-240
  R.k++;
  if (ADJ) R.pk *= (1.0f - s);
  const u4 u = philox({(uint32_t)R.id, (uint32_t)(R.id >> 32), R.event, 0u}, D.seed, D.key1);
  R.event++;
  // rejection_control::check_weight_reweight_or_kill (src/rti/mc/rejection_control.hpp:27-61)
  if (D.roulette && R.w < 0.1f) {
    if (u01(u.x) < 1.0f - R.w / 0.3f) return false;
    if (ADJ) R.rfac *= 0.3f / R.w;
    R.w = 0.3f;
  }
  R.o = R.o + R.d * tp;  // meta_geometry::get_new_origin (src/rti/geo/meta_geometry.hpp:17-23)
  // cos^2 of the diffuse direction is drawn from (0, 1]: from [0, 1) a
  // 24-bit draw of 0 (one bounce in 2^24) would give a direction exactly in
  // the surface's tangent plane, a ray that travels inside the plane of the
  // surface it left and, on a flat periodic top surface, never meets it again
  // (tens of thousands of boundary segments); the reference's 64-bit draws
  // (cos_hemi.hpp:25-26 over mt19937_64) make that a 2^-64 event
  if (u01(u.w) < D.diffuse_prob) R.d = diffuse_dir(n, u01(u.y), u01_open0(u.z));
  else R.d = R.d - n * (2.0f * dot(R.d, n));  // specular::use (src/rti/reflection/specular.hpp:16-40)
  return true;
}

template <bool ADJ, bool FP64ACC, bool COUNT>
__device__ __forceinline__ bool finish_segment(const tparams& P, ray_state& R, const trav_state& T, counters& C) {
  const bool more = shade_segment<ADJ, FP64ACC, COUNT>(P, R, T, C);
  const bool cap = more && R.segs >= kMaxSegmentsPerRay;
  C.trunc += cap ? 1u : 0u;
  if (!more || cap) C.longest = R.segs > C.longest ? R.segs : C.longest;
  return more && !cap;
}

// Wave reduction of the counters, then one atomic per wave and counter. The
// counters stay in registers: every index below is a compile-time constant.
__device__ __forceinline__ void flush_one(const tparams& P, int i, uint32_t v) {
  unsigned long long x = v;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  if ((threadIdx.x & 63) == 0 && x) atomicAdd(P.stats + i, x);
}

template <bool COUNT>
__device__ __forceinline__ void flush_counters(const tparams& P, const counters& C) {
  {
    uint32_t m = C.longest;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const uint32_t o = (uint32_t)__shfl_xor((int)m, off, 64);
      m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0 && m) atomicMax(P.stats + 8, (unsigned long long)m);
  }
  flush_one(P, 0, C.seg);
  flush_one(P, 1, C.hit);
  flush_one(P, 2, C.miss);
  flush_one(P, 3, C.trunc);
  if (COUNT) {
    flush_one(P, 4, C.nodes);
    flush_one(P, 5, C.prims);
    flush_one(P, 6, C.nbt);
    flush_one(P, 7, C.deps);
  }
}

// Persistent, wavefront-scheduled trace. Each wave holds up to 64 live rays
// and alternates two phases:
//   traverse: every lane with an open closest-hit query advances one node per
//     iteration, until at most a fraction (64 - compact) / 64 of the live
//     lanes still have an open query;
//   shade + refill: the lanes whose query finished process the hit and start
//     their next segment; lanes whose ray ended take new rays from the wave's
//     reservation of queue ids, topped up by one atomic per wave when it runs
//     short (ballot + popcount assign consecutive ids).
// Register budget: 7 resident waves per SIMD (72 VGPRs) measured 16 % faster
// than the unconstrained 88-VGPR allocation on the triangle-hole workload;
// the compiler spills about 20 registers, all outside the traversal loop
// (around shading and the disc-neighbour loop, once per shading pass), and
// that loop's only scratch traffic is the stack beyond its LDS entries. The
// adjoint variant carries four more live values per ray and gets 6 (80
// VGPRs). The launch can still run fewer waves per CU
// (waves_per_cu) where the memory system prefers it.
//
// With postponed leaves (PL, four-wide tree) each traversal iteration is
// either a node phase (lanes with a node left visit it and append its hit
// leaves to the wave's list) or a leaf phase (up to 64 list entries tested,
// one per lane), chosen for the whole wave: a leaf phase runs once the
// entries it would take reach leaf_min / 64 of the lanes with a node left,
// when the list could not take another node phase's leaves, or when no node
// is left. A lane's query is open while it has a node or a listed leaf left.
template <int LS, int W, bool ADJ, bool FP64ACC, bool COUNT, bool PL, bool MESH>
__global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(ADJ ? 6 : 7))) void k_trace_persistent(
    tparams P) {
  lane_stack<LS, W, MESH> S;
  counters C;
  ray_state R;
  trav_state T;
  bool alive = false, exhausted = false, tracing = false;
  const int lane = threadIdx.x & 63;
  const int compact = P.compact;
  // Ray ids come from the global queue in blocks: a wave reserves
  // kRayReserve ids beyond what its lanes need, so most refills hand out ids
  // without an atomic. Reserving stops once the queue is within tail ids of
  // its end, so that the last rays spread over all waves. Ids are handed out
  // in increasing order per wave, each exactly once; which lane traces a ray
  // does not change any result.
  unsigned long long res_base = 0, res_end = 0;  // reserved ids [res_base, res_end), wave-uniform
  const unsigned long long tail = (unsigned long long)gridDim.x * (P.block >> 6) * 64ull * 16ull;
  // the wave's leaf list (postponed leaves): wave-uniform head group of the
  // ring, its sequence number, groups and leaves listed
  const uint32_t list = PL ? leaf_list_base<LS>(P.block) : 0u;
  const uint32_t keys = PL ? hit_key_base<LS>(P.block) : 0u;
  const uint32_t owner = 0x80000000u | ((uint32_t)lane << kLeafPrimBits);
  uint32_t lhead = 0, lseq = 0, lcount = 0, lleaves = 0;
  for (;;) {
    // ended rays are replaced once refill_min lanes wait for one (a refill
    // costs the wave the same whatever the number of lanes it serves), or
    // when no lane is alive
    const unsigned long long need = __ballot(!alive && !exhausted);
    if (need && (__popcll(need) >= P.refill_min || __ballot(alive) == 0ull)) {
      const unsigned long long cnt = (unsigned long long)__popcll(need);
      const unsigned long long rank = lanes_below(need);
      const unsigned long long have = res_end - res_base;
      unsigned long long my;
      if (cnt <= have) {
        my = res_base + rank;
        res_base += cnt;
      } else {
        const int leader = __ffsll((long long)need) - 1;
        const unsigned long long more = cnt - have;
        const unsigned long long extra = (res_end + tail < P.ray_hi) ? kRayReserve : 0ull;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(P.queue, more + extra);
        base = __shfl(base, leader, 64);
        my = rank < have ? res_base + rank : base + (rank - have);
        res_base = base + more;
        res_end = base + more + extra;
      }
      if (!alive && !exhausted) {
        if (my < P.ray_hi) {
          start_ray(P, my, R);
          begin_segment<W, MESH>(P, R, T, C, lseq);
          alive = true;
          tracing = true;
        } else {
          exhausted = true;
        }
      }
    }
    const unsigned long long live = __ballot(alive);
    if (live == 0ull) break;
    // At most (64 - compact) / 64 of the live lanes may still be tracing when
    // the wave shades. With a full wave that is 64 - compact lanes; while the
    // wave drains (queue empty, fewer live lanes) the bound shrinks with it,
    // so a few long paths still advance many nodes per shading pass.
    const int still = ((64 - compact) * __popcll(live)) >> 6;
    if constexpr (PL) {
      // tr mirrors the lanes' tracing flags as a wave mask. It is kept up to
      // date from ballots of the plain comparisons (each one compare into a
      // scalar mask, shared with the lanes' own branch conditions) rather
      // than re-balloted from the flags, whose ballot would first turn the
      // flag into a 0/1 register and compare that.
      unsigned long long tr = __ballot(tracing);
      for (;;) {
        const uint32_t nl = (uint32_t)__popcll(tr & __ballot(T.cur >= 0));
        const uint32_t take = lleaves < 64u ? lleaves : 64u;
        const bool leaf_phase_now =
            lcount != 0u && (nl == 0u || lcount + nl > kLeafGroups || take * 64u >= (uint32_t)P.leaf_min * nl);
        if (leaf_phase_now) {
          uint32_t took;
          const uint32_t ng = leaf_phase<COUNT>(P, R.o, R.d, T, list, keys, lhead, lcount, (uint32_t)lane, took, C);
          lhead = (lhead + ng) & (kLeafGroups - 1u);
          lseq += ng;
          lcount -= ng;
          lleaves -= took;
        } else {
          uint4 g = make_uint4(0u, 0u, 0u, 0u);
          if (tracing && T.cur >= 0) g = S.template visit<COUNT>(P, T, owner, C);
          // append the visit's group: one 16-byte write per lane with hit leaves
          const unsigned long long bx = __ballot(g.x != 0u), by = __ballot(g.y != 0u);
          const unsigned long long bz = __ballot(g.z != 0u), bw = __ballot(g.w != 0u);
          const unsigned long long bg = bx | by | bz | bw;
          if ((g.x | g.y | g.z | g.w) != 0u) {
            const uint32_t rank = lanes_below(bg);
            // the list's end: lhead + lcount, lcount + nl <= kLeafGroups here
            u32x4 v;
            v.x = g.x; v.y = g.y; v.z = g.z; v.w = g.w;
            *lds_quad(list + ((lhead + lcount + rank) & (kLeafGroups - 1u)) * 16u) = v;
            T.lastq = lseq + lcount + rank + 1u;
          }
          lcount += (uint32_t)__popcll(bg);
          lleaves += (uint32_t)(__popcll(bx) + __popcll(by) + __popcll(bz) + __popcll(bw));
        }
        const bool pend = (int32_t)(T.lastq - lseq) > 0;
        tr &= __ballot(T.cur >= 0) | __ballot(pend);
        tracing = tracing && (T.cur >= 0 || pend);
        if (__popcll(tr) <= still) break;
      }
    } else {
      // a finished query leaves T.cur = -1, so the wave mask of tracing
      // lanes follows from one compare (see the postponed-leaf loop above)
      unsigned long long tr = __ballot(tracing);
      for (;;) {
        if (tracing) S.template step<COUNT>(P, R.o, R.d, T, C);
        tr &= __ballot(T.cur >= 0);
        tracing = tracing && T.cur >= 0;
        if (__popcll(tr) <= still) break;
      }
    }
    if (alive && !tracing) {
      alive = finish_segment<ADJ, FP64ACC, COUNT>(P, R, T, C);
      if (alive) {
        begin_segment<W, MESH>(P, R, T, C, lseq);
        tracing = true;
      }
    }
  }
  flush_counters<COUNT>(P, C);
}

// Plain trace: one thread per ray, every segment traversed to the end.
template <int LS, int W, bool ADJ, bool FP64ACC, bool COUNT, bool MESH>
__global__ __launch_bounds__(kMaxBlock) void k_trace_plain(tparams P) {
  lane_stack<LS, W, MESH> S;
  counters C;
  ray_state R;
  trav_state T;
  const unsigned long long gid = P.ray_lo + (unsigned long long)blockIdx.x * P.block + threadIdx.x;
  if (gid < P.ray_hi) {
    start_ray(P, gid, R);
    for (;;) {
      begin_segment<W, MESH>(P, R, T, C, 0u);
      while (!S.template step<COUNT>(P, R.o, R.d, T, C)) {
      }
      if (!finish_segment<ADJ, FP64ACC, COUNT>(P, R, T, C)) break;
    }
  }
  flush_counters<COUNT>(P, C);
}

using kfn = void (*)(tparams);

template <int LS, int W, bool PL, bool MESH>
kfn pick_persistent(bool adj, bool fp64, bool count) {
  if (count) return k_trace_persistent<LS, W, false, false, true, PL, MESH>;
  if (adj)
    return fp64 ? k_trace_persistent<LS, W, true, true, false, PL, MESH>
                : k_trace_persistent<LS, W, true, false, false, PL, MESH>;
  return fp64 ? k_trace_persistent<LS, W, false, true, false, PL, MESH>
              : k_trace_persistent<LS, W, false, false, false, PL, MESH>;
}

template <int LS, int W, bool MESH>
kfn pick(bool persistent, bool postpone, bool adj, bool fp64, bool count) {
  if (persistent) {
    if constexpr (W == 4) {
      if (postpone) return pick_persistent<LS, W, true, MESH>(adj, fp64, count);
    }
    return pick_persistent<LS, W, false, MESH>(adj, fp64, count);
  }
  if (count) return k_trace_plain<LS, W, false, false, true, MESH>;
  if (adj) return fp64 ? k_trace_plain<LS, W, true, true, false, MESH> : k_trace_plain<LS, W, true, false, false, MESH>;
  return fp64 ? k_trace_plain<LS, W, false, true, false, MESH> : k_trace_plain<LS, W, false, false, false, MESH>;
}

template <int W, bool MESH>
kfn pick_ls(int lds_stack, bool persistent, bool postpone, bool adj, bool fp64, bool count) {
  switch (lds_stack) {
    case 4: return pick<4, W, MESH>(persistent, postpone, adj, fp64, count);
    case 6: return pick<6, W, MESH>(persistent, postpone, adj, fp64, count);
    case 8: return pick<8, W, MESH>(persistent, postpone, adj, fp64, count);
    case 16: return pick<16, W, MESH>(persistent, postpone, adj, fp64, count);
    case 24: return pick<24, W, MESH>(persistent, postpone, adj, fp64, count);
    case 32: return pick<32, W, MESH>(persistent, postpone, adj, fp64, count);
    default: return nullptr;
  }
}

// mesh: every instance of the launch is a triangle mesh (leaves contiguous
// in every node; the four-wide kernels then drop the disc clouds' leaf word).
kfn pick_kernel(int width, int lds_stack, bool persistent, bool postpone, bool adj, bool fp64, bool count,
                bool mesh) {
  if (width == 2) return pick_ls<2, false>(lds_stack, persistent, false, adj, fp64, count);
  if (width == 4)
    return mesh ? pick_ls<4, true>(lds_stack, persistent, postpone, adj, fp64, count)
                : pick_ls<4, false>(lds_stack, persistent, postpone, adj, fp64, count);
  return nullptr;
}

// Dynamic LDS: the lane stacks, then (postponed leaves) the waves' leaf
// lists and the threads' closest-hit keys.
size_t trace_lds_bytes(int width, int block_size, int lds_stack, bool postpone) {
  size_t b = (size_t)block_size * lds_stack * (width == 4 ? 8 : 4);
  if (width == 4 && postpone) b += (size_t)(block_size / 64) * kLeafGroups * 16 + (size_t)block_size * 8;
  return b;
}

}  // namespace

// Six entries beside a leaf queue keep the block's LDS within what all the
// register-limited waves need; eight without one.
int default_lds_stack(int width, bool postpone) { return width == 4 ? (postpone ? 6 : 8) : 16; }
int default_compact_lanes() { return 56; }
// Postponed leaves pay off where a leaf test is expensive: triangle stores
// (measured +27 % on a 199k-triangle hole, flat within 1 % for thresholds
// 64..192) but not stores of discs only, whose leaf test is a few
// instructions (batch of disc clouds -2 %, rough surface -2 % at 64).
int default_leaf_batch(bool triangles) { return triangles ? 128 : 0; }
uint64_t postponed_leaf_prim_limit() { return 1ull << kLeafPrimBits; }
int wide_stack_capacity() { return kWideStack; }

hipError_t launch_trace(const trace_launch& L, hipStream_t stream) {
  tparams P;
  P.desc = L.desc;
  P.n_inst = L.n_inst;
  P.ray_lo = L.ray_lo;
  P.ray_hi = L.ray_hi;
  P.prims = L.prims;
  P.prim_shift = L.prim_shift;
  P.nodes = L.nodes;
  P.nodes4 = L.nodes4;
  P.nb_col = L.nb_col;
  P.acc = L.acc;
  P.adj = L.adj;
  P.mom = L.mom;
  P.queue = L.queue;
  P.stats = L.stats;
  P.compact = L.compact > 0 ? (L.compact > 64 ? 64 : L.compact) : default_compact_lanes();
  P.leaf_min = L.leaf_batch > 256 ? 256 : L.leaf_batch;
  P.refill_min = L.refill_lanes < 1 ? 1 : (L.refill_lanes > 64 ? 64 : L.refill_lanes);
  const bool persistent = L.kernel == 0;
  const bool postpone = persistent && L.width == 4 && L.leaf_batch > 0;
  const bool count = L.count != 0;
  const int bs = L.block_size;
  if (bs < 64 || bs > kMaxBlock || (bs & 63)) return hipErrorInvalidValue;
  P.block = (uint32_t)bs;
  const kfn f = pick_kernel(L.width, L.lds_stack, persistent, postpone, L.adjoint, L.accum_fp64, count, L.mesh != 0);
  if (!f) return hipErrorInvalidValue;
  const size_t shmem = trace_lds_bytes(L.width, bs, L.lds_stack, postpone);
  uint64_t grid;
  if (persistent) {
    grid = (uint64_t)L.grid_blocks;
  } else {
    grid = (L.ray_hi - L.ray_lo + bs - 1) / bs;
    if (grid == 0) grid = 1;
  }
  hipLaunchKernelGGL(f, dim3((unsigned)grid), dim3(bs), shmem, stream, P);
  return hipGetLastError();
}

hipError_t trace_occupancy(int lds_stack, int block_size, int persistent, int leaf_batch, int adjoint, int fp64,
                           int width, int mesh, int* blocks_per_cu) {
  const bool postpone = persistent && width == 4 && leaf_batch > 0;
  const kfn f = pick_kernel(width, lds_stack, persistent, postpone, adjoint, fp64, false, mesh != 0);
  if (!f) return hipErrorInvalidValue;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, block_size,
                                                      trace_lds_bytes(width, block_size, lds_stack, postpone));
}

}  // namespace rti_impl
