// Device-side data layout and helpers shared by the gfx950 kernels.
//
// Layout in HBM (one "scene store" holds one geometry or a whole batch; all
// arrays are concatenated over the instances and indexed by absolute ids):
//   prims[(p << s) + k]  float4  one record per primitive, so that a
//                  primitive test, a neighbour test and a surface hit each
//                  touch one aligned block. A store of discs only has
//                  32-byte records (s = 1), any other store 64-byte records
//                  (s = 2):
//                    k = 0  discs: centre xyz + radius;  triangles: v0, w = -1
//                    k = 1  discs: unit normal, 0;       triangles: e1 = v1 - v0
//                    k = 2  discs: unused;               triangles: e2 = v2 - v0
//                    k = 3  unit normal (discs as given; triangles from the winding)
//                  The last float4 of a record, k = (1 << s) - 1, is always
//                  the unit normal, with the start of the primitive's
//                  neighbour row (nb_ptr[p], u32 bits) in w. One sentinel
//                  record after the last primitive holds only that w (the
//                  end of the last row), so row p is [w of p, w of p + 1).
//   nodes4[i]      bvh4_node 64 B, four quantised child boxes (the traversal tree)
//   nodes[i]       bvh_node 64 B, both children's boxes in the parent (the build's tree)
//   nb_ptr/nb_col  neighbour CSR over absolute prim ids (discs; the trace
//                  reads the row bounds from the records, nb_ptr serves the
//                  build, the smoothing and the exports)
//   acc[4a + k]    u64  k = 0 sum of deposits (fixed point 2^shift), 1 sum of
//                  squares (fixed point 2^shift2), or both as FP64 bits, 2
//                  deposit count, 3 unused: one 32-byte record, so a
//                  deposit's three atomics share a cache line. a is the
//                  accumulator slot, p + acc_off of the primitive's instance:
//                  instances (particle species) that share one geometry share
//                  its primitive records and trees but not their slots
//   adj[2a + k]    u64  adjoint sums d/ds (k = 0) and d/dn (k = 1), signed fixed point
//   mom[2a + k]    f64  sums of deposit^3 (k = 0) and deposit^4 (k = 1)
// Primitives are stored in Morton order of their instance (disc clouds) or
// in the depth-first order of its four-wide tree over Morton order (each
// node's leaves contiguous; triangle meshes), so neighbouring primitives in
// space are neighbours in memory (local leaf and accumulator traffic);
// `perm` maps a stored position back to the input id.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rti_dev {

constexpr uint32_t kPrimShiftDiscs = 1;  // log2 float4 per record, store of discs only
constexpr uint32_t kPrimShiftMixed = 2;  // log2 float4 per record, any triangle in the store
constexpr int kAccStride = 4;   // u64 per accumulator record

// Internal node of the binary LBVH. Child c's box is
//   x: (lo_hi_xy[c].x, lo_hi_xy[c].y), y: (lo_hi_xy[c].z, lo_hi_xy[c].w),
//   z: (lo_hi_z.{x,y} for child 0, lo_hi_z.{z,w} for child 1).
// child[c] >= 0 is an internal node index; child[c] < 0 is leaf ~prim.
struct __align__(16) bvh_node {
  float4 xy0;     // child 0: lo.x, hi.x, lo.y, hi.y
  float4 xy1;     // child 1: lo.x, hi.x, lo.y, hi.y
  float4 z01;     // child 0 lo.z, hi.z, child 1 lo.z, hi.z
  int4 child;     // child0, child1, unused, unused
};
static_assert(sizeof(bvh_node) == 64, "node must be one 64-byte record");

// Four-wide node, collapsed from the binary tree, with its children's boxes
// quantised to 8 bits per plane in the frame of the node's own box:
//   plane = fmaf(q, 2^(e_axis - 127), org_axis),
// where q is byte c (child c) of the plane's word and e_axis the axis's
// biased exponent. The build rounds lower planes down and upper planes up
// through exactly this expression, so every quantised box contains the
// child's exact box. Children are ordered internal first: slots 0..nint-1
// are the four-wide nodes int_base + c (siblings are contiguous), slots
// nint..nchild-1 are leaves, primitive leaf[c] (stored position); slots
// nchild..3 hold an inverted box (low bytes 255, high 0) and no child.
// 64 bytes; a visit loads the first three 16-byte words (boxes, internal
// children and leaf_info). The primitives of a triangle mesh are stored in
// its tree's depth-first order, each node's leaves contiguous
// (kNodeLeavesContiguous in meta), so leaf c is leaf_info + c - nint and the
// fourth word is never loaded; in the tree of a disc cloud (Morton order)
// leaf_info marks the internal children that have leaf children, and the
// visit of a marked node loads the fourth word with the other three.
struct __align__(16) bvh4_node {
  float ox, oy, oz;
  uint32_t meta;       // e_x | e_y << 8 | e_z << 16 | (nchild - 1) << 24 | nint << 26 | kNodeLeavesContiguous
  uint32_t lo_x, lo_y, lo_z, hi_x;
  uint32_t hi_y, hi_z;
  uint32_t int_base;   // first internal child (absolute four-wide index)
  uint32_t leaf_info;  // triangles: first leaf (absolute stored position); discs: bit c = internal
                       // child c has leaf children
  uint32_t leaf[4];    // leaf slots: absolute stored position of the primitive; other slots 0
};
constexpr uint32_t kNodeLeavesContiguous = 1u << 29;
// A reference to a four-wide node in the traversal of a tree of discs
// carries, in this bit, its parent's mark that the node has leaf children
// (node indices stay below it).
constexpr uint32_t kNodeHasLeaves = 1u << 30;

__host__ __device__ inline int bvh4_nchild(const bvh4_node& nd) { return 1 + (int)((nd.meta >> 24) & 3u); }
__host__ __device__ inline int bvh4_nint(const bvh4_node& nd) { return (int)((nd.meta >> 26) & 7u); }
static_assert(sizeof(bvh4_node) == 64, "four-wide node must be one 64-byte record");

// Per-instance description, read by each lane for the instance its ray
// belongs to. 128 bytes.
struct __align__(16) inst_desc {
  float xmin, xmax, ymin, ymax;     // boundary faces
  float zmin, zmax, src_z, tnear;   // boundary z extent, source plane, ray epsilon
  float sticking, cos_exponent, diffuse_prob, pow_e;  // pow_e = 2 / (n + 1)
  int32_t bc_x, bc_y, prim_kind, root;                // root: absolute node index
  uint64_t ray_begin;               // first id of the instance's slice of the global ray queue
  uint64_t ray_id0;                 // Philox ray id of the slice's first ray (the run's first ray id)
  uint32_t seed, key1, roulette;
  int32_t acc_off;                  // accumulator slot of primitive p: p + acc_off (species
                                    // sharing one geometry have their own slots)
  uint32_t n_prims;
  float fx_scale2;                  // fixed-point scale of the sum of squares 2^shift2
  uint32_t pad1, pad2;
  float score_c, fx_scale;          // score_c = 1 / (n + 1); fixed-point scale 2^shift
  int32_t root4, source;            // root4: absolute four-wide node index; source: rti_source
};
static_assert(sizeof(inst_desc) == 128, "descriptor size");

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 make(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 normalize(f3 a) {
  float l = sqrtf(dot(a, a));
  return {a.x / l, a.y / l, a.z / l};
}

// Philox4x32-10 (Salmon et al., SC'11): 10 rounds of the published
// multipliers, Weyl key schedule.
struct u4 {
  uint32_t x, y, z, w;
};
__device__ __forceinline__ u4 philox(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one full 32 x 32 -> 64-bit product per multiplier (v_mad_u64_u32)
    // instead of separate low and high multiplies
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    c = {hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}
// 24-bit uniforms, exactly representable in float.
__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }
__device__ __forceinline__ float u01_open0(uint32_t x) { return (float)((x >> 8) + 1u) * 5.9604644775390625e-08f; }

}  // namespace rti_dev
