// Host-side declarations shared by the translation units of librti_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "rti_device.hpp"

namespace rti_impl {

// Device buffers of one instance's build inputs (input order) and where its
// outputs go inside the scene store.
struct build_job {
  int prim_kind;            // 0 discs, 1 triangles
  uint32_t n;               // primitives
  uint32_t prim_begin;      // first absolute prim slot
  uint32_t node_begin;      // first absolute node slot (max(n - 1, 1) nodes)
  uint32_t node4_begin;     // first absolute four-wide slot (wide_capacity(n) slots, even)
  const float* d_xyzr;      // discs, input order, n x 4
  const float* d_nrm;       // discs, n x 3
  const float* d_verts;     // triangles
  const uint32_t* d_tris;   // triangles, n x 3
  float cmin[3], cmax[3];   // centroid bounds (for Morton quantisation)
  float rmax;               // largest disc radius (neighbour query reach)
  float pad;                // absolute padding of every primitive box (coordinate scale of the instance)
  float area_box[4];        // xmin, ymin, xmax, ymax for the exposed area
  int builder;              // binary tree topology: kBuildKarras or kBuildPloc
};
constexpr int kBuildKarras = 1;  // Karras radix tree over the Morton codes
constexpr int kBuildPloc = 2;    // locally-ordered clustering (PLOC) in Morton order

// Scene store (absolute indexing; see rti_device.hpp).
struct store_ptrs {
  float4* prims;            // 1 << prim_shift float4 per primitive (see rti_device.hpp)
  uint32_t prim_shift;      // kPrimShiftDiscs or kPrimShiftMixed
  uint32_t* perm;           // stored position -> input id (per instance)
  rti_dev::bvh_node* nodes;
  rti_dev::bvh4_node* nodes4;  // four-wide traversal tree, same slot ranges as nodes
  uint32_t* nb_ptr;         // n_total + 1
  uint32_t* nb_col;
  float* areas;             // per stored position
};

// Scratch for builds (sized for the largest instance).
struct build_scratch {
  uint32_t cap = 0;
  uint32_t* keys[2] = {nullptr, nullptr};
  uint32_t* vals[2] = {nullptr, nullptr};
  uint32_t* hist = nullptr;       // 256 x blocks
  uint32_t* scan_tmp = nullptr;
  int32_t* parent = nullptr;      // 2n - 1: internal nodes then leaves (local ids)
  uint32_t* flags = nullptr;      // n - 1
  float4* leaf_lo = nullptr;      // n
  float4* leaf_hi = nullptr;      // n
  uint32_t* counts = nullptr;     // n + 1 (neighbour degrees; PLOC survivor flags)
  uint32_t* depth = nullptr;      // 1
  float4* clo[2] = {nullptr, nullptr};  // n: PLOC cluster boxes (ping-pong)
  float4* chi[2] = {nullptr, nullptr};
  uint32_t* lead = nullptr;       // n + 1: PLOC merge flags
};

// Four-wide slots an instance of n primitives may use: its nodes plus one
// padding slot per family (each family starts on a 128-byte line) and the
// root's padding slot.
inline uint32_t wide_capacity(uint32_t n) { return 2u * (n > 2u ? n - 1u : 1u) + 2u; }

hipError_t scratch_reserve(build_scratch& s, uint32_t n);
void scratch_free(build_scratch& s);

// Builds prims (Morton order), perm, BVH nodes and areas for one instance.
// Neighbour CSR is built separately because its size is data dependent.
hipError_t build_instance(const build_job& job, const store_ptrs& st, build_scratch& s, uint32_t* depth_out,
                          hipStream_t stream);
// Four-wide traversal tree collapsed from the instance's binary tree, into
// nodes4[node4_begin ..). n_wide: slots written; stack_bound: the most stack
// entries a closest-hit traversal of it can hold.
hipError_t build_wide(const build_job& job, const store_ptrs& st, build_scratch& s, uint32_t* n_wide,
                      uint32_t* stack_bound, hipStream_t stream);
// Degrees of the overlap graph (discs), into s.counts[0..n).
hipError_t neighbor_count(const build_job& job, const store_ptrs& st, build_scratch& s, hipStream_t stream);
// Columns, with st.nb_ptr already holding absolute row offsets.
hipError_t neighbor_fill(const build_job& job, const store_ptrs& st, hipStream_t stream);
// Exclusive scan of n uint32 values (out may alias in). tmp: scan_tmp_size(n) words.
hipError_t exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, hipStream_t stream);
size_t scan_tmp_size(uint32_t n);

// ---- trace ----
// Run counters of a trace: segments, surface hits, misses, truncated rays,
// (instrumented runs) node visits, primitive tests, neighbour tests,
// deposits, then the segments of the longest ray path.
constexpr int kStatsWords = 9;

struct trace_launch {
  const rti_dev::inst_desc* desc;   // n_inst descriptors
  uint32_t n_inst;
  uint64_t ray_lo, ray_hi;          // global ray ids [ray_lo, ray_hi) traced by this launch
  const float4* prims;
  uint32_t prim_shift;              // log2 float4 per primitive record
  const rti_dev::bvh_node* nodes;
  const rti_dev::bvh4_node* nodes4;
  const uint32_t* nb_col;           // neighbour columns; row bounds are in the prims records
  unsigned long long* acc;          // kAccStride per accumulator slot: sum, sum of squares (fixed point or double
                                    // bits), count
  unsigned long long* adj;          // 2 per slot: d/ds, d/dn (signed fixed point); may be null
  double* mom;                      // 2 per slot: sums of deposit^3, ^4 (FP64 atomics); may be null
  unsigned long long* queue;        // [0]: next ray id
  unsigned long long* stats;        // kStatsWords run counters
  double fx_scale;                  // 2^shift
  int adjoint;
  int accum_fp64;
  int kernel;                       // 0 persistent, 1 plain
  int block_size;
  int lds_stack;
  int grid_blocks;                  // persistent grid
  int count;                        // 1 = instrumented run (node / prim / neighbour counters)
  int compact;                      // persistent: lanes done before a wave shades (1..64, 0 = default)
  int width;                        // traversal tree: 2 binary, 4 four-wide
  int leaf_batch;                   // persistent four-wide: postponed leaf tests on the wave's leaf list, a leaf
                                    // phase once the entries it takes reach this many 64ths of the lanes with
                                    // a node left (1..256); 0 = at each node
  int refill_lanes;                 // persistent: lanes waiting for a new ray before a wave refills (1..64)
  int mesh;                         // every instance is a triangle mesh (four-wide nodes without a leaf word)
};
hipError_t launch_trace(const trace_launch& L, hipStream_t stream);
hipError_t trace_occupancy(int lds_stack, int block_size, int persistent, int leaf_batch, int adjoint, int fp64,
                           int width, int mesh, int* blocks_per_cu);
int default_lds_stack(int width, bool postpone);
// Stack entries (LDS + private) a four-wide traversal holds at most.
int wide_stack_capacity();
int default_compact_lanes();
int default_leaf_batch(bool triangles);  // 0 = leaf tests at each node
// Postponed leaf tests list primitive ids below this bound.
uint64_t postponed_leaf_prim_limit();

// ---- post ----
hipError_t fill_u64(unsigned long long* p, uint64_t n, unsigned long long v, hipStream_t stream);
// Converts one instance's accumulators (slots acc_begin ..; primitives at
// stored positions prim_begin ..): raw_stored[j] (local stored order) and, in
// input order via perm, raw_out / sigma_out / cnt_out (each may be null).
// inv_scale / inv_scale2: inverse fixed-point scales of the sum and the sum
// of squares.
hipError_t finalize_instance(const unsigned long long* acc, uint32_t acc_begin, const uint32_t* perm,
                             uint32_t prim_begin, uint32_t n, double inv_scale, double inv_scale2, int accum_fp64,
                             uint64_t n_rays, double* raw_stored, double* raw_out, double* sigma_out,
                             unsigned long long* cnt_out, hipStream_t stream);
// Relative error (rel_out) and, when s3 / s4 are given, the variance of the
// variance (vov_out) per primitive in input order; total = all deposits of
// the instance. Either output may be null.
hipError_t error_stats(const unsigned long long* acc, const double* mom, uint32_t acc_begin, const uint32_t* perm,
                       uint32_t prim_begin, uint32_t n, double inv_scale, double inv_scale2, int accum_fp64,
                       uint64_t total, double* rel_out, double* vov_out, hipStream_t stream);
// Signed fixed-point (adjoint) accumulator component k of adj[2a + k] to
// double, input order.
hipError_t finalize_signed(const unsigned long long* adj, int k, uint32_t acc_begin, const uint32_t* perm,
                           uint32_t prim_begin, uint32_t n, double inv_scale, double* out, hipStream_t stream);
// Smoothed, max-normalised estimate (reference device.hpp:153-187), written
// in input order. raw_stored / areas are in local stored order; the CSR uses
// absolute ids. vtmp: n doubles; maxbuf: 1 double.
hipError_t smooth_normalize(const double* raw_stored, const float* areas_stored, const uint32_t* perm,
                            const uint32_t* nb_ptr, const uint32_t* nb_col, uint32_t prim_begin, uint32_t n,
                            double* vtmp, double* maxbuf, float* out, hipStream_t stream);
// Permute a stored-order float array into input order.
hipError_t permute_f32(const float* stored, const uint32_t* perm, uint32_t n, float* out, hipStream_t stream);

}  // namespace rti_impl
