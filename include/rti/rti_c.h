/*
 * rti_c.h -- C ABI of the MI355X (gfx950) Monte Carlo flux tracer.
 *
 * This is the boundary the C++ drop-in `rti::device` (include/rti/device.hpp)
 * and any other host binding (ctypes, cgo, JNI) call. Plain pointers and
 * sizes only; the library copies inputs in and results out, owns all device
 * memory, and every call is synchronous on return.
 *
 * Which reference interface each entry point replaces (reference paths are
 * relative to the rti source tree):
 *   rti_ctx_create / rti_ctx_destroy  -> rti::device::device() and the per-run
 *                                        rtcNewDevice/rtcReleaseDevice
 *                                        (src/rti/device.hpp:42-45, 84-85, 101)
 *   rti_set_discs                     -> device::set_points + set_normals +
 *                                        set_grid_spacing, packed as
 *                                        (x,y,z,r) by
 *                                        combine_points_with_grid_spacing...
 *                                        (device.hpp:47-60, 206-222)
 *   rti_set_triangles                 -> geo::triangle_geometry buffers
 *                                        (src/rti/geo/triangle_geometry.hpp:56-99)
 *   rti_set_params                    -> device::set_number_of_rays, set_x,
 *                                        set_y, set(i_direction&) and the
 *                                        particle's sticking probability
 *                                        (device.hpp:62-80;
 *                                        src/rti/particle/i_particle.hpp:8-16)
 *   rti_run                           -> device::run (device.hpp:82-107) =
 *                                        geometry + boundary + source + tracer
 *                                        + post-processing
 *   rti_get_flux                      -> device::get_mc_estimates
 *                                        (device.hpp:109-112, 153-187) and the
 *                                        raw hit_accumulator::get_values
 *                                        (src/rti/trace/hit_accumulator.hpp:140-142)
 *   rti_get_counts                    -> device::get_hit_cnts (device.hpp:114-117)
 *   rti_get_areas                     -> hit_accumulator::get_exposed_areas
 *                                        (tracer.hpp:281-282, 385-413)
 *   rti_get_neighbors                 -> point_cloud_disc_geometry::get_neighbors
 *                                        (point_cloud_disc_geometry.hpp:391-394)
 *   rti_get_stats                     -> trace::result (src/rti/trace/result.hpp:11-33)
 *   rti_get_adjoint                   -> (no reference counterpart: d flux / ds
 *                                        and d flux / dn of the same run)
 *   rti_batch_*                       -> (no reference counterpart: many
 *                                        geometries in one persistent launch)
 *   rti_get_counters                  -> the tracer's hit / non-hit counters
 *                                        (tracer.hpp:106-107, 186, 229)
 *
 * Error convention: every function returns RTI_OK (0) or an RTI_ERR_* code;
 * rti_last_error(ctx) returns a message for the last failure on that ctx.
 * One ctx per host thread; calls on one ctx must not overlap.
 */
#ifndef RTI_C_H
#define RTI_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rti_ctx rti_ctx;

enum rti_status {
  RTI_OK = 0,
  RTI_ERR_INVALID_ARG = 1,   /* null pointer, size mismatch, value out of range */
  RTI_ERR_NO_GEOMETRY = 2,   /* run before set_discs / set_triangles, or N = 0  */
  RTI_ERR_HIP = 3,           /* a HIP runtime call failed                        */
  RTI_ERR_OVERFLOW = 4,      /* fixed-point accumulator could have overflowed    */
  RTI_ERR_STATE = 5,         /* getter called before a successful run            */
  RTI_ERR_ALLOC = 6          /* device allocation failed                         */
};

enum rti_bound_condition { RTI_BC_REFLECTIVE = 0, RTI_BC_PERIODIC = 1 };
enum rti_kernel {
  RTI_KERNEL_PERSISTENT = 0,  /* persistent, wavefront-scheduled (default)              */
  RTI_KERNEL_PLAIN = 1,       /* one thread per ray, one launch per geometry            */
  RTI_KERNEL_COUNT = 2        /* persistent with node / primitive / neighbour counters  */
};
enum rti_accum { RTI_ACCUM_FIXED64 = 0, RTI_ACCUM_FP64 = 1 };
/* Where primary rays start on the source plane. RECTANGLE: uniform over the
 * boundary box (rectangle_origin_z, src/rti/ray/rectangle_origin_z.hpp:23-37,
 * as placed by device::create_rectangular_source_from_bounding_box,
 * src/rti/device.hpp:257-267; what device::run uses). DISC: uniform over the
 * disc centred on the box with radius half its x extent (disc_origin_z,
 * src/rti/ray/disc_origin_z.hpp:17-33, as placed by
 * device::create_circular_source_from_bounding_box, device.hpp:269-280). */
enum rti_source { RTI_SOURCE_RECTANGLE = 0, RTI_SOURCE_DISC = 1 };

/* Topology of the binary tree the traversal trees are built from (both over
 * the primitives' Morton order, with a bottom-up box refit; replaces the
 * Embree BVH build of tracer.hpp:74-100). LBVH: the Karras radix tree of the
 * Morton codes. PLOC: locally-ordered clustering (mutual nearest neighbours
 * by merged surface area within 16 positions), fewer node visits per ray. */
enum rti_tree_builder { RTI_TREE_LBVH = 1, RTI_TREE_PLOC = 2 };

typedef struct rti_params {
  uint64_t n_rays;        /* primary rays (reference default 1024, device.hpp:289)    */
  uint32_t seed;          /* Philox key word 0                                        */
  uint32_t instance;      /* Philox key word 1 (batch: added to the instance index)   */
  int32_t bc_x, bc_y;     /* rti_bound_condition (reference default reflective)       */
  float sticking;         /* sticking probability s in (0, 1]                          */
  float cos_exponent;     /* source direction pdf ~ cos^n (1 = cosine, the default)    */
  float diffuse_prob;     /* probability of a diffuse bounce; 1 = diffuse, 0 = mirror  */
  float tnear;            /* ray epsilon, reference 1e-4 (tracer.hpp:179)              */
  int32_t roulette;       /* 1 = rejection control 0.1/0.3 (rejection_control.hpp)    */
  float src_z;            /* source plane height; NaN = derived from the geometry      */
  int32_t kernel;         /* rti_kernel                                                 */
  int32_t adjoint;        /* 1 = also accumulate d/ds and d/dn                          */
  int32_t accum;          /* rti_accum                                                  */
  int32_t block_size;     /* 0 = tuned default; else 64, 128, 192 or 256                */
  int32_t lds_stack;      /* 0 = tuned default; else 4, 6, 8, 16, 24 or 32 LDS stack slots */
  int32_t waves_per_cu;   /* persistent kernel: resident waves per CU, 0 = default     */
  int32_t compact_lanes;  /* persistent kernel: lanes of a wave whose closest-hit query
                             must be finished before the wave shades and refills
                             (1..64; 0 = tuned default)                                */
  int32_t bvh_width;      /* traversal tree: 0 = default (4), 4 = four-wide quantised,
                             2 = binary LBVH                                           */
  int32_t moments;        /* 1 = also accumulate sums of deposit^3 and deposit^4 (FP64)
                             for the variance of the variance                          */
  int32_t source;         /* rti_source (default RTI_SOURCE_RECTANGLE)                  */
  int32_t leaf_batch;     /* persistent kernel, four-wide tree: hit leaves go on a list
                             shared by the wave's lanes and are tested in leaf phases, up
                             to 64 at a time, one per lane; a leaf phase runs once the
                             entries it would take reach this many 64ths of the lanes
                             with a node left (1..256), or no node is left;
                             0 = default (128 when the store holds triangles, else -1);
                             -1 = test leaves at each node visit                       */
  int32_t tree_builder;   /* rti_tree_builder; 0 = default (PLOC for triangle meshes,
                             LBVH for disc clouds)                                      */
  int32_t refill_lanes;   /* persistent kernel: a wave replaces ended rays once this many
                             of its lanes wait for one, or when none is alive (1..64);
                             0 = default (1, at every shading pass)                     */
  uint64_t ray_begin;     /* id of the first ray: the run traces rays [ray_begin,
                             ray_begin + n_rays) of the sequence a larger run with the
                             same seed and instance would trace (a shard of a job)     */
  uint64_t rays_total;    /* rays of the whole job this run is a shard of; sets the
                             fixed-point scale, so that the integer sums of shards
                             [0, a) + [a, b) + ... add up to the single run's bit for
                             bit; 0 = n_rays; otherwise at least n_rays               */
} rti_params;

typedef struct rti_stats {
  uint64_t n_prims;
  uint64_t n_rays;
  uint64_t segments;      /* closest-hit queries (primary, reflected, boundary, back face, miss) */
  uint64_t surface_hits;  /* front-face primitive hits                                      */
  uint64_t misses;        /* rays that left the domain                                       */
  uint64_t truncated;     /* rays stopped by the per-ray segment cap                         */
  uint64_t n_edges;       /* directed neighbour pairs                                        */
  uint32_t bvh_depth;     /* deepest leaf of the hierarchy                                   */
  int32_t fixed_shift;    /* fixed-point scale is 2^fixed_shift                              */
  double trace_ms;        /* trace kernel(s) time (HIP events)                               */
  double build_ms;        /* upload + LBVH + neighbours + areas                              */
  double post_ms;         /* finalisation + download                                         */
  float box[6];           /* boundary xmin, xmax, ymin, ymax, zmin, zmax                     */
  float src_z;
  uint32_t wide_stack;    /* most stack entries a four-wide traversal can hold               */
  uint64_t wide_nodes;    /* four-wide traversal-tree nodes                                  */
  int32_t fixed_shift2;   /* fixed-point scale of the sum of squared deposits is 2^fixed_shift2 */
  int32_t pad0;
  uint64_t max_segments;  /* segments of the longest ray path of the last trace (the paths
                             that set how long a launch's last waves run)               */
} rti_stats;

/* One geometry of a batch. Exactly one of (xyzr, nrm) or (verts, tris) is set.
 * Instances whose geometry buffers are the very same pointers and sizes as an
 * earlier instance's (particle species on one geometry) share that instance's
 * primitive records, trees and neighbour graph in device memory; each keeps
 * its own parameters and accumulators, and results are the same as with
 * separate copies. */
typedef struct rti_instance {
  const float* xyzr;      /* n_prims x 4 */
  const float* nrm;       /* n_prims x 3, unit length */
  const float* verts;     /* n_verts x 3 */
  uint64_t n_verts;
  const uint32_t* tris;   /* n_prims x 3 */
  uint64_t n_prims;
  rti_params params;      /* per-instance rays, sticking, reflection, source, bcs; kernel knobs from rti_set_params */
} rti_instance;

void rti_default_params(rti_params* p);

/* No reference counterpart: the device's architecture name (e.g. "gfx950")
 * and compute-unit count, for tools that key measurements by device (the
 * autotuner's cache). name may be NULL; it receives at most cap bytes,
 * NUL-terminated. */
int rti_device_info(int device, char* name, uint64_t cap, int* num_cus);

int rti_ctx_create(rti_ctx** out, int device);
int rti_ctx_destroy(rti_ctx* ctx);
const char* rti_last_error(const rti_ctx* ctx);

int rti_set_discs(rti_ctx* ctx, const float* xyzr, const float* nrm, uint64_t n);
int rti_set_triangles(rti_ctx* ctx, const float* verts, uint64_t n_verts, const uint32_t* tris, uint64_t n_tris);
int rti_set_params(rti_ctx* ctx, const rti_params* p);

/* Build (if the geometry changed), trace, finalise; results stay on the ctx. */
int rti_run(rti_ctx* ctx);
/* Split form for timing with inputs resident: rti_prepare builds the
 * acceleration structures; rti_trace clears the accumulators and traces on
 * the device without downloading; rti_finalize post-processes and downloads. */
int rti_prepare(rti_ctx* ctx);
int rti_trace(rti_ctx* ctx);
int rti_finalize(rti_ctx* ctx);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the trace of rti_trace on a context already built by
 * rti_prepare, calling no hipMalloc / hipFree. The accumulator records go to
 * the caller's device buffer `acc` of acc_words >= 4 x (primitives of all
 * instances) 64-bit words, cleared by the call: per primitive, in stored
 * order (rti_get_hierarchy's perm maps a stored position to its input id),
 * {sum of deposits, sum of squared deposits, deposit count, 0}, the sums in
 * fixed point with scale 2^fixed_shift (rti_get_stats). The instance
 * descriptors, ray queue and counters live in a static per-device array of
 * 33,024 bytes, which holds at most 256 instances (more: RTI_ERR_INVALID_ARG);
 * every call on one GPU shares it, so calls that overlap on one GPU race.
 * params.adjoint and params.moments are rejected (RTI_ERR_INVALID_ARG); an
 * unbuilt context returns RTI_ERR_STATE. rti_get_counters reports the trace;
 * the context's own results and rti_finalize need a new rti_trace/rti_run. */
int rti_trace_noalloc(rti_ctx* ctx, unsigned long long* acc, uint64_t acc_words);

/* Any output pointer may be NULL. smoothed_norm: reference mc estimates
 * (neighbourhood mean of value/area, divided by the max); raw: sum of
 * deposits per primitive (reference `value`); sigma: Monte Carlo standard
 * deviation of raw (per-deposit second moment). Arrays hold n_prims entries
 * in input order. */
int rti_get_flux(rti_ctx* ctx, float* smoothed_norm, double* raw, double* sigma);
int rti_get_counts(rti_ctx* ctx, uint64_t* counts);
int rti_get_areas(rti_ctx* ctx, float* areas);
/* row_ptr: n_prims + 1 entries; col: up to cap entries (input order ids). */
int rti_get_neighbors(rti_ctx* ctx, uint32_t* row_ptr, uint32_t* col, uint64_t cap, uint64_t* n_edges);
int rti_get_stats(rti_ctx* ctx, rti_stats* out);
/* Per-primitive statistics of the last run, input order, as the reference's
 * VTP output has them (hit_accumulator::get_relative_error / get_vov,
 * src/rti/trace/hit_accumulator.hpp:152-249, written by io/vtp_writer.hpp:207-237):
 * rel_err = sqrt(S2/S1^2 - 1/N), vov = (S4 - 4 S1 S3/N + 8 S2 S1^2/N^2 - 4 S1^4/N^3
 * - S2^2/N) N^2 / (N S2 - S1^2)^2, with S_k the sums of deposit^k on the
 * primitive and N all deposits of the run; DBL_MAX where undefined (S1 = 0;
 * no deposit or zero denominator). vov needs params.moments = 1. Either
 * output may be NULL. */
int rti_get_error_stats(rti_ctx* ctx, double* rel_err, double* vov);

/* d(raw)/ds and d(raw)/dn of the last run made with params.adjoint = 1. */
int rti_get_adjoint(rti_ctx* ctx, double* draw_ds, double* draw_dn);
/* Counters of the last trace: [0] segments, [1] surface hits, [2] misses,
 * [3] truncated rays; after an RTI_KERNEL_COUNT run also [4] BVH node
 * visits, [5] primitive tests, [6] neighbour tests, [7] deposits. */
int rti_get_counters(rti_ctx* ctx, uint64_t out[8]);
/* The built hierarchy of the (first) geometry, for validation and tools:
 * n_nodes 64-byte nodes {float4 child0 x/y box, float4 child1 x/y box,
 * float4 z boxes, int4 children (>= 0 node, < 0 leaf ~stored position)},
 * perm[stored position] = input id, prim_a (n_prims x float4, stored
 * order: disc centre + radius, or triangle v0 with w = -1). Node indices in
 * the children are relative to this geometry. Any output may be NULL. */
int rti_get_hierarchy(rti_ctx* ctx, void* nodes, uint64_t cap_nodes, uint32_t* perm, uint64_t* n_nodes,
                      float* prim_a);

/* The four-wide traversal tree of the (single) geometry: 64-byte records
 * { float org[3]; uint32 meta, lo_x, lo_y, lo_z, hi_x, hi_y, hi_z,
 *   int_base, leaf_info, leaf[4] },
 * meta = e_x | e_y << 8 | e_z << 16 | (nchild - 1) << 24 | nint << 26
 *        (| 1 << 29 in a triangle mesh's tree: leaves contiguous);
 * child c's box planes fmaf(byte c of the plane word, 2^(e - 127), org).
 * Children are ordered internal first: slot c < nint is node int_base + c
 * (node index relative to this geometry), slot nint <= c < nchild is the
 * primitive at stored position leaf[c]; slots beyond nchild hold an inverted
 * box. In a geometry of triangles the stored positions follow the tree
 * depth first, each node's leaves contiguous from leaf_info; for discs
 * (Morton order) bit c of leaf_info marks an internal child with leaf
 * children. Root is record 0. Each family of siblings starts on an even record
 * (a 128-byte line); records no node references are padding (an inverted
 * box in all four slots). The reference's Embree BVH is internal to Embree
 * and has no such export; this one serves validation. */
int rti_get_wide_hierarchy(rti_ctx* ctx, void* nodes, uint64_t cap_nodes, uint64_t* n_nodes);

/* Batch: many geometries in one persistent launch (captured in a HIP graph
 * and replayed by rti_batch_replay). Results per instance by index. */
int rti_batch_set(rti_ctx* ctx, const rti_instance* inst, uint32_t n);
int rti_batch_run(rti_ctx* ctx);
int rti_batch_replay(rti_ctx* ctx, uint32_t times);
int rti_batch_get_flux(rti_ctx* ctx, uint32_t i, float* smoothed_norm, double* raw, double* sigma);
int rti_batch_get_counts(rti_ctx* ctx, uint32_t i, uint64_t* counts);
int rti_batch_get_stats(rti_ctx* ctx, rti_stats* out);
int rti_batch_get_error_stats(rti_ctx* ctx, uint32_t i, double* rel_err, double* vov);
/* d(raw)/ds and d(raw)/dn of instance i after a batch run with params.adjoint = 1. */
int rti_batch_get_adjoint(rti_ctx* ctx, uint32_t i, double* draw_ds, double* draw_dn);

#ifdef __cplusplus
}
#endif

#endif /* RTI_C_H */
