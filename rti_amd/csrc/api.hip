// C ABI of librti_hip.so (declared in include/rti/rti_c.h).
//
// Host orchestration of one run, following rti::device::run (reference
// src/rti/device.hpp:82-107): pack the geometry, derive the bounding box,
// source plane and boundary (device.hpp:88-93, 224-267;
// src/rti/geo/point_cloud_disc_geometry.hpp:83-87), build the acceleration
// structure and the disc neighbourhood, trace, then post-process.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rti/rti_c.h"
#include "internal.hpp"

using rti_dev::bvh_node;
using rti_dev::bvh4_node;
using rti_dev::inst_desc;

namespace {

// Owning device allocation.
struct dbuf {
  void* p = nullptr;
  size_t bytes = 0;
  dbuf() = default;
  dbuf(const dbuf&) = delete;
  dbuf& operator=(const dbuf&) = delete;
  ~dbuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    release();
    size_t want = b ? b : 16;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    bytes = want;
    return hipSuccess;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct instance {
  int kind = 0;  // 0 discs, 1 triangles
  std::vector<float> xyzr, nrm, verts;
  std::vector<uint32_t> tris;
  uint64_t n = 0;
  rti_params params{};
  // geometry bounds (set with the geometry)
  float geo_lo[3] = {0, 0, 0}, geo_hi[3] = {0, 0, 0};
  float ztop = 0;
  int share = -1;  // >= 0: this instance uses the geometry of instance `share` (same buffers)
  // derived at prepare
  uint32_t prim_begin = 0, node_begin = 0, node4_begin = 0;
  uint32_t acc_begin = 0;  // first accumulator slot
  uint64_t ray_begin = 0;
  float box[6] = {0, 0, 0, 0, 0, 0};
  float src_z = 0;
  float rmax = 0;
  uint32_t depth = 0;
  uint32_t n_wide = 0, wide_stack = 0;
  uint64_t n_edges = 0;
  int shift = 32, shift2 = 32;  // fixed-point scales of the sum and of the sum of squares
  // results (input order)
  std::vector<double> raw, sigma, dds, ddn, rel, vov;
  std::vector<float> smoothed, areas;
  std::vector<uint64_t> counts;
  uint64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

}  // namespace

struct rti_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  rti_params params{};
  std::vector<instance> inst;
  bool dirty = true;        // geometry changed since the last prepare
  bool batch_mode = false;  // instances came from rti_batch_set (own physics parameters)
  bool have_results = false;
  bool acc_traced = false;  // the context's accumulators hold a trace of the current store
  // scene store
  dbuf prims, perm, nodes, nodes4, nb_ptr, nb_col, areas, desc;
  uint32_t prim_shift = rti_dev::kPrimShiftMixed;  // log2 float4 per primitive record
  dbuf acc, adj, mom, queue, stats;  // accumulator records (rti_device.hpp)
  // finalisation scratch
  dbuf raw_st, vtmp, maxbuf, out_raw, out_sigma, out_cnt, out_smooth, out_f, out_adj, out_err;
  rti_impl::build_scratch scratch;
  uint64_t n_total = 0, nodes_total = 0, edges_total = 0, rays_total = 0;
  uint64_t acc_total = 0;  // accumulator slots: the primitives of every instance
  double build_ms = 0, trace_ms = 0, post_ms = 0;
  int last_adjoint = 0, last_fp64 = 0, last_count = 0, last_moments = 0;
  uint64_t run_stats[rti_impl::kStatsWords] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::string graph_key;
  std::vector<inst_desc> desc_host;  // source of the descriptor upload; lives as long as the context
};

namespace {

// ---- static scratch of rti_trace_noalloc ----
// The no-allocation trace keeps its instance descriptors, ray queue and
// counters in this file-scope device array: one copy per GPU, held until the
// process ends, shared by every such call on that GPU (calls that overlap on
// one GPU race on it). Layout: [0, 8) ray queue, [64, 136) the nine
// counters, [256, ...) up to kNoAllocMaxInstances descriptors. In host code
// the name `scratch` is a host-side stand-in, not a device address; host code
// uses only the per-device address resolved by resolve_scratch.
constexpr uint32_t kNoAllocMaxInstances = 256;
constexpr size_t kScratchStatsOffset = 64;
constexpr size_t kScratchDescOffset = 256;
constexpr size_t kScratchBytes = kScratchDescOffset + kNoAllocMaxInstances * sizeof(inst_desc);  // 33,024 B
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static_assert(kScratchStatsOffset + 8 * rti_impl::kStatsWords <= kScratchDescOffset, "counters overlap the descriptors");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per device

// Resolves the device address of `scratch` for `device`, once per device;
// the caller has made `device` current.
hipError_t resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = p;
  return hipSuccess;
}

void* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

int fail(rti_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                            \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      return fail((ctx), e_ == hipErrorOutOfMemory ? RTI_ERR_ALLOC : RTI_ERR_HIP, "%s failed: %s (%s:%d)", \
                  #expr, hipGetErrorString(e_), __FILE__, __LINE__);                                  \
  } while (0)

// Deepest binary tree the binary traversal and the neighbour search accept
// (their stacks hold 64 entries). Karras trees over 30-bit Morton codes stay
// below it; PLOC trees are not bounded, and the four-wide traversal checks its
// own exact stack bound instead.
constexpr int kMaxBinaryDepth = 64;

int check_params(rti_ctx* ctx, const rti_params& p) {
  if (!(p.sticking > 0.0f && p.sticking <= 1.0f))
    return fail(ctx, RTI_ERR_INVALID_ARG, "sticking must be in (0, 1], got %g", (double)p.sticking);
  if (!(p.cos_exponent >= 0.0f) || !std::isfinite(p.cos_exponent))
    return fail(ctx, RTI_ERR_INVALID_ARG, "cos_exponent must be finite and >= 0, got %g", (double)p.cos_exponent);
  if (!(p.diffuse_prob >= 0.0f && p.diffuse_prob <= 1.0f))
    return fail(ctx, RTI_ERR_INVALID_ARG, "diffuse_prob must be in [0, 1]");
  if (!(p.tnear >= 0.0f) || !std::isfinite(p.tnear)) return fail(ctx, RTI_ERR_INVALID_ARG, "tnear must be >= 0");
  if (p.bc_x != RTI_BC_REFLECTIVE && p.bc_x != RTI_BC_PERIODIC) return fail(ctx, RTI_ERR_INVALID_ARG, "bad bc_x");
  if (p.bc_y != RTI_BC_REFLECTIVE && p.bc_y != RTI_BC_PERIODIC) return fail(ctx, RTI_ERR_INVALID_ARG, "bad bc_y");
  if (p.kernel != RTI_KERNEL_PERSISTENT && p.kernel != RTI_KERNEL_PLAIN && p.kernel != RTI_KERNEL_COUNT)
    return fail(ctx, RTI_ERR_INVALID_ARG, "bad kernel");
  if (p.accum != RTI_ACCUM_FIXED64 && p.accum != RTI_ACCUM_FP64) return fail(ctx, RTI_ERR_INVALID_ARG, "bad accum");
  if (p.block_size != 0 && (p.block_size < 64 || p.block_size > 256 || (p.block_size & 63)))
    return fail(ctx, RTI_ERR_INVALID_ARG, "block_size must be 64, 128, 192 or 256");
  if (p.lds_stack != 0 && p.lds_stack != 4 && p.lds_stack != 6 && p.lds_stack != 8 && p.lds_stack != 16 &&
      p.lds_stack != 24 && p.lds_stack != 32)
    return fail(ctx, RTI_ERR_INVALID_ARG, "lds_stack must be 4, 6, 8, 16, 24 or 32");
  if (p.waves_per_cu < 0 || p.waves_per_cu > 64) return fail(ctx, RTI_ERR_INVALID_ARG, "waves_per_cu out of range");
  if (p.compact_lanes < 0 || p.compact_lanes > 64) return fail(ctx, RTI_ERR_INVALID_ARG, "compact_lanes out of range");
  if (p.leaf_batch < -1 || p.leaf_batch > 256) return fail(ctx, RTI_ERR_INVALID_ARG, "leaf_batch must be -1 or 0..256");
  if (p.refill_lanes < 0 || p.refill_lanes > 64) return fail(ctx, RTI_ERR_INVALID_ARG, "refill_lanes must be 0..64");
  if (p.tree_builder != 0 && p.tree_builder != RTI_TREE_LBVH && p.tree_builder != RTI_TREE_PLOC)
    return fail(ctx, RTI_ERR_INVALID_ARG, "tree_builder must be 0, RTI_TREE_LBVH or RTI_TREE_PLOC");
  if (p.bvh_width != 0 && p.bvh_width != 2 && p.bvh_width != 4)
    return fail(ctx, RTI_ERR_INVALID_ARG, "bvh_width must be 0, 2 or 4");
  if (p.moments != 0 && p.moments != 1) return fail(ctx, RTI_ERR_INVALID_ARG, "moments must be 0 or 1");
  if (p.source != RTI_SOURCE_RECTANGLE && p.source != RTI_SOURCE_DISC)
    return fail(ctx, RTI_ERR_INVALID_ARG, "source must be RTI_SOURCE_RECTANGLE or RTI_SOURCE_DISC");
  if (p.rays_total != 0 && p.rays_total < p.n_rays)
    return fail(ctx, RTI_ERR_INVALID_ARG, "rays_total (%llu) must be 0 or at least n_rays (%llu)",
                (unsigned long long)p.rays_total, (unsigned long long)p.n_rays);
  if (p.ray_begin + p.n_rays < p.ray_begin) return fail(ctx, RTI_ERR_INVALID_ARG, "ray range wraps around");
  return RTI_OK;
}

int check_discs(rti_ctx* ctx, const float* xyzr, const float* nrm, uint64_t n) {
  if (n == 0) return fail(ctx, RTI_ERR_NO_GEOMETRY, "no primitives");
  if (!xyzr || !nrm) return fail(ctx, RTI_ERR_INVALID_ARG, "null disc buffer");
  if (n >= (1ull << 31)) return fail(ctx, RTI_ERR_INVALID_ARG, "too many primitives");
  for (uint64_t i = 0; i < n; ++i) {
    const float* q = nrm + 3 * i;
    float l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    // the reference asserts normalised normals (point_cloud_disc_geometry.hpp:125-128)
    if (!(std::fabs(l - 1.0f) <= 1e-4f))
      return fail(ctx, RTI_ERR_INVALID_ARG, "normal %llu is not unit length (|n| = %g)", (unsigned long long)i,
                  (double)l);
    const float* p = xyzr + 4 * i;
    if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]) && p[3] > 0.0f && std::isfinite(p[3])))
      return fail(ctx, RTI_ERR_INVALID_ARG, "disc %llu has a non-finite centre or non-positive radius",
                  (unsigned long long)i);
  }
  return RTI_OK;
}

int check_tris(rti_ctx* ctx, const float* verts, uint64_t nv, const uint32_t* tris, uint64_t nt) {
  if (nt == 0) return fail(ctx, RTI_ERR_NO_GEOMETRY, "no primitives");
  if (!verts || !tris) return fail(ctx, RTI_ERR_INVALID_ARG, "null triangle buffer");
  if (nt >= (1ull << 31)) return fail(ctx, RTI_ERR_INVALID_ARG, "too many primitives");
  for (uint64_t i = 0; i < 3 * nt; ++i)
    if (tris[i] >= nv) return fail(ctx, RTI_ERR_INVALID_ARG, "triangle index out of range");
  for (uint64_t i = 0; i < 3 * nv; ++i)
    if (!std::isfinite(verts[i])) return fail(ctx, RTI_ERR_INVALID_ARG, "non-finite vertex");
  return RTI_OK;
}

int bit_length(uint64_t v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

// Geometry bounds of one instance (done once, when the geometry is set).
void geometry_bounds(instance& I) {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  float rmax = 0;
  if (I.kind == 0) {
    for (uint64_t i = 0; i < I.n; ++i) {
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], I.xyzr[4 * i + k]);
        hi[k] = std::max(hi[k], I.xyzr[4 * i + k]);
      }
      rmax = std::max(rmax, I.xyzr[4 * i + 3]);
    }
  } else {
    for (uint64_t i = 0; i < I.verts.size() / 3; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], I.verts[3 * i + k]);
        hi[k] = std::max(hi[k], I.verts[3 * i + k]);
      }
  }
  I.rmax = rmax;
  for (int k = 0; k < 3; ++k) {
    I.geo_lo[k] = lo[k];
    I.geo_hi[k] = hi[k];
  }
  if (I.kind == 0) {
    // bbox z-max padded by the max radius (point_cloud_disc_geometry.hpp:86)
    // and again by the max radius (device.hpp:89-90, 224-237)
    I.ztop = (hi[2] + rmax) + rmax;
  } else {
    I.ztop = hi[2] + 0.01f * std::max(hi[0] - lo[0], hi[1] - lo[1]);
  }
}

// Boundary box and source plane: the box spans the x/y extent of the disc
// centres (boundary_x_y.hpp:194-207 via device.hpp:93); the source sits at the
// padded top (device.hpp:257-267) unless params.src_z overrides it, in which
// case the lateral faces are extended up to the source.
// The boundary faces reach up to the source plane, as in the reference,
// where the source sits on the top of the bounding box (device.hpp:89-93,
// 257-267; main/dev.cpp, which pads the box top once instead of twice).
void derive_box(instance& I) {
  float sz = std::isnan(I.params.src_z) ? I.ztop : I.params.src_z;
  I.src_z = sz;
  I.box[0] = I.geo_lo[0];
  I.box[1] = I.geo_hi[0];
  I.box[2] = I.geo_lo[1];
  I.box[3] = I.geo_hi[1];
  I.box[4] = I.geo_lo[2];
  I.box[5] = sz;
}

void centroid_bounds(const instance& I, float cmin[3], float cmax[3]) {
  for (int k = 0; k < 3; ++k) {
    cmin[k] = FLT_MAX;
    cmax[k] = -FLT_MAX;
  }
  if (I.kind == 0) {
    for (uint64_t i = 0; i < I.n; ++i)
      for (int k = 0; k < 3; ++k) {
        cmin[k] = std::min(cmin[k], I.xyzr[4 * i + k]);
        cmax[k] = std::max(cmax[k], I.xyzr[4 * i + k]);
      }
  } else {
    for (uint64_t t = 0; t < I.n; ++t)
      for (int k = 0; k < 3; ++k) {
        float c = (I.verts[3 * I.tris[3 * t] + k] + I.verts[3 * I.tris[3 * t + 1] + k] +
                   I.verts[3 * I.tris[3 * t + 2] + k]) *
                  (1.0f / 3.0f);
        cmin[k] = std::min(cmin[k], c);
        cmax[k] = std::max(cmax[k], c);
      }
  }
}

__global__ void k_add_u32(uint32_t* p, uint32_t n, uint32_t add) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] += add;
}

__global__ void k_set_u32(uint32_t* p, uint32_t n, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Record p's last float4 carries the start of its neighbour row in w (for
// p = n, the sentinel record after the store: the end of the last row), so
// that the shading of a surface hit finds the row bounds beside the normal it
// loads instead of in a separate array.
__global__ void k_link_rows(float4* prims, uint32_t shift, const uint32_t* nb_ptr, uint32_t n) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p <= n) prims[(((size_t)p + 1) << shift) - 1].w = __uint_as_float(nb_ptr[p]);
}

__global__ void k_set_one_u64(unsigned long long* p, unsigned long long v) { *p = v; }

// Binary tree topology. Default: PLOC for triangle meshes (measured +1.5 %
// on a 199k-triangle hole), the Karras tree for disc clouds (PLOC measured
// -1 % on a batch of disc trenches and holes, -4 % on a 4.2M-disc rough
// surface, whose regular grids Morton order already clusters well).
int tree_builder(const rti_params& p, int kind) {
  if (p.tree_builder == RTI_TREE_LBVH) return rti_impl::kBuildKarras;
  if (p.tree_builder == RTI_TREE_PLOC) return rti_impl::kBuildPloc;
  return kind == 1 ? rti_impl::kBuildPloc : rti_impl::kBuildKarras;
}

int prepare(rti_ctx* ctx) {
  if (ctx->inst.empty()) return fail(ctx, RTI_ERR_NO_GEOMETRY, "no geometry set");
  if (!ctx->dirty) return RTI_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipEvent_t b0, b1;
  HIP_TRY(ctx, hipEventCreate(&b0));
  HIP_TRY(ctx, hipEventCreate(&b1));
  HIP_TRY(ctx, hipEventRecord(b0, ctx->stream));
  // layout
  // an instance that shares another's geometry takes that instance's
  // primitive records, trees and neighbour rows; every instance has its own
  // accumulator slots
  uint64_t np = 0, nn = 0, nn4 = 0, maxn = 0, na = 0;
  for (auto& I : ctx->inst) {
    if (I.share < 0) {
      I.prim_begin = (uint32_t)np;
      I.node_begin = (uint32_t)nn;
      I.node4_begin = (uint32_t)nn4;  // even: each instance's families start on 128-byte lines
      np += I.n;
      nn += std::max<uint64_t>(I.n - 1, 1);
      nn4 += rti_impl::wide_capacity((uint32_t)I.n);
      maxn = std::max(maxn, I.n);
    } else {
      I.prim_begin = ctx->inst[I.share].prim_begin;
      I.node_begin = ctx->inst[I.share].node_begin;
      I.node4_begin = ctx->inst[I.share].node4_begin;
    }
    I.acc_begin = (uint32_t)std::min<uint64_t>(na, UINT32_MAX);
    na += I.n;
    derive_box(I);
  }
  if (np >= (1ull << 31)) return fail(ctx, RTI_ERR_INVALID_ARG, "too many primitives in total");
  if (na >= (1ull << 31)) return fail(ctx, RTI_ERR_INVALID_ARG, "too many primitives over all instances");
  // four-wide node references keep bit 30 for the parent's leaf mark
  if (nn4 >= rti_dev::kNodeHasLeaves) return fail(ctx, RTI_ERR_INVALID_ARG, "too many tree nodes in total");
  ctx->n_total = np;
  ctx->acc_total = na;
  ctx->nodes_total = nn;
  // 32-byte records when every primitive is a disc, else 64-byte records
  bool discs_only = true;
  for (auto& I : ctx->inst) discs_only = discs_only && I.kind == 0;
  ctx->prim_shift = discs_only ? rti_dev::kPrimShiftDiscs : rti_dev::kPrimShiftMixed;
  // one sentinel record after the last primitive (its neighbour-row end)
  HIP_TRY(ctx, ctx->prims.ensure(((np + 1) * 16) << ctx->prim_shift));
  HIP_TRY(ctx, ctx->perm.ensure(np * 4));
  HIP_TRY(ctx, ctx->nodes.ensure(nn * sizeof(bvh_node)));
  HIP_TRY(ctx, ctx->nodes4.ensure(nn4 * sizeof(bvh4_node)));
  HIP_TRY(ctx, ctx->nb_ptr.ensure((np + 1) * 4));
  HIP_TRY(ctx, ctx->areas.ensure(np * 4));
  HIP_TRY(ctx, ctx->acc.ensure(na * 8 * rti_dev::kAccStride));
  HIP_TRY(ctx, ctx->queue.ensure(64));
  HIP_TRY(ctx, ctx->stats.ensure(rti_impl::kStatsWords * 8));
  HIP_TRY(ctx, ctx->desc.ensure(ctx->inst.size() * sizeof(inst_desc)));
  rti_impl::store_ptrs st{ctx->prims.as<float4>(), ctx->prim_shift, ctx->perm.as<uint32_t>(), ctx->nodes.as<bvh_node>(), ctx->nodes4.as<bvh4_node>(),
                          ctx->nb_ptr.as<uint32_t>(), nullptr, ctx->areas.as<float>()};
  HIP_TRY(ctx, hipError_t(rti_impl::scratch_reserve(ctx->scratch, (uint32_t)maxn)));
  dbuf in_a, in_b;
  std::vector<rti_impl::build_job> jobs(ctx->inst.size());
  uint64_t edges = 0;
  for (size_t k = 0; k < ctx->inst.size(); ++k) {
    instance& I = ctx->inst[k];
    rti_impl::build_job& J = jobs[k];
    J = rti_impl::build_job{};
    if (I.share >= 0) continue;  // built with the instance it shares
    J.prim_kind = I.kind;
    J.n = (uint32_t)I.n;
    J.prim_begin = I.prim_begin;
    J.node_begin = I.node_begin;
    J.node4_begin = I.node4_begin;
    centroid_bounds(I, J.cmin, J.cmax);
    J.rmax = I.rmax;
    // Leaf boxes are padded by 2e-6 of the instance's coordinate scale (its
    // extent and source plane), four times the rounding error of a slab
    // distance against that of a primitive's own hit distance for ray origins
    // of that scale: a primitive hit at exactly the closest distance found so
    // far (coplanar discs tie exactly) then always lies inside the boxes the
    // traversal keeps, whatever the order of the tests.
    {
      float m = std::max(1.0f, std::fabs(I.ztop));
      for (int k = 0; k < 3; ++k) m = std::max(m, std::max(std::fabs(I.geo_lo[k]), std::fabs(I.geo_hi[k])));
      if (!std::isnan(I.params.src_z)) m = std::max(m, std::fabs(I.params.src_z));
      J.pad = 2e-6f * m;
    }
    J.area_box[0] = I.box[0];
    J.area_box[1] = I.box[2];
    J.area_box[2] = I.box[1];
    J.area_box[3] = I.box[3];
    J.builder = tree_builder(ctx->params, I.kind);
    if (I.kind == 0) {
      HIP_TRY(ctx, in_a.ensure(I.n * 16));
      HIP_TRY(ctx, in_b.ensure(I.n * 12));
      HIP_TRY(ctx, hipMemcpyAsync(in_a.p, I.xyzr.data(), I.n * 16, hipMemcpyHostToDevice, ctx->stream));
      HIP_TRY(ctx, hipMemcpyAsync(in_b.p, I.nrm.data(), I.n * 12, hipMemcpyHostToDevice, ctx->stream));
      J.d_xyzr = in_a.as<float>();
      J.d_nrm = in_b.as<float>();
    } else {
      HIP_TRY(ctx, in_a.ensure(I.verts.size() * 4));
      HIP_TRY(ctx, in_b.ensure(I.tris.size() * 4));
      HIP_TRY(ctx, hipMemcpyAsync(in_a.p, I.verts.data(), I.verts.size() * 4, hipMemcpyHostToDevice, ctx->stream));
      HIP_TRY(ctx, hipMemcpyAsync(in_b.p, I.tris.data(), I.tris.size() * 4, hipMemcpyHostToDevice, ctx->stream));
      J.d_verts = in_a.as<float>();
      J.d_tris = in_b.as<uint32_t>();
    }
    uint32_t depth = 0;
    HIP_TRY(ctx, rti_impl::build_instance(J, st, ctx->scratch, &depth, ctx->stream));
    I.depth = depth;
    // the neighbour search walks the binary tree with a 64-entry stack
    if (depth > kMaxBinaryDepth && I.kind == 0)
      return fail(ctx, RTI_ERR_INVALID_ARG,
                  "hierarchy depth %u exceeds the neighbour search's bound %d: use tree_builder = RTI_TREE_LBVH",
                  depth, kMaxBinaryDepth);
    HIP_TRY(ctx, rti_impl::build_wide(J, st, ctx->scratch, &I.n_wide, &I.wide_stack, ctx->stream));
    // neighbour degrees -> row offsets (absolute)
    uint32_t* rowp = ctx->nb_ptr.as<uint32_t>() + I.prim_begin;
    if (I.kind == 0) {
      HIP_TRY(ctx, rti_impl::neighbor_count(J, st, ctx->scratch, ctx->stream));
      k_set_u32<<<1, 1, 0, ctx->stream>>>(ctx->scratch.counts + I.n, 1, 0u);
      HIP_TRY(ctx, rti_impl::exclusive_scan_u32(ctx->scratch.counts, rowp, (uint32_t)I.n + 1,
                                                ctx->scratch.scan_tmp, ctx->stream));
      uint32_t e_local = 0;
      HIP_TRY(ctx, hipMemcpyAsync(&e_local, rowp + I.n, 4, hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      k_add_u32<<<(unsigned)((I.n + 1 + 255) / 256), 256, 0, ctx->stream>>>(rowp, (uint32_t)I.n + 1,
                                                                               (uint32_t)edges);
      I.n_edges = e_local;
      edges += e_local;
      if (edges >= (1ull << 32)) return fail(ctx, RTI_ERR_INVALID_ARG, "neighbour graph too large");
    } else {
      k_set_u32<<<(unsigned)((I.n + 1 + 255) / 256), 256, 0, ctx->stream>>>(rowp, (uint32_t)I.n + 1,
                                                                              (uint32_t)edges);
      I.n_edges = 0;
    }
    HIP_TRY(ctx, hipGetLastError());
  }
  ctx->edges_total = edges;
  for (auto& I : ctx->inst) {
    if (I.share < 0) continue;
    const instance& O = ctx->inst[I.share];
    I.depth = O.depth;
    I.n_wide = O.n_wide;
    I.wide_stack = O.wide_stack;
    I.n_edges = O.n_edges;
  }
  k_link_rows<<<(unsigned)((np + 1 + 255) / 256), 256, 0, ctx->stream>>>(ctx->prims.as<float4>(), ctx->prim_shift,
                                                                         ctx->nb_ptr.as<uint32_t>(), (uint32_t)np);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, ctx->nb_col.ensure(std::max<uint64_t>(edges, 1) * 4));
  st.nb_col = ctx->nb_col.as<uint32_t>();
  for (size_t k = 0; k < ctx->inst.size(); ++k) {
    instance& I = ctx->inst[k];
    if (I.kind != 0 || I.n_edges == 0 || I.share >= 0) continue;
    // inputs are not needed: the fill pass reads the stored prims
    HIP_TRY(ctx, rti_impl::neighbor_fill(jobs[k], st, ctx->stream));
  }
  HIP_TRY(ctx, hipEventRecord(b1, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIP_TRY(ctx, hipEventElapsedTime(&ms, b0, b1));
  (void)hipEventDestroy(b0);
  (void)hipEventDestroy(b1);
  ctx->build_ms = ms;
  ctx->dirty = false;
  ctx->acc_traced = false;
  return RTI_OK;
}

// Expected front-face hits of one path: the weight after k hits is (1-s)^k
// and rejection control starts below 0.1 (rejection_control.hpp:27-61).
double path_cost(const rti_params& p) {
  const double s = p.sticking;
  if (s >= 1.0) return 1.0;
  return std::min(1e4, std::log(0.1) / std::log1p(-s) + 1.0);
}

// Descriptors for the current per-instance parameters (cheap; done per
// trace). The global ray queue holds the instances' ray ranges in order of
// decreasing expected path length, so the longest paths start first and the
// launch does not end on a tail of long paths. The order does not change any
// result: an instance's paths depend on its own key and ray ids only.
int upload_descriptors(rti_ctx* ctx, inst_desc* dst) {
  const size_t n = ctx->inst.size();
  std::vector<size_t> order(n);
  for (size_t k = 0; k < n; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return path_cost(ctx->inst[a].params) > path_cost(ctx->inst[b].params);
  });
  std::vector<inst_desc>& D = ctx->desc_host;  // the asynchronous copy reads it after this returns
  D.assign(n, inst_desc{});
  uint64_t rb = 0;
  for (size_t j = 0; j < n; ++j) {
    const size_t k = order[j];
    instance& I = ctx->inst[k];
    derive_box(I);  // source height may have changed with the parameters
    const rti_params& p = I.params;
    inst_desc& d = D[j];
    std::memset(&d, 0, sizeof d);
    d.xmin = I.box[0];
    d.xmax = I.box[1];
    d.ymin = I.box[2];
    d.ymax = I.box[3];
    d.zmin = I.box[4];
    d.zmax = I.box[5];
    d.src_z = I.src_z;
    d.tnear = p.tnear;
    d.sticking = p.sticking;
    d.cos_exponent = p.cos_exponent;
    d.diffuse_prob = p.diffuse_prob;
    d.pow_e = 2.0f / (p.cos_exponent + 1.0f);
    d.score_c = 1.0f / (p.cos_exponent + 1.0f);
    d.bc_x = p.bc_x;
    d.bc_y = p.bc_y;
    d.prim_kind = I.kind;
    d.root = (int32_t)I.node_begin;
    d.root4 = (int32_t)I.node4_begin;
    d.source = p.source;
    d.ray_begin = rb;
    d.ray_id0 = p.ray_begin;
    d.seed = p.seed;
    d.key1 = p.instance + (uint32_t)k;
    d.roulette = p.roulette ? 1u : 0u;
    d.acc_off = (int32_t)I.acc_begin - (int32_t)I.prim_begin;
    d.n_prims = (uint32_t)I.n;
    // 2^shift with shift = 58 - bitlen(rays of the job), at most 40: each
    // deposit adds at most 2^shift, so an element can only overflow after
    // more than 32 deposits per ray on average (rti_finalize checks)
    I.shift = std::min(40, 58 - bit_length(p.rays_total ? p.rays_total : p.n_rays));
    d.fx_scale = std::ldexp(1.0f, I.shift);
    // sum of squares: 2^shift2 with s^2 2^shift2 <= 2^(shift - 1) for s < 1
    // (2^shift for s = 1), so a squared deposit (<= s^2, rounded in FP32) adds
    // no more than a deposit can (the same overflow bound), and the smallest
    // squared deposit with rejection control, (0.1 s)^2, still spans more
    // than 2^(shift - 9) grid units
    const int lg = (int)std::floor(std::log2(1.0 / ((double)p.sticking * (double)p.sticking)));
    I.shift2 = I.shift + std::max(0, lg - 1);
    d.fx_scale2 = std::ldexp(1.0f, I.shift2);
    I.ray_begin = rb;
    rb += p.n_rays;
  }
  ctx->rays_total = rb;
  HIP_TRY(ctx, hipMemcpyAsync(dst, D.data(), D.size() * sizeof(inst_desc), hipMemcpyHostToDevice, ctx->stream));
  return RTI_OK;
}

// Every instance of the context is a triangle mesh: each four-wide node
// keeps its leaves contiguous, and the kernels need no leaf word.
int all_meshes(const rti_ctx* ctx) {
  for (const auto& I : ctx->inst)
    if (I.kind != 1) return 0;
  return 1;
}

struct kernel_cfg {
  int block, lds, grid, width;
  int leaf;  // postponed-leaf threshold (64ths), 0 = leaf tests at each node
};

int pick_config(rti_ctx* ctx, kernel_cfg& cfg, bool adj, bool fp64) {
  const rti_params& p = ctx->params;
  cfg.block = p.block_size ? p.block_size : 256;
  cfg.width = p.bvh_width ? p.bvh_width : 4;
  if (cfg.width == 2) {
    // the binary traversal's lane stack holds 64 entries
    for (auto& I : ctx->inst)
      if ((int)I.depth > kMaxBinaryDepth)
        return fail(ctx, RTI_ERR_INVALID_ARG,
                    "binary traversal needs a hierarchy of depth <= %d, this one has %u: use bvh_width = 4 or "
                    "tree_builder = RTI_TREE_LBVH",
                    kMaxBinaryDepth, I.depth);
  }
  cfg.leaf = p.leaf_batch < 0 ? 0
             : (p.leaf_batch ? p.leaf_batch : rti_impl::default_leaf_batch(ctx->prim_shift == rti_dev::kPrimShiftMixed));
  if (cfg.width == 4 && cfg.leaf > 0 && ctx->n_total >= rti_impl::postponed_leaf_prim_limit()) {
    // leaf words on the wave's list hold 25-bit primitive ids
    if (p.leaf_batch > 0)
      return fail(ctx, RTI_ERR_INVALID_ARG, "leaf_batch > 0 needs fewer than %llu primitives in the store (%llu)",
                  (unsigned long long)rti_impl::postponed_leaf_prim_limit(), (unsigned long long)ctx->n_total);
    cfg.leaf = 0;
  }
  cfg.lds = p.lds_stack ? p.lds_stack : rti_impl::default_lds_stack(cfg.width, cfg.width == 4 && cfg.leaf > 0);
  if (cfg.width == 4) {
    for (auto& I : ctx->inst)
      if ((int)I.wide_stack > rti_impl::wide_stack_capacity())
        return fail(ctx, RTI_ERR_INVALID_ARG,
                    "four-wide traversal needs %u stack entries (capacity %d): use bvh_width = 2", I.wide_stack,
                    rti_impl::wide_stack_capacity());
  }
  int per_cu = 0;
  HIP_TRY(ctx, rti_impl::trace_occupancy(cfg.lds, cfg.block, 1, cfg.leaf, adj, fp64, cfg.width, all_meshes(ctx),
                                         &per_cu));
  if (per_cu < 1) per_cu = 1;
  if (p.waves_per_cu > 0) per_cu = std::max(1, std::min(per_cu, p.waves_per_cu * 64 / cfg.block));
  cfg.grid = per_cu * ctx->num_cus;
  return RTI_OK;
}

int alloc_accumulators(rti_ctx* ctx, bool adj) {
  if (adj) HIP_TRY(ctx, ctx->adj.ensure(ctx->acc_total * 16));
  if (ctx->params.moments) HIP_TRY(ctx, ctx->mom.ensure(ctx->acc_total * 16));
  return RTI_OK;
}

// Device memory one trace writes besides the scene store: the accumulator
// records, the instance descriptors, the ray queue and the counters (the
// context's own buffers, or the caller's records and the static scratch of
// the no-allocation entry point).
struct trace_bufs {
  unsigned long long* acc;
  inst_desc* desc;
  unsigned long long* queue;
  unsigned long long* stats;  // kStatsWords counters
};

trace_bufs ctx_bufs(rti_ctx* ctx) {
  return {ctx->acc.as<unsigned long long>(), ctx->desc.as<inst_desc>(), ctx->queue.as<unsigned long long>(),
          ctx->stats.as<unsigned long long>()};
}

// Clears accumulators / queue / counters and launches the trace on the stream.
// `only` selects one instance for the plain per-geometry path (-1: all).
int enqueue_trace(rti_ctx* ctx, const trace_bufs& B, const kernel_cfg& cfg, bool adj, bool fp64, bool count,
                  int kernel, long only) {
  const uint64_t np = ctx->acc_total;  // accumulator slots
  HIP_TRY(ctx, rti_impl::fill_u64(B.acc, np * rti_dev::kAccStride, 0ull, ctx->stream));
  if (adj) HIP_TRY(ctx, rti_impl::fill_u64(ctx->adj.as<unsigned long long>(), np * 2, 0ull, ctx->stream));
  const bool mom = ctx->params.moments != 0;
  if (mom)  // 0.0 has all bits zero
    HIP_TRY(ctx, rti_impl::fill_u64(ctx->mom.as<unsigned long long>(), np * 2, 0ull, ctx->stream));
  HIP_TRY(ctx, rti_impl::fill_u64(B.stats, rti_impl::kStatsWords, 0ull, ctx->stream));
  rti_impl::trace_launch L{};
  L.desc = B.desc;
  L.n_inst = (uint32_t)ctx->inst.size();
  L.prims = ctx->prims.as<float4>();
  L.prim_shift = ctx->prim_shift;
  L.nodes = ctx->nodes.as<bvh_node>();
  L.nodes4 = ctx->nodes4.as<bvh4_node>();
  L.nb_col = ctx->nb_col.as<uint32_t>();
  L.acc = B.acc;
  L.adj = adj ? ctx->adj.as<unsigned long long>() : nullptr;
  L.mom = mom ? ctx->mom.as<double>() : nullptr;
  L.queue = B.queue;
  L.stats = B.stats;
  L.adjoint = adj ? 1 : 0;
  L.accum_fp64 = fp64 ? 1 : 0;
  L.kernel = kernel;
  L.block_size = cfg.block;
  L.lds_stack = cfg.lds;
  L.grid_blocks = cfg.grid;
  L.count = count ? 1 : 0;
  L.compact = ctx->params.compact_lanes;
  L.width = cfg.width;
  L.leaf_batch = cfg.leaf;
  L.refill_lanes = ctx->params.refill_lanes ? ctx->params.refill_lanes : 1;
  L.mesh = all_meshes(ctx);
  if (kernel == RTI_KERNEL_PERSISTENT) {
    L.ray_lo = 0;
    L.ray_hi = ctx->rays_total;
    k_set_one_u64<<<1, 1, 0, ctx->stream>>>(L.queue, 0ull);
    HIP_TRY(ctx, hipGetLastError());
    if (L.ray_hi > 0) HIP_TRY(ctx, rti_impl::launch_trace(L, ctx->stream));
  } else {
    // plain: one launch per geometry
    for (size_t k = 0; k < ctx->inst.size(); ++k) {
      if (only >= 0 && (long)k != only) continue;
      const instance& I = ctx->inst[k];
      if (I.params.n_rays == 0) continue;
      L.ray_lo = I.ray_begin;
      L.ray_hi = I.ray_begin + I.params.n_rays;
      HIP_TRY(ctx, rti_impl::launch_trace(L, ctx->stream));
    }
  }
  return RTI_OK;
}

int trace(rti_ctx* ctx) {
  int rc = prepare(ctx);
  if (rc) return rc;
  const bool adj = ctx->params.adjoint != 0;
  const bool fp64 = ctx->params.accum == RTI_ACCUM_FP64;
  const bool count = ctx->params.kernel == RTI_KERNEL_COUNT;
  const trace_bufs B = ctx_bufs(ctx);
  rc = upload_descriptors(ctx, B.desc);
  if (rc) return rc;
  rc = alloc_accumulators(ctx, adj);
  if (rc) return rc;
  kernel_cfg cfg;
  rc = pick_config(ctx, cfg, adj, fp64);
  if (rc) return rc;
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  rc = enqueue_trace(ctx, B, cfg, adj, fp64, count, ctx->params.kernel == RTI_KERNEL_PLAIN ? 1 : 0, -1);
  if (rc) return rc;
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->trace_ms = ms;
  ctx->last_adjoint = adj;
  ctx->last_fp64 = fp64;
  ctx->last_count = count;
  ctx->last_moments = ctx->params.moments != 0;
  HIP_TRY(ctx, hipMemcpy(ctx->run_stats, ctx->stats.p, sizeof ctx->run_stats, hipMemcpyDeviceToHost));
  ctx->have_results = false;
  ctx->acc_traced = true;
  return RTI_OK;
}

int finalize(rti_ctx* ctx) {
  if (ctx->dirty || !ctx->acc_traced) return fail(ctx, RTI_ERR_STATE, "no trace for the current geometry");
  hipEvent_t a, b;
  HIP_TRY(ctx, hipEventCreate(&a));
  HIP_TRY(ctx, hipEventCreate(&b));
  HIP_TRY(ctx, hipEventRecord(a, ctx->stream));
  uint64_t maxn = 0;
  for (auto& I : ctx->inst) maxn = std::max(maxn, I.n);
  HIP_TRY(ctx, ctx->raw_st.ensure(maxn * 8));
  HIP_TRY(ctx, ctx->vtmp.ensure(maxn * 8));
  HIP_TRY(ctx, ctx->maxbuf.ensure(64));
  HIP_TRY(ctx, ctx->out_raw.ensure(maxn * 8));
  HIP_TRY(ctx, ctx->out_sigma.ensure(maxn * 8));
  HIP_TRY(ctx, ctx->out_cnt.ensure(maxn * 8));
  HIP_TRY(ctx, ctx->out_smooth.ensure(maxn * 4));
  HIP_TRY(ctx, ctx->out_f.ensure(maxn * 4));
  if (ctx->last_adjoint) HIP_TRY(ctx, ctx->out_adj.ensure(maxn * 16));
  HIP_TRY(ctx, ctx->out_err.ensure(maxn * 16));
  for (auto& I : ctx->inst) {
    const uint32_t n = (uint32_t)I.n;
    const double inv = std::ldexp(1.0, -I.shift), inv2 = std::ldexp(1.0, -I.shift2);
    HIP_TRY(ctx, rti_impl::finalize_instance(
                     ctx->acc.as<unsigned long long>(), I.acc_begin, ctx->perm.as<uint32_t>(), I.prim_begin, n, inv,
                     inv2, ctx->last_fp64, I.params.n_rays, ctx->raw_st.as<double>(), ctx->out_raw.as<double>(),
                     ctx->out_sigma.as<double>(), ctx->out_cnt.as<unsigned long long>(), ctx->stream));
    HIP_TRY(ctx, rti_impl::smooth_normalize(ctx->raw_st.as<double>(), ctx->areas.as<float>(),
                                            ctx->perm.as<uint32_t>(), I.kind == 0 ? ctx->nb_ptr.as<uint32_t>() : nullptr,
                                            ctx->nb_col.as<uint32_t>(), I.prim_begin, n, ctx->vtmp.as<double>(),
                                            ctx->maxbuf.as<double>(), ctx->out_smooth.as<float>(), ctx->stream));
    HIP_TRY(ctx, rti_impl::permute_f32(ctx->areas.as<float>() + I.prim_begin, ctx->perm.as<uint32_t>() + I.prim_begin,
                                       n, ctx->out_f.as<float>(), ctx->stream));
    I.raw.resize(n);
    I.sigma.resize(n);
    I.counts.resize(n);
    I.smoothed.resize(n);
    I.areas.resize(n);
    HIP_TRY(ctx, hipMemcpyAsync(I.raw.data(), ctx->out_raw.p, n * 8ull, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(I.sigma.data(), ctx->out_sigma.p, n * 8ull, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(I.counts.data(), ctx->out_cnt.p, n * 8ull, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(I.smoothed.data(), ctx->out_smooth.p, n * 4ull, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(I.areas.data(), ctx->out_f.p, n * 4ull, hipMemcpyDeviceToHost, ctx->stream));
    if (ctx->last_adjoint) {
      const double inva = std::ldexp(1.0, -(I.shift - 8));
      double* o1 = ctx->out_adj.as<double>();
      double* o2 = o1 + maxn;
      HIP_TRY(ctx, rti_impl::finalize_signed(ctx->adj.as<unsigned long long>(), 0, I.acc_begin,
                                             ctx->perm.as<uint32_t>(), I.prim_begin, n, inva, o1, ctx->stream));
      HIP_TRY(ctx, rti_impl::finalize_signed(ctx->adj.as<unsigned long long>(), 1, I.acc_begin,
                                             ctx->perm.as<uint32_t>(), I.prim_begin, n, inva, o2, ctx->stream));
      I.dds.resize(n);
      I.ddn.resize(n);
      HIP_TRY(ctx, hipMemcpyAsync(I.dds.data(), o1, n * 8ull, hipMemcpyDeviceToHost, ctx->stream));
      HIP_TRY(ctx, hipMemcpyAsync(I.ddn.data(), o2, n * 8ull, hipMemcpyDeviceToHost, ctx->stream));
    } else {
      I.dds.clear();
      I.ddn.clear();
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    // overflow guard: each deposit adds at most 2^shift to either sum, so no
    // element can have wrapped unless its deposit count reaches 2^(63 - shift)
    if (!ctx->last_fp64) {
      uint64_t mc = 0;
      for (uint64_t c : I.counts) mc = std::max(mc, c);
      if (mc >= (1ull << (63 - I.shift)))
        return fail(ctx, RTI_ERR_OVERFLOW, "fixed-point range exceeded (%llu deposits on one element)",
                    (unsigned long long)mc);
    }
    // relative error (always) and variance of the variance (with the moment
    // sums), N = all deposits of the instance (hit_accumulator.hpp:176)
    uint64_t total = 0;
    for (uint64_t c : I.counts) total += c;
    double* er = ctx->out_err.as<double>();
    double* ev = er + maxn;
    const bool mom = ctx->last_moments != 0;
    HIP_TRY(ctx, rti_impl::error_stats(ctx->acc.as<unsigned long long>(), mom ? ctx->mom.as<double>() : nullptr,
                                       I.acc_begin, ctx->perm.as<uint32_t>(), I.prim_begin, n, inv, inv2,
                                       ctx->last_fp64, total, er, mom ? ev : nullptr, ctx->stream));
    I.rel.resize(n);
    HIP_TRY(ctx, hipMemcpyAsync(I.rel.data(), er, n * 8ull, hipMemcpyDeviceToHost, ctx->stream));
    if (mom) {
      I.vov.resize(n);
      HIP_TRY(ctx, hipMemcpyAsync(I.vov.data(), ev, n * 8ull, hipMemcpyDeviceToHost, ctx->stream));
    } else {
      I.vov.clear();
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  HIP_TRY(ctx, hipEventRecord(b, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIP_TRY(ctx, hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  ctx->post_ms = ms;
  ctx->have_results = true;
  return RTI_OK;
}

void reset_graph(rti_ctx* ctx) {
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
  ctx->graph_key.clear();
}

int set_single(rti_ctx* ctx, instance&& I) {
  I.params = ctx->params;
  geometry_bounds(I);
  ctx->inst.clear();
  ctx->inst.push_back(std::move(I));
  ctx->batch_mode = false;
  ctx->dirty = true;
  ctx->have_results = false;
  reset_graph(ctx);
  return RTI_OK;
}

void fill_stats(rti_ctx* ctx, const instance* I, rti_stats* out) {
  std::memset(out, 0, sizeof *out);
  uint64_t np = 0, nr = 0, ne = 0;
  uint32_t depth = 0, wstack = 0;
  uint64_t nw = 0;
  for (auto& J : ctx->inst) {
    np += J.n;
    nr += J.params.n_rays;
    depth = std::max(depth, J.depth);
    wstack = std::max(wstack, J.wide_stack);
    if (J.share >= 0) continue;  // its geometry is counted with the instance it shares
    ne += J.n_edges;
    nw += J.n_wide;
  }
  out->n_prims = I ? I->n : np;
  out->n_rays = I ? I->params.n_rays : nr;
  out->segments = ctx->run_stats[0];
  out->surface_hits = ctx->run_stats[1];
  out->misses = ctx->run_stats[2];
  out->truncated = ctx->run_stats[3];
  out->max_segments = ctx->run_stats[8];
  out->n_edges = I ? I->n_edges : ne;
  out->bvh_depth = I ? I->depth : depth;
  const instance& J = I ? *I : ctx->inst[0];
  out->fixed_shift = J.shift;
  out->trace_ms = ctx->trace_ms;
  out->build_ms = ctx->build_ms;
  out->post_ms = ctx->post_ms;
  std::memcpy(out->box, J.box, sizeof out->box);
  out->src_z = J.src_z;
  out->wide_stack = I ? I->wide_stack : wstack;
  out->wide_nodes = I ? I->n_wide : nw;
  out->fixed_shift2 = J.shift2;
}

int get_error_impl(rti_ctx* ctx, const instance& I, double* rel_err, double* vov) {
  if (!ctx->have_results) return fail(ctx, RTI_ERR_STATE, "no results: run first");
  if (vov && I.vov.empty())
    return fail(ctx, RTI_ERR_STATE, "variance of the variance needs a run made with params.moments = 1");
  if (rel_err) std::memcpy(rel_err, I.rel.data(), I.n * 8);
  if (vov) std::memcpy(vov, I.vov.data(), I.n * 8);
  return RTI_OK;
}

}  // namespace

extern "C" {

void rti_default_params(rti_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->n_rays = 1024;  // device.hpp:289
  p->seed = 1;
  p->instance = 0;
  p->bc_x = RTI_BC_REFLECTIVE;  // device.hpp:292-293
  p->bc_y = RTI_BC_REFLECTIVE;
  p->sticking = 1.0f;
  p->cos_exponent = 1.0f;
  p->diffuse_prob = 1.0f;
  p->tnear = 1e-4f;  // tracer.hpp:179
  p->roulette = 1;
  p->src_z = NAN;
  p->kernel = RTI_KERNEL_PERSISTENT;
  p->adjoint = 0;
  p->accum = RTI_ACCUM_FIXED64;
}

int rti_device_info(int device, char* name, uint64_t cap, int* num_cus) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RTI_ERR_HIP;
  if (device < 0 || device >= ndev) return RTI_ERR_INVALID_ARG;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return RTI_ERR_HIP;
  if (name && cap > 0) {
    std::strncpy(name, prop.gcnArchName, cap - 1);
    name[cap - 1] = '\0';
  }
  if (num_cus) *num_cus = prop.multiProcessorCount;
  return RTI_OK;
}

int rti_ctx_create(rti_ctx** out, int device) {
  if (!out) return RTI_ERR_INVALID_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RTI_ERR_HIP;
  if (device < 0 || device >= ndev) return RTI_ERR_INVALID_ARG;
  rti_ctx* ctx = new rti_ctx;
  ctx->device = device;
  rti_default_params(&ctx->params);
  // the device address of the static scratch is resolved here, once per
  // device, with the device current
  if (hipSetDevice(device) != hipSuccess || resolve_scratch(device) != hipSuccess ||
      hipStreamCreate(&ctx->stream) != hipSuccess || hipEventCreate(&ctx->ev0) != hipSuccess ||
      hipEventCreate(&ctx->ev1) != hipSuccess ||
      hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RTI_ERR_HIP;
  }
  *out = ctx;
  return RTI_OK;
}

int rti_ctx_destroy(rti_ctx* ctx) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  reset_graph(ctx);
  rti_impl::scratch_free(ctx->scratch);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RTI_OK;
}

const char* rti_last_error(const rti_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rti_set_discs(rti_ctx* ctx, const float* xyzr, const float* nrm, uint64_t n) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  int rc = check_discs(ctx, xyzr, nrm, n);
  if (rc) return rc;
  instance I;
  I.kind = 0;
  I.n = n;
  I.xyzr.assign(xyzr, xyzr + 4 * n);
  I.nrm.assign(nrm, nrm + 3 * n);
  return set_single(ctx, std::move(I));
}

int rti_set_triangles(rti_ctx* ctx, const float* verts, uint64_t n_verts, const uint32_t* tris, uint64_t n_tris) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  int rc = check_tris(ctx, verts, n_verts, tris, n_tris);
  if (rc) return rc;
  instance I;
  I.kind = 1;
  I.n = n_tris;
  I.verts.assign(verts, verts + 3 * n_verts);
  I.tris.assign(tris, tris + 3 * n_tris);
  return set_single(ctx, std::move(I));
}

int rti_set_params(rti_ctx* ctx, const rti_params* p) {
  if (!ctx || !p) return fail(ctx, RTI_ERR_INVALID_ARG, "null argument");
  rti_params q = *p;
  int rc = check_params(ctx, q);
  if (rc) return rc;
  // another tree topology needs another build
  if (q.tree_builder != ctx->params.tree_builder) ctx->dirty = true;
  ctx->params = q;
  if (!ctx->batch_mode)
    for (auto& I : ctx->inst) I.params = q;
  ctx->have_results = false;
  reset_graph(ctx);
  return RTI_OK;
}

int rti_prepare(rti_ctx* ctx) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  return prepare(ctx);
}

int rti_trace(rti_ctx* ctx) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  return trace(ctx);
}

int rti_finalize(rti_ctx* ctx) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  return finalize(ctx);
}

int rti_run(rti_ctx* ctx) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = trace(ctx);
  if (rc) return rc;
  return finalize(ctx);
}

static int get_flux_impl(rti_ctx* ctx, const instance& I, float* smoothed, double* raw, double* sigma) {
  if (!ctx->have_results) return fail(ctx, RTI_ERR_STATE, "no results: call rti_run first");
  if (smoothed) std::memcpy(smoothed, I.smoothed.data(), I.n * 4);
  if (raw) std::memcpy(raw, I.raw.data(), I.n * 8);
  if (sigma) std::memcpy(sigma, I.sigma.data(), I.n * 8);
  return RTI_OK;
}

int rti_get_flux(rti_ctx* ctx, float* smoothed_norm, double* raw, double* sigma) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  if (ctx->inst.empty()) return fail(ctx, RTI_ERR_NO_GEOMETRY, "no geometry");
  return get_flux_impl(ctx, ctx->inst[0], smoothed_norm, raw, sigma);
}

int rti_get_counts(rti_ctx* ctx, uint64_t* counts) {
  if (!ctx || !counts) return fail(ctx, RTI_ERR_INVALID_ARG, "null argument");
  if (!ctx->have_results) return fail(ctx, RTI_ERR_STATE, "no results: call rti_run first");
  std::memcpy(counts, ctx->inst[0].counts.data(), ctx->inst[0].n * 8);
  return RTI_OK;
}

int rti_get_areas(rti_ctx* ctx, float* areas) {
  if (!ctx || !areas) return fail(ctx, RTI_ERR_INVALID_ARG, "null argument");
  if (!ctx->have_results) return fail(ctx, RTI_ERR_STATE, "no results: call rti_run first");
  std::memcpy(areas, ctx->inst[0].areas.data(), ctx->inst[0].n * 4);
  return RTI_OK;
}

int rti_get_neighbors(rti_ctx* ctx, uint32_t* row_ptr, uint32_t* col, uint64_t cap, uint64_t* n_edges) {
  if (!ctx || !row_ptr) return fail(ctx, RTI_ERR_INVALID_ARG, "null argument");
  if (ctx->inst.empty()) return fail(ctx, RTI_ERR_NO_GEOMETRY, "no geometry");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = prepare(ctx);
  if (rc) return rc;
  const instance& I = ctx->inst[0];
  const uint64_t n = I.n;
  std::vector<uint32_t> rp(n + 1), pm(n);
  HIP_TRY(ctx, hipMemcpy(rp.data(), ctx->nb_ptr.as<uint32_t>() + I.prim_begin, (n + 1) * 4, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(pm.data(), ctx->perm.as<uint32_t>() + I.prim_begin, n * 4, hipMemcpyDeviceToHost));
  const uint64_t e = rp[n] - rp[0];
  if (n_edges) *n_edges = e;
  std::vector<uint32_t> cl(std::max<uint64_t>(e, 1));
  if (e) HIP_TRY(ctx, hipMemcpy(cl.data(), ctx->nb_col.as<uint32_t>() + rp[0], e * 4, hipMemcpyDeviceToHost));
  // rows in input order, columns as input ids, ascending
  std::vector<uint32_t> deg(n);
  for (uint64_t j = 0; j < n; ++j) deg[pm[j]] = rp[j + 1] - rp[j];
  row_ptr[0] = 0;
  for (uint64_t i = 0; i < n; ++i) row_ptr[i + 1] = row_ptr[i] + deg[i];
  if (col) {
    if (cap < e) return fail(ctx, RTI_ERR_INVALID_ARG, "col capacity %llu < %llu edges", (unsigned long long)cap,
                             (unsigned long long)e);
    for (uint64_t j = 0; j < n; ++j) {
      uint32_t i = pm[j];
      uint32_t w = row_ptr[i];
      for (uint32_t q = rp[j]; q < rp[j + 1]; ++q) col[w++] = pm[cl[q - rp[0]] - I.prim_begin];
      std::sort(col + row_ptr[i], col + row_ptr[i + 1]);
    }
  }
  return RTI_OK;
}

int rti_get_stats(rti_ctx* ctx, rti_stats* out) {
  if (!ctx || !out) return fail(ctx, RTI_ERR_INVALID_ARG, "null argument");
  if (ctx->inst.empty()) return fail(ctx, RTI_ERR_NO_GEOMETRY, "no geometry");
  fill_stats(ctx, &ctx->inst[0], out);
  return RTI_OK;
}

int rti_get_adjoint(rti_ctx* ctx, double* draw_ds, double* draw_dn) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  if (!ctx->have_results) return fail(ctx, RTI_ERR_STATE, "no results: call rti_run first");
  if (ctx->inst[0].dds.empty()) return fail(ctx, RTI_ERR_STATE, "last run was not made with params.adjoint = 1");
  if (draw_ds) std::memcpy(draw_ds, ctx->inst[0].dds.data(), ctx->inst[0].n * 8);
  if (draw_dn) std::memcpy(draw_dn, ctx->inst[0].ddn.data(), ctx->inst[0].n * 8);
  return RTI_OK;
}

int rti_get_error_stats(rti_ctx* ctx, double* rel_err, double* vov) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  if (ctx->inst.empty()) return fail(ctx, RTI_ERR_NO_GEOMETRY, "no geometry");
  return get_error_impl(ctx, ctx->inst[0], rel_err, vov);
}

int rti_batch_get_error_stats(rti_ctx* ctx, uint32_t i, double* rel_err, double* vov) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  if (i >= ctx->inst.size()) return fail(ctx, RTI_ERR_INVALID_ARG, "instance index out of range");
  return get_error_impl(ctx, ctx->inst[i], rel_err, vov);
}

int rti_batch_get_adjoint(rti_ctx* ctx, uint32_t i, double* draw_ds, double* draw_dn) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  if (i >= ctx->inst.size()) return fail(ctx, RTI_ERR_INVALID_ARG, "instance index out of range");
  if (!ctx->have_results) return fail(ctx, RTI_ERR_STATE, "no results: call rti_batch_run first");
  const instance& I = ctx->inst[i];
  if (I.dds.empty()) return fail(ctx, RTI_ERR_STATE, "last run was not made with params.adjoint = 1");
  if (draw_ds) std::memcpy(draw_ds, I.dds.data(), I.n * 8);
  if (draw_dn) std::memcpy(draw_dn, I.ddn.data(), I.n * 8);
  return RTI_OK;
}

int rti_get_counters(rti_ctx* ctx, uint64_t out[8]) {
  if (!ctx || !out) return RTI_ERR_INVALID_ARG;
  std::memcpy(out, ctx->run_stats, 64);
  return RTI_OK;
}

int rti_get_hierarchy(rti_ctx* ctx, void* nodes, uint64_t cap_nodes, uint32_t* perm, uint64_t* n_nodes,
                      float* prim_a) {
  if (!ctx || !n_nodes) return fail(ctx, RTI_ERR_INVALID_ARG, "null argument");
  if (ctx->inst.empty()) return fail(ctx, RTI_ERR_NO_GEOMETRY, "no geometry");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = prepare(ctx);
  if (rc) return rc;
  const instance& I = ctx->inst[0];
  const uint64_t nn = std::max<uint64_t>(I.n - 1, 1);
  *n_nodes = nn;
  if (nodes) {
    if (cap_nodes < nn) return fail(ctx, RTI_ERR_INVALID_ARG, "node capacity too small");
    HIP_TRY(ctx, hipMemcpy(nodes, ctx->nodes.as<bvh_node>() + I.node_begin, nn * sizeof(bvh_node),
                           hipMemcpyDeviceToHost));
  }
  if (perm) HIP_TRY(ctx, hipMemcpy(perm, ctx->perm.as<uint32_t>() + I.prim_begin, I.n * 4, hipMemcpyDeviceToHost));
  if (prim_a)  // the first float4 of each primitive record
    HIP_TRY(ctx, hipMemcpy2D(prim_a, 16, ctx->prims.as<float4>() + ((size_t)I.prim_begin << ctx->prim_shift),
                             (size_t)16 << ctx->prim_shift, 16, I.n, hipMemcpyDeviceToHost));
  return RTI_OK;
}

int rti_get_wide_hierarchy(rti_ctx* ctx, void* nodes, uint64_t cap_nodes, uint64_t* n_nodes) {
  if (!ctx || !n_nodes) return fail(ctx, RTI_ERR_INVALID_ARG, "null argument");
  if (ctx->inst.empty()) return fail(ctx, RTI_ERR_NO_GEOMETRY, "no geometry");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = prepare(ctx);
  if (rc) return rc;
  const instance& I = ctx->inst[0];
  *n_nodes = I.n_wide;
  if (nodes) {
    if (cap_nodes < I.n_wide) return fail(ctx, RTI_ERR_INVALID_ARG, "node capacity too small");
    HIP_TRY(ctx, hipMemcpy(nodes, ctx->nodes4.as<bvh4_node>() + I.node4_begin, I.n_wide * sizeof(bvh4_node),
                           hipMemcpyDeviceToHost));
    bvh4_node* w = static_cast<bvh4_node*>(nodes);
    for (uint32_t i = 0; i < I.n_wide; ++i) {  // relative to this geometry
      const int m = rti_dev::bvh4_nchild(w[i]), nint = rti_dev::bvh4_nint(w[i]);
      w[i].int_base -= nint ? I.node4_begin : w[i].int_base;
      for (int c = nint; c < m; ++c) w[i].leaf[c] -= I.prim_begin;
      if (w[i].meta & rti_dev::kNodeLeavesContiguous) w[i].leaf_info -= I.prim_begin;
    }
  }
  return RTI_OK;
}

int rti_batch_set(rti_ctx* ctx, const rti_instance* in, uint32_t n) {
  if (!ctx || !in || n == 0) return fail(ctx, RTI_ERR_INVALID_ARG, "empty batch");
  std::vector<instance> v(n);
  for (uint32_t k = 0; k < n; ++k) {
    const rti_instance& s = in[k];
    int rc = check_params(ctx, s.params);
    if (rc) return rc;
    instance& I = v[k];
    // the very same geometry buffers as an earlier instance: share its
    // device copy (particle species on one geometry)
    int owner = -1;
    for (uint32_t j = 0; j < k && owner < 0; ++j) {
      const rti_instance& o = in[j];
      if (v[j].share >= 0 || o.n_prims != s.n_prims) continue;
      const bool same = s.xyzr ? (o.xyzr == s.xyzr && o.nrm == s.nrm)
                               : (!o.xyzr && o.verts == s.verts && o.tris == s.tris && o.n_verts == s.n_verts);
      if (same) owner = (int)j;
    }
    if (owner >= 0) {
      const instance& O = v[owner];
      I.kind = O.kind;
      I.n = O.n;
      I.share = owner;
      I.params = s.params;
      I.rmax = O.rmax;
      I.ztop = O.ztop;
      for (int a = 0; a < 3; ++a) {
        I.geo_lo[a] = O.geo_lo[a];
        I.geo_hi[a] = O.geo_hi[a];
      }
      continue;
    }
    if (s.xyzr) {
      rc = check_discs(ctx, s.xyzr, s.nrm, s.n_prims);
      if (rc) return rc;
      I.kind = 0;
      I.xyzr.assign(s.xyzr, s.xyzr + 4 * s.n_prims);
      I.nrm.assign(s.nrm, s.nrm + 3 * s.n_prims);
    } else {
      rc = check_tris(ctx, s.verts, s.n_verts, s.tris, s.n_prims);
      if (rc) return rc;
      I.kind = 1;
      I.verts.assign(s.verts, s.verts + 3 * s.n_verts);
      I.tris.assign(s.tris, s.tris + 3 * s.n_prims);
    }
    I.n = s.n_prims;
    I.params = s.params;
    geometry_bounds(I);
  }
  ctx->inst = std::move(v);
  ctx->batch_mode = true;
  ctx->dirty = true;
  ctx->have_results = false;
  reset_graph(ctx);
  return RTI_OK;
}

int rti_batch_run(rti_ctx* ctx) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  // batch physics come from the instances; kernel knobs from rti_set_params
  int rc = trace(ctx);
  if (rc) return rc;
  return finalize(ctx);
}

int rti_batch_replay(rti_ctx* ctx, uint32_t times) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = prepare(ctx);
  if (rc) return rc;
  const bool adj = ctx->params.adjoint != 0;
  const bool fp64 = ctx->params.accum == RTI_ACCUM_FP64;
  const int kernel = ctx->params.kernel == RTI_KERNEL_PLAIN ? 1 : 0;
  const trace_bufs B = ctx_bufs(ctx);
  if (!ctx->graph_exec) {
    rc = upload_descriptors(ctx, B.desc);
    if (rc) return rc;
    rc = alloc_accumulators(ctx, adj);
    if (rc) return rc;
    kernel_cfg cfg;
    rc = pick_config(ctx, cfg, adj, fp64);
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_trace(ctx, B, cfg, adj, fp64, false, kernel, -1);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) return rc;
    HIP_TRY(ctx, e);
    ctx->graph = g;
    HIP_TRY(ctx, hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
  }
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (uint32_t t = 0; t < times; ++t) HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->trace_ms = times ? ms / times : 0.0;
  ctx->last_adjoint = adj;
  ctx->last_fp64 = fp64;
  ctx->last_count = 0;
  ctx->last_moments = ctx->params.moments != 0;
  HIP_TRY(ctx, hipMemcpy(ctx->run_stats, ctx->stats.p, sizeof ctx->run_stats, hipMemcpyDeviceToHost));
  ctx->have_results = false;
  if (times) ctx->acc_traced = true;
  return RTI_OK;
}

int rti_trace_noalloc(rti_ctx* ctx, unsigned long long* acc, uint64_t acc_words) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (ctx->inst.empty()) return fail(ctx, RTI_ERR_NO_GEOMETRY, "no geometry set");
  if (ctx->dirty)
    return fail(ctx, RTI_ERR_STATE, "the scene is not built for the current geometry: call rti_prepare first");
  if (!acc) return fail(ctx, RTI_ERR_INVALID_ARG, "null accumulator buffer");
  const uint64_t need = ctx->acc_total * rti_dev::kAccStride;
  if (acc_words < need)
    return fail(ctx, RTI_ERR_INVALID_ARG, "accumulator buffer holds %llu words, the scene needs %llu",
                (unsigned long long)acc_words, (unsigned long long)need);
  if (ctx->inst.size() > kNoAllocMaxInstances)
    return fail(ctx, RTI_ERR_INVALID_ARG, "%zu instances: the no-allocation trace takes at most %u", ctx->inst.size(),
                kNoAllocMaxInstances);
  if (ctx->params.adjoint || ctx->params.moments)
    return fail(ctx, RTI_ERR_INVALID_ARG,
                "adjoint and moments need accumulators the no-allocation trace does not take: use rti_trace");
  unsigned char* base = static_cast<unsigned char*>(scratch_ptr(ctx->device));
  if (!base) return fail(ctx, RTI_ERR_STATE, "static scratch not resolved for device %d", ctx->device);
  const trace_bufs B{acc, reinterpret_cast<inst_desc*>(base + kScratchDescOffset),
                     reinterpret_cast<unsigned long long*>(base),
                     reinterpret_cast<unsigned long long*>(base + kScratchStatsOffset)};
  const bool fp64 = ctx->params.accum == RTI_ACCUM_FP64;
  const bool count = ctx->params.kernel == RTI_KERNEL_COUNT;
  int rc = upload_descriptors(ctx, B.desc);
  if (rc) return rc;
  kernel_cfg cfg;
  rc = pick_config(ctx, cfg, false, fp64);
  if (rc) return rc;
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  rc = enqueue_trace(ctx, B, cfg, false, fp64, count, ctx->params.kernel == RTI_KERNEL_PLAIN ? 1 : 0, -1);
  if (rc) return rc;
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->trace_ms = ms;
  ctx->last_adjoint = 0;
  ctx->last_fp64 = fp64;
  ctx->last_count = count;
  ctx->last_moments = 0;
  HIP_TRY(ctx, hipMemcpy(ctx->run_stats, B.stats, sizeof ctx->run_stats, hipMemcpyDeviceToHost));
  // the context's getters and rti_finalize refer to its own accumulators,
  // which this trace did not write
  ctx->have_results = false;
  ctx->acc_traced = false;
  return RTI_OK;
}

int rti_batch_get_flux(rti_ctx* ctx, uint32_t i, float* smoothed_norm, double* raw, double* sigma) {
  if (!ctx) return RTI_ERR_INVALID_ARG;
  if (i >= ctx->inst.size()) return fail(ctx, RTI_ERR_INVALID_ARG, "instance index out of range");
  return get_flux_impl(ctx, ctx->inst[i], smoothed_norm, raw, sigma);
}

int rti_batch_get_counts(rti_ctx* ctx, uint32_t i, uint64_t* counts) {
  if (!ctx || !counts) return RTI_ERR_INVALID_ARG;
  if (i >= ctx->inst.size()) return fail(ctx, RTI_ERR_INVALID_ARG, "instance index out of range");
  if (!ctx->have_results) return fail(ctx, RTI_ERR_STATE, "no results: call rti_batch_run first");
  std::memcpy(counts, ctx->inst[i].counts.data(), ctx->inst[i].n * 8);
  return RTI_OK;
}

int rti_batch_get_stats(rti_ctx* ctx, rti_stats* out) {
  if (!ctx || !out) return RTI_ERR_INVALID_ARG;
  if (ctx->inst.empty()) return fail(ctx, RTI_ERR_NO_GEOMETRY, "no geometry");
  fill_stats(ctx, nullptr, out);
  return RTI_OK;
}

}  // extern "C"
