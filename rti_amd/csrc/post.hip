// Accumulator clearing and finalisation for gfx950.
//
// fill_u64 clears accumulators with a kernel (no memset node, so the batch
// graph can be replayed). finalize_instance converts the 64-bit fixed-point
// sums to FP64 (the only FP64 arithmetic of the path) and scatters them to
// input order; smooth_normalize restates
// device::extract_mc_estimates_normalized_smoothed (reference
// src/rti/device.hpp:153-187): v_i = value_i / area_i, mean over i and its
// neighbours, divided by the maximum.

#include <hip/hip_runtime.h>

#include <cfloat>

#include "internal.hpp"

namespace rti_impl {
namespace {

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__device__ __forceinline__ double acc_value(unsigned long long x, double inv_scale, int fp64) {
  return fp64 ? __longlong_as_double((long long)x) : (double)(long long)x * inv_scale;
}

// Primitive j of an instance: stored position pb + j (perm, areas), accumulator
// slot ab + j (acc, adj, mom).
__global__ void k_finalize(const unsigned long long* acc, uint32_t ab, const uint32_t* perm, uint32_t pb, uint32_t n,
                           double inv_scale, double inv_scale2, int fp64, double nrays, double* raw_stored,
                           double* raw_out, double* sigma_out, unsigned long long* cnt_out) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  size_t g = (size_t)pb + j;
  const unsigned long long* a = acc + (size_t)rti_dev::kAccStride * ((size_t)ab + j);
  double v = acc_value(a[0], inv_scale, fp64);
  raw_stored[j] = v;
  uint32_t i = perm[g];
  if (raw_out) raw_out[i] = v;
  if (sigma_out) {
    // standard deviation of the sum over nrays histories, from the
    // per-deposit second moment: var = S2 - S1^2 / N
    double q = acc_value(a[1], inv_scale2, fp64);
    double var = q - v * v / nrays;
    sigma_out[i] = var > 0.0 ? sqrt(var) : 0.0;
  }
  if (cnt_out) cnt_out[i] = a[2];
}

// Relative error and variance of the variance per primitive, input order
// (hit_accumulator::get_relative_error / get_vov,
// src/rti/trace/hit_accumulator.hpp:152-249), with N = all deposits of the
// instance:
//   rel = sqrt(S2 / S1^2 - 1/N)                        (DBL_MAX where S1 = 0)
//   vov = (S4 - 4 S1 S3/N + 8 S2 S1^2/N^2 - 4 S1^4/N^3 - S2^2/N) N^2 / (N S2 - S1^2)^2
//                                                      (DBL_MAX where the count or
//                                                       the denominator is 0)
// The powers of N are formed in double.
__global__ void k_error_stats(const unsigned long long* acc, const double* mom, uint32_t ab, const uint32_t* perm,
                              uint32_t pb, uint32_t n, double inv_scale, double inv_scale2, int fp64, double nn,
                              double* rel, double* vov) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const size_t g = (size_t)pb + j;
  const size_t sl = (size_t)ab + j;
  const uint32_t i = perm[g];
  const double big = DBL_MAX;
  const unsigned long long* a = acc + (size_t)rti_dev::kAccStride * sl;
  const double a1 = acc_value(a[0], inv_scale, fp64);
  const double a2 = acc_value(a[1], inv_scale2, fp64);
  const double sq = a1 * a1;
  if (rel) rel[i] = (nn == 0.0 || sq == 0.0) ? big : sqrt(a2 / sq - 1.0 / nn);
  if (vov) {
    if (a[2] == 0) {
      vov[i] = big;
      return;
    }
    const double a3 = mom[2 * sl], a4 = mom[2 * sl + 1];
    const double q4 = sq * sq, a2sq = a2 * a2;
    const double num = a4 - (4 * a1 * a3 / nn) + (8 * a2 * sq / (nn * nn)) - (4 * q4 / (nn * nn * nn)) - (a2sq / nn);
    const double r2 = nn * a2 - sq;
    const double den2 = r2 * r2;
    vov[i] = den2 == 0.0 ? big : num / den2 * nn * nn;
  }
}

__global__ void k_finalize_signed(const unsigned long long* adj, int k, uint32_t ab, const uint32_t* perm, uint32_t pb,
                                  uint32_t n, double inv_scale, double* out) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  size_t g = (size_t)pb + j;
  out[perm[g]] = (double)(long long)adj[2 * ((size_t)ab + j) + k] * inv_scale;
}

__global__ void k_smooth(const double* raw, const float* areas, const uint32_t* nb_ptr, const uint32_t* nb_col,
                         uint32_t pb, uint32_t n, double* v, unsigned long long* maxbits) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  double vv = 0.0;
  if (j < n) {
    size_t g = (size_t)pb + j;
    vv = raw[j] / (double)areas[g];
    uint32_t b = nb_ptr ? nb_ptr[g] : 0u, e = nb_ptr ? nb_ptr[g + 1] : 0u;
    for (uint32_t q = b; q < e; ++q) {
      uint32_t k = nb_col[q] - pb;
      vv += raw[k] / (double)areas[(size_t)pb + k];
    }
    vv /= (double)(e - b + 1);
    v[j] = vv;
  }
  // max over non-negative doubles: their bit patterns order like integers
  double m = vv > 0.0 ? vv : 0.0;
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(m));
}

__global__ void k_normalize(const double* v, const uint32_t* perm, uint32_t pb, uint32_t n,
                            const unsigned long long* maxbits, float* out) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double m = __longlong_as_double((long long)*maxbits);
  out[perm[(size_t)pb + j]] = (float)(v[j] / m);
}

__global__ void k_permute_f32(const float* stored, const uint32_t* perm, uint32_t n, float* out) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  out[perm[j]] = stored[j];
}

inline unsigned nblk(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

hipError_t fill_u64(unsigned long long* p, uint64_t n, unsigned long long v, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  unsigned blocks = nblk(n, 256);
  if (blocks > 4096) blocks = 4096;
  k_fill_u64<<<blocks, 256, 0, stream>>>(p, n, v);
  return hipGetLastError();
}

hipError_t finalize_instance(const unsigned long long* acc, uint32_t acc_begin, const uint32_t* perm,
                             uint32_t prim_begin, uint32_t n, double inv_scale, double inv_scale2, int accum_fp64,
                             uint64_t n_rays, double* raw_stored, double* raw_out, double* sigma_out,
                             unsigned long long* cnt_out, hipStream_t stream) {
  k_finalize<<<nblk(n, 256), 256, 0, stream>>>(acc, acc_begin, perm, prim_begin, n, inv_scale, inv_scale2, accum_fp64,
                                               (double)(n_rays ? n_rays : 1), raw_stored, raw_out, sigma_out,
                                               cnt_out);
  return hipGetLastError();
}

hipError_t error_stats(const unsigned long long* acc, const double* mom, uint32_t acc_begin, const uint32_t* perm,
                       uint32_t prim_begin, uint32_t n, double inv_scale, double inv_scale2, int accum_fp64,
                       uint64_t total, double* rel_out, double* vov_out, hipStream_t stream) {
  k_error_stats<<<nblk(n, 256), 256, 0, stream>>>(acc, mom, acc_begin, perm, prim_begin, n, inv_scale, inv_scale2,
                                                  accum_fp64, (double)total, rel_out, vov_out);
  return hipGetLastError();
}

hipError_t finalize_signed(const unsigned long long* adj, int k, uint32_t acc_begin, const uint32_t* perm,
                           uint32_t prim_begin, uint32_t n, double inv_scale, double* out, hipStream_t stream) {
  k_finalize_signed<<<nblk(n, 256), 256, 0, stream>>>(adj, k, acc_begin, perm, prim_begin, n, inv_scale, out);
  return hipGetLastError();
}

hipError_t smooth_normalize(const double* raw_stored, const float* areas_stored, const uint32_t* perm,
                            const uint32_t* nb_ptr, const uint32_t* nb_col, uint32_t prim_begin, uint32_t n,
                            double* vtmp, double* maxbuf, float* out, hipStream_t stream) {
  hipError_t e = fill_u64(reinterpret_cast<unsigned long long*>(maxbuf), 1, 0ull, stream);
  if (e != hipSuccess) return e;
  k_smooth<<<nblk(n, 256), 256, 0, stream>>>(raw_stored, areas_stored, nb_ptr, nb_col, prim_begin, n, vtmp,
                                             reinterpret_cast<unsigned long long*>(maxbuf));
  k_normalize<<<nblk(n, 256), 256, 0, stream>>>(vtmp, perm, prim_begin, n,
                                                reinterpret_cast<const unsigned long long*>(maxbuf), out);
  return hipGetLastError();
}

hipError_t permute_f32(const float* stored, const uint32_t* perm, uint32_t n, float* out, hipStream_t stream) {
  k_permute_f32<<<nblk(n, 256), 256, 0, stream>>>(stored, perm, n, out);
  return hipGetLastError();
}

}  // namespace rti_impl
