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

__global__ void k_finalize(const unsigned long long* s1, const unsigned long long* s2, const unsigned long long* cnt,
                           const uint32_t* perm, uint32_t pb, uint32_t n, double inv_scale, int fp64, double nrays,
                           double* raw_stored, double* raw_out, double* sigma_out, unsigned long long* cnt_out) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  size_t g = (size_t)pb + j;
  double v = acc_value(s1[g], inv_scale, fp64);
  raw_stored[j] = v;
  uint32_t i = perm[g];
  if (raw_out) raw_out[i] = v;
  if (sigma_out) {
    // standard deviation of the sum over nrays histories, from the
    // per-deposit second moment: var = S2 - S1^2 / N
    double q = acc_value(s2[g], inv_scale, fp64);
    double var = q - v * v / nrays;
    sigma_out[i] = var > 0.0 ? sqrt(var) : 0.0;
  }
  if (cnt_out) cnt_out[i] = cnt[g];
}

__global__ void k_finalize_signed(const unsigned long long* acc, const uint32_t* perm, uint32_t pb, uint32_t n,
                                  double inv_scale, double* out) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  size_t g = (size_t)pb + j;
  out[perm[g]] = (double)(long long)acc[g] * inv_scale;
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

hipError_t finalize_instance(const unsigned long long* s1, const unsigned long long* s2,
                             const unsigned long long* cnt, const uint32_t* perm, uint32_t prim_begin, uint32_t n,
                             double inv_scale, int accum_fp64, uint64_t n_rays, double* raw_stored, double* raw_out,
                             double* sigma_out, unsigned long long* cnt_out, hipStream_t stream) {
  k_finalize<<<nblk(n, 256), 256, 0, stream>>>(s1, s2, cnt, perm, prim_begin, n, inv_scale, accum_fp64,
                                               (double)(n_rays ? n_rays : 1), raw_stored, raw_out, sigma_out,
                                               cnt_out);
  return hipGetLastError();
}

hipError_t finalize_signed(const unsigned long long* acc, const uint32_t* perm, uint32_t prim_begin, uint32_t n,
                           double inv_scale, double* out, hipStream_t stream) {
  k_finalize_signed<<<nblk(n, 256), 256, 0, stream>>>(acc, perm, prim_begin, n, inv_scale, out);
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
