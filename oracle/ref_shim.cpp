// C entry points over the reference's own Embree-free headers, compiled from
// the reference tree where it lies (oracle/Makefile target `ref`, output in
// oracle/_ref/). TEST INFRASTRUCTURE ONLY: used by CPU tests to pin the
// oracle restatement; nothing ships or runs this on the GPU path.
//
// Headers used (all under src/rti/ of the reference):
//   geo/disc_bounding_box_intersector.hpp  (exposed disc area)
//   geo/disc_neighborhood.hpp              (overlap graph, divide and conquer)
//   util/utils.hpp                         (orthonormal basis)
//   ray/cos_hemi.hpp                       (cosine lobe about a basis)
//   ray/power_cosine_direction_z.hpp       (cos^n source direction)
//   ray/rectangle_origin_z.hpp             (source origin)
//   mc/rejection_control.hpp               (Russian roulette)
//   trace/hit_accumulator.hpp              (S1..S4, relative error, VOV)
// The only code here is a scripted implementation of the reference's own
// i_rng interface (it returns caller-given integers) and the C wrappers.

#include <algorithm>
#include <cassert>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <map>
#include <vector>

#include "rti/util/utils.hpp"
#include "rti/rng/i_rng.hpp"
#include "rti/geo/disc_bounding_box_intersector.hpp"
#include "rti/geo/disc_neighborhood.hpp"
#include "rti/ray/cos_hemi.hpp"
#include "rti/ray/power_cosine_direction_z.hpp"
#include "rti/ray/rectangle_origin_z.hpp"
#include "rti/mc/rejection_control.hpp"
#include "rti/trace/hit_accumulator.hpp"

namespace {

// i_rng that replays a caller-provided sequence of 64-bit draws.
struct scripted_rng : public rti::rng::i_rng {
  struct state : public rti::rng::i_rng::i_state {
    const uint64_t* vals = nullptr;
    size_t n = 0, pos = 0;
    std::unique_ptr<rti::rng::i_rng::i_state> clone() const override { return std::make_unique<state>(*this); }
  };
  uint64_t get(rti::rng::i_rng::i_state& s) const override {
    auto& st = static_cast<state&>(s);
    return st.pos < st.n ? st.vals[st.pos++] : 0;
  }
  uint64_t min() const override { return 0; }
  uint64_t max() const override { return UINT64_MAX; }
};

}  // namespace

extern "C" {

float ref_area_inside(float xmin, float ymin, float xmax, float ymax, const float disc[4], const float nrm[3]) {
  rti::geo::disc_bounding_box_intersector dbbi(xmin, ymin, xmax, ymax);
  std::array<float, 4> d = {disc[0], disc[1], disc[2], disc[3]};
  std::array<float, 3> n = {nrm[0], nrm[1], nrm[2]};
  return dbbi.area_inside(d, n);
}

// Reference overlap graph; writes degrees (pass 1, col == null) or columns
// ascending per row (pass 2). Returns the number of directed edges.
int64_t ref_neighbors(const float* xyzr, uint64_t n, uint32_t* row_ptr, uint32_t* col) {
  std::vector<rti::util::quadruple<float>> pts(n);
  rti::util::triple<float> mn = {FLT_MAX, FLT_MAX, FLT_MAX}, mx = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (uint64_t i = 0; i < n; ++i) {
    pts[i] = {xyzr[4 * i], xyzr[4 * i + 1], xyzr[4 * i + 2], xyzr[4 * i + 3]};
    for (int k = 0; k < 3; ++k) {
      mn[k] = std::min(mn[k], pts[i][k]);
      mx[k] = std::max(mx[k], pts[i][k]);
    }
  }
  rti::geo::disc_neighborhood<float> nb;
  nb.setup_neighborhood(pts, mn, mx);
  row_ptr[0] = 0;
  for (uint64_t i = 0; i < n; ++i) row_ptr[i + 1] = row_ptr[i] + (uint32_t)nb.get_neighbors(i).size();
  if (col) {
    for (uint64_t i = 0; i < n; ++i) {
      auto v = nb.get_neighbors(i);
      std::sort(v.begin(), v.end());
      for (size_t q = 0; q < v.size(); ++q) col[row_ptr[i] + q] = (uint32_t)v[q];
    }
  }
  return row_ptr[n];
}

void ref_orthonormal_basis(const float nrm[3], float out[9]) {
  rti::util::triple<float> v = {nrm[0], nrm[1], nrm[2]};
  auto b = rti::util::get_orthonormal_basis(v);
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) out[3 * i + k] = b[i][k];
}

void ref_cos_hemi(const float nrm[3], const uint64_t draws[2], float out[3]) {
  rti::util::triple<float> v = {nrm[0], nrm[1], nrm[2]};
  auto b = rti::util::get_orthonormal_basis(v);
  scripted_rng rng;
  scripted_rng::state st;
  st.vals = draws;
  st.n = 2;
  auto d = rti::ray::cos_hemi::get<float>(b, rng, st);
  for (int k = 0; k < 3; ++k) out[k] = d[k];
}

void ref_power_cosine_direction(float exponent, const uint64_t draws[2], float out[3]) {
  rti::ray::power_cosine_direction_z<float> dir(exponent);
  scripted_rng rng;
  scripted_rng::state s1, s2;
  s1.vals = draws;
  s1.n = 1;
  s2.vals = draws + 1;
  s2.n = 1;
  auto d = dir.get(rng, s1, s2);
  for (int k = 0; k < 3; ++k) out[k] = d[k];
}

void ref_rectangle_origin(float z, float x1, float y1, float x2, float y2, const uint64_t draws[2], float out[3]) {
  rti::ray::rectangle_origin_z<float> org(z, {x1, y1}, {x2, y2});
  scripted_rng rng;
  scripted_rng::state s1, s2;
  s1.vals = draws;
  s1.n = 1;
  s2.vals = draws + 1;
  s2.n = 1;
  auto p = org.get(rng, s1, s2);
  for (int k = 0; k < 3; ++k) out[k] = p[k];
}

// Returns 1 (continue) or 0 (kill); *weight updated as the reference does.
int ref_rejection_control(float* weight, uint64_t draw) {
  scripted_rng rng;
  scripted_rng::state st;
  st.vals = &draw;
  st.n = 1;
  float init = 1.0f;
  return rti::mc::rejection_control<float>::check_weight_reweight_or_kill(*weight, init, rng, st) ? 1 : 0;
}

// Relative error and variance of variance of a sequence of deposits on
// n_prims elements (ids[k], values[k]).
void ref_hit_accumulator(uint64_t n_prims, const uint32_t* ids, const float* values, uint64_t n,
                         double* s1, uint64_t* cnt, double* relerr, double* vov) {
  rti::trace::hit_accumulator<float> acc(n_prims);
  for (uint64_t k = 0; k < n; ++k) acc.use(ids[k], values[k]);
  auto v = acc.get_values();
  auto c = acc.get_cnts();
  auto re = acc.get_relative_error();
  auto vv = acc.get_vov();
  for (uint64_t i = 0; i < n_prims; ++i) {
    s1[i] = v[i];
    cnt[i] = c[i];
    relerr[i] = re[i];
    vov[i] = vv[i];
  }
}

}  // extern "C"
