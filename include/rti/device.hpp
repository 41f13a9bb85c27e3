// rti::device -- drop-in for the reference's library facade
// (src/rti/device.hpp:34-297) on an MI355X: the same class template, setters
// and getters; run() calls the gfx950 tracer through the C ABI in rti_c.h
// (link with librti_hip.so).
//
// Kept from the reference:
//   * setters copy their arguments, getters return copies (device.hpp:47-117);
//   * set_grid_spacing gives the disc radii verbatim (device.hpp:206-222);
//   * defaults: 1024 rays, reflective x and y, cosine source
//     (device.hpp:289-296);
//   * get_mc_estimates: value / area, averaged over the disc and its
//     neighbours, divided by the maximum (device.hpp:153-187);
//     get_hit_cnts: deposits per disc including neighbour deposits.
// Decided differently (each is a documented divergence):
//   * set(i_direction&) is honoured (cos^n exponent from
//     direction.cos_exponent()); in the reference it assigns through a base
//     reference and leaves the cosine source in place (device.hpp:77-80, 296);
//   * particle_type::get_sticking_probability is evaluated once on the host
//     and used for every hit (exact for constant-sticking particles, the only
//     kind in the reference tree: main/dev.cpp:215-228, main/test2.cpp:6-20);
//   * reflection_type must be (derived from) rti::reflection::diffuse or
//     ::specular: the GPU implements those two laws;
//   * errors are reported by exceptions (std::runtime_error) instead of
//     asserts.
// Extension (no public counterpart in the reference):
//   * set_circular_source(true) starts the primary rays on the disc of
//     device::create_circular_source_from_bounding_box (device.hpp:269-280,
//     ray/disc_origin_z.hpp:17-33) instead of the rectangle run() uses.
#pragma once

#include <array>
#include <cmath>
#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "rti_c.h"
#include "types.hpp"

namespace rti {
using ::rti::geo::bound_condition;
using ::rti::particle::i_particle;
using ::rti::reflection::i_reflection;

template <typename numeric_type, typename particle_type, typename reflection_type>
class device final {
  static_assert(std::is_base_of<i_particle<numeric_type>, particle_type>::value, "Precondition");
  static_assert(std::is_base_of<i_reflection<numeric_type>, reflection_type>::value, "Precondition");
  static_assert(std::is_base_of<reflection::diffuse<numeric_type>, reflection_type>::value ||
                    std::is_base_of<reflection::specular<numeric_type>, reflection_type>::value,
                "the GPU tracer implements diffuse and specular reflection");

 public:
  device() = default;
  ~device() {
    if (ctx) rti_ctx_destroy(ctx);
  }
  device(const device&) = delete;
  device& operator=(const device&) = delete;

  void set_points(std::vector<std::array<numeric_type, 3>> points_) { points = points_; }
  void set_normals(std::vector<std::array<numeric_type, 3>> normals_) { normals = normals_; }
  void set_grid_spacing(std::vector<numeric_type> spacing_) { spacing = spacing_; }
  void set_number_of_rays(size_t numofrays_) { numofrays = numofrays_; }
  void set_x(bound_condition cond) { xCond = cond; }
  void set_y(bound_condition cond) { yCond = cond; }
  void set(ray::i_direction<numeric_type>& srcDirection) { cosExponent = srcDirection.cos_exponent(); }
  void set_circular_source(bool circular) { circularSource = circular; }

  void run() {
    if (points.size() != normals.size() || points.size() != spacing.size())
      throw std::invalid_argument("rti::device: points, normals and grid spacing differ in length");
    if (points.empty()) throw std::invalid_argument("rti::device: no points");
    if (std::isnan((double)cosExponent))
      throw std::invalid_argument(
          "rti::device: the source direction is not a cos^n distribution about -z (a direction class must "
          "override cos_exponent())");
    // host-side plugin evaluation first: no device work for an invalid particle
    const numeric_type sticking = sticking_probability();
    if (!ctx) check(rti_ctx_create(&ctx, 0), "rti_ctx_create");
    const size_t n = points.size();
    std::vector<float> xyzr(4 * n), nrm(3 * n);
    for (size_t i = 0; i < n; ++i) {
      for (int k = 0; k < 3; ++k) {
        xyzr[4 * i + k] = (float)points[i][k];
        nrm[3 * i + k] = (float)normals[i][k];
      }
      xyzr[4 * i + 3] = (float)spacing[i];
    }
    check(rti_set_discs(ctx, xyzr.data(), nrm.data(), n), "rti_set_discs");
    rti_params p;
    rti_default_params(&p);
    p.n_rays = numofrays;
    p.bc_x = xCond == bound_condition::PERIODIC ? RTI_BC_PERIODIC : RTI_BC_REFLECTIVE;
    p.bc_y = yCond == bound_condition::PERIODIC ? RTI_BC_PERIODIC : RTI_BC_REFLECTIVE;
    p.cos_exponent = (float)cosExponent;
    p.source = circularSource ? RTI_SOURCE_DISC : RTI_SOURCE_RECTANGLE;
    p.sticking = (float)sticking;
    p.diffuse_prob = std::is_base_of<reflection::diffuse<numeric_type>, reflection_type>::value ? 1.0f : 0.0f;
    check(rti_set_params(ctx, &p), "rti_set_params");
    check(rti_run(ctx), "rti_run");
    std::vector<float> est(n);
    std::vector<uint64_t> cnt(n);
    check(rti_get_flux(ctx, est.data(), nullptr, nullptr), "rti_get_flux");
    check(rti_get_counts(ctx, cnt.data()), "rti_get_counts");
    mcestimates.assign(est.begin(), est.end());
    hitcnts.assign(cnt.begin(), cnt.end());
  }

  std::vector<numeric_type> get_mc_estimates() { return mcestimates; }
  std::vector<size_t> get_hit_cnts() { return hitcnts; }

 private:
  // The disc cloud as the particle sees it (point_cloud_disc_geometry's
  // meta_geometry view, src/rti/geo/point_cloud_disc_geometry.hpp:40-101):
  // normals by primitive id, a reference into the stored normals, and print
  // in the reference's "(:class ...)" form with one "(x y z r)" per disc.
  struct disc_cloud_geometry : geo::meta_geometry<numeric_type> {
    const std::vector<util::triple<numeric_type>>* pts = nullptr;
    std::vector<util::triple<numeric_type>>* nrm = nullptr;
    const std::vector<numeric_type>* rad = nullptr;
    void print(std::ostream& os) override {
      os << "(:class rti::device::disc_cloud_geometry";
      for (size_t i = 0; i < pts->size(); ++i)
        os << " (" << (*pts)[i][0] << " " << (*pts)[i][1] << " " << (*pts)[i][2] << " " << (*rad)[i] << ")";
      os << ")";
    }
    util::triple<numeric_type> get_normal(unsigned int primID) override { return nrm->at(primID); }
    util::triple<numeric_type>& get_normal_ref(unsigned int primID) override { return nrm->at(primID); }
  };

  // The particle's sticking probability, queried once with a hit at the
  // first disc (user particles see a valid ray, hit and geometry).
  numeric_type sticking_probability() {
    disc_cloud_geometry geom;
    geom.pts = &points;
    geom.nrm = &normals;
    geom.rad = &spacing;
    RTCRay ray{};
    ray.org_x = (float)points[0][0];
    ray.org_y = (float)points[0][1];
    ray.org_z = (float)points[0][2] + 1.0f;
    ray.dir_z = -1.0f;
    ray.tnear = 1e-4f;
    ray.tfar = 1.0f;
    RTCHit hit{};
    hit.Ng_x = (float)normals[0][0];
    hit.Ng_y = (float)normals[0][1];
    hit.Ng_z = (float)normals[0][2];
    hit.primID = 0;
    hit.geomID = 1;
    rng::mt64_rng rng;
    rng::mt64_rng::state st(31);
    particle_type particle{};
    particle.init_new();
    numeric_type s = particle.get_sticking_probability(ray, hit, geom, rng, st);
    if (!(s > 0 && s <= 1)) throw std::invalid_argument("rti::device: sticking probability must be in (0, 1]");
    return s;
  }

  void check(int rc, const char* what) {
    if (rc != RTI_OK)
      throw std::runtime_error(std::string("rti::device: ") + what + " failed: " + rti_last_error(ctx));
  }

  std::vector<util::triple<numeric_type>> points;
  std::vector<util::triple<numeric_type>> normals;
  std::vector<numeric_type> spacing;
  std::vector<numeric_type> mcestimates;
  std::vector<size_t> hitcnts;
  size_t numofrays = 1024;
  bound_condition xCond = geo::bound_condition::REFLECTIVE;
  bound_condition yCond = geo::bound_condition::REFLECTIVE;
  numeric_type cosExponent = 1;
  bool circularSource = false;
  rti_ctx* ctx = nullptr;
};
}  // namespace rti
