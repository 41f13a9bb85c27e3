// User plugin classes written against the reference's interfaces
// (src/rti/particle/i_particle.hpp:8-16, src/rti/geo/meta_geometry.hpp:8-25,
// src/rti/ray/i_direction.hpp:9-19), compiled against the drop-in.
// Usage: dropin_plugins <mode>
//   direction   a direction class with only get(): run() must refuse it
//               (the GPU samples cos^n directions only) before any device work
//   particle    a particle that reads the geometry through get_normal_ref and
//               print, then returns an invalid probability: run() must refuse
//               it before any device work; prints what the particle saw
//   gpu         the same particle returning 0.25: a full run on the GPU
#include <cstdio>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "rti/device.hpp"

using numeric_type = float;

class plain_direction : public rti::ray::i_direction<numeric_type> {
 public:
  rti::util::triple<numeric_type> get(rti::rng::i_rng&, rti::rng::i_rng::i_state&,
                                      rti::rng::i_rng::i_state&) const override {
    return {0, 0, -1};
  }
};

struct seen_t {
  bool called = false, ref_ok = false, ref_is_stored = false, print_ok = false;
  unsigned prim = 99;
};
static seen_t seen;
static numeric_type answer = 2.0f;

class probing_particle : public rti::particle::i_particle<numeric_type> {
 public:
  numeric_type get_sticking_probability(RTCRay&, RTCHit& hit, rti::geo::meta_geometry<numeric_type>& geometry,
                                        rti::rng::i_rng&, rti::rng::i_rng::i_state&) override final {
    seen.called = true;
    seen.prim = hit.primID;
    rti::util::triple<numeric_type>& r = geometry.get_normal_ref(hit.primID);
    const rti::util::triple<numeric_type> v = geometry.get_normal(hit.primID);
    seen.ref_ok = r == v && r[2] == 1.0f;
    seen.ref_is_stored = &geometry.get_normal_ref(hit.primID) == &r;
    std::ostringstream os;
    geometry.print(os);
    const std::string s = os.str();
    seen.print_ok = s.rfind("(:class", 0) == 0 && s.find("(0.5 0.25 0 0.1)") != std::string::npos;
    return answer;
  }
  void init_new() override final {}
};

static void fill(rti::device<numeric_type, probing_particle, rti::reflection::diffuse<numeric_type>>& dev) {
  std::vector<std::array<numeric_type, 3>> pts, nrm;
  std::vector<numeric_type> rad;
  for (int i = 0; i < 10; ++i)
    for (int j = 0; j < 10; ++j) {
      pts.push_back({0.5f + 0.1f * i, 0.25f + 0.1f * j, 0.0f});
      nrm.push_back({0.0f, 0.0f, 1.0f});
      rad.push_back(0.1f);
    }
  dev.set_points(pts);
  dev.set_normals(nrm);
  dev.set_grid_spacing(rad);
  dev.set_number_of_rays(20000);
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "";
  rti::device<numeric_type, probing_particle, rti::reflection::diffuse<numeric_type>> dev;
  fill(dev);
  if (mode == "direction") {
    plain_direction d;
    dev.set(d);
    try {
      dev.run();
    } catch (const std::invalid_argument& e) {
      const bool ok = std::strstr(e.what(), "cos_exponent") != nullptr && !seen.called;
      std::printf("%s: %s\n", ok ? "refused" : "unexpected", e.what());
      return ok ? 0 : 1;
    }
    std::printf("not refused\n");
    return 1;
  }
  if (mode == "particle") {
    try {
      dev.run();
    } catch (const std::invalid_argument& e) {
      std::printf("called %d prim %u ref_ok %d ref_is_stored %d print_ok %d: %s\n", seen.called, seen.prim,
                  seen.ref_ok, seen.ref_is_stored, seen.print_ok, e.what());
      const bool ok = seen.called && seen.prim == 0 && seen.ref_ok && seen.ref_is_stored && seen.print_ok &&
                      std::strstr(e.what(), "sticking probability") != nullptr;
      return ok ? 0 : 1;
    }
    return 1;
  }
  if (mode == "gpu") {
    answer = 0.25f;
    dev.run();
    size_t total = 0;
    for (size_t c : dev.get_hit_cnts()) total += c;
    std::printf("deposits %zu ref_ok %d print_ok %d\n", total, seen.ref_ok, seen.print_ok);
    return total > 0 && seen.ref_ok && seen.print_ok ? 0 : 1;
  }
  std::fprintf(stderr, "usage: %s direction|particle|gpu\n", argv[0]);
  return 2;
}
