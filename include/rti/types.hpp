// Plugin-interface types of the rti::device drop-in (own definitions; no
// Embree, no VTK). Names, namespaces and virtual signatures follow the
// reference so that user particle / reflection / direction classes written
// against it compile unchanged:
//   rti::geo::bound_condition            src/rti/geo/bound_condition.hpp:4-6
//   rti::rng::i_rng (+ i_state)          src/rti/rng/i_rng.hpp:11-30
//   rti::rng::mt64_rng                   src/rti/rng/mt64_rng.hpp:8-50
//   rti::geo::meta_geometry              src/rti/geo/meta_geometry.hpp:8-25 (no Embree handles)
//   rti::particle::i_particle            src/rti/particle/i_particle.hpp:8-16
//   rti::reflection::i_reflection        src/rti/reflection/i_reflection.hpp:7-19
//   rti::reflection::diffuse / specular  src/rti/reflection/diffuse.hpp, specular.hpp
//   rti::ray::i_direction                src/rti/ray/i_direction.hpp:8-20
//   rti::ray::cosine_direction_z         src/rti/ray/cosine_direction_z.hpp:13-27
//   rti::ray::power_cosine_direction_z   src/rti/ray/power_cosine_direction_z.hpp:17-32
//   rti::util::pair / triple / quadruple src/rti/util/utils.hpp:17-24
// RTCRay / RTCHit are layout-compatible stand-ins for Embree 3's public
// structs when Embree is not included (the interfaces take them by
// reference). The GPU path never calls these virtuals per hit: constant
// sticking probabilities and the reflection kind are read once on the host.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <ostream>
#include <random>
#include <vector>

#if !defined(RTC_VERSION_MAJOR)
struct RTCRay {
  float org_x, org_y, org_z, tnear;
  float dir_x, dir_y, dir_z, time;
  float tfar;
  unsigned int mask, id, flags;
};
struct RTCHit {
  float Ng_x, Ng_y, Ng_z;
  float u, v;
  unsigned int primID, geomID;
  unsigned int instID[1];
};
#endif

namespace rti {
namespace util {
template <typename T>
using pair = std::array<T, 2>;
template <typename T>
using triple = std::array<T, 3>;
template <typename T>
using quadruple = std::array<T, 4>;
}  // namespace util

namespace geo {
enum class bound_condition { REFLECTIVE, PERIODIC };

// The geometry a particle or reflection sees (meta_geometry.hpp:8-25). The
// reference's get_rtc_device / get_rtc_geometry return Embree handles and
// have no counterpart without Embree; print, get_normal, get_normal_ref and
// get_new_origin keep their signatures.
template <typename T>
class meta_geometry {
 public:
  virtual ~meta_geometry() {}
  virtual void print(std::ostream& os) = 0;
  virtual util::triple<T> get_normal(unsigned int primID) = 0;
  virtual util::triple<T>& get_normal_ref(unsigned int primID) = 0;
  virtual util::triple<T> get_new_origin(RTCRay& r, unsigned int) {
    return {(T)(r.org_x + r.dir_x * r.tfar), (T)(r.org_y + r.dir_y * r.tfar), (T)(r.org_z + r.dir_z * r.tfar)};
  }
};
}  // namespace geo

namespace rng {
class i_rng {
 public:
  virtual ~i_rng() {}
  struct i_state {
    virtual ~i_state() {}
    virtual std::unique_ptr<i_state> clone() const = 0;
  };
  virtual uint64_t get(i_state& s) const = 0;
  virtual uint64_t min() const = 0;
  virtual uint64_t max() const = 0;
};

class mt64_rng : public i_rng {
 public:
  struct state : public i_rng::i_state {
    state() : mMT(std::mt19937_64::default_seed) {}
    explicit state(unsigned int seed) : mMT(seed) {}
    std::unique_ptr<i_rng::i_state> clone() const override { return std::make_unique<state>(*this); }
    std::mt19937_64 mMT;
  };
  uint64_t get(i_rng::i_state& s) const override { return static_cast<state&>(s).mMT(); }
  uint64_t min() const override { return std::mt19937_64::min(); }
  uint64_t max() const override { return std::mt19937_64::max(); }
};
}  // namespace rng

namespace particle {
template <typename T>
class i_particle {
 public:
  virtual ~i_particle() {}
  // Called once per geometry on the host by the drop-in; the value is used
  // as the sticking probability of every hit on the GPU.
  virtual T get_sticking_probability(RTCRay& rayin, RTCHit& hitin, geo::meta_geometry<T>& geometry, rng::i_rng& rng,
                                     rng::i_rng::i_state& rngstate) = 0;
  virtual void init_new() = 0;
};
}  // namespace particle

namespace reflection {
template <typename T>
class i_reflection {
 public:
  virtual ~i_reflection() {}
  virtual util::pair<util::triple<T>> use(RTCRay& rayin, RTCHit& hitin, geo::meta_geometry<T>& geometry,
                                          rng::i_rng& rng, rng::i_rng::i_state& rngstate) = 0;
};

// Lambertian reflection about the surface normal (cos_hemi::get).
template <typename T>
class diffuse : public i_reflection<T> {
 public:
  util::pair<util::triple<T>> use(RTCRay& r, RTCHit& h, geo::meta_geometry<T>& g, rng::i_rng& rng,
                                  rng::i_rng::i_state& st) override final {
    auto n = g.get_normal(h.primID);
    T r1 = (T)rng.get(st) / ((T)rng.max() + 1), r2 = (T)rng.get(st) / ((T)rng.max() + 1);
    T sign = std::copysign((T)1, n[2]);
    T a = (T)-1 / (sign + n[2]);
    T b = n[0] * n[1] * a;
    util::triple<T> b1 = {(T)1 + sign * n[0] * n[0] * a, sign * b, -sign * n[0]};
    util::triple<T> b2 = {b, sign + n[1] * n[1] * a, -n[1]};
    T c1 = std::sqrt(r2), st2 = std::sqrt((T)1 - r2);
    T c2 = std::cos((T)(2 * M_PI) * r1) * st2, c3 = std::sin((T)(2 * M_PI) * r1) * st2;
    util::triple<T> d = {n[0] * c1 + b1[0] * c2 + b2[0] * c3, n[1] * c1 + b1[1] * c2 + b2[1] * c3,
                         n[2] * c1 + b1[2] * c2 + b2[2] * c3};
    return {g.get_new_origin(r, h.primID), d};
  }
};

// Mirror reflection (specular::use).
template <typename T>
class specular : public i_reflection<T> {
 public:
  util::pair<util::triple<T>> use(RTCRay& r, RTCHit& h, geo::meta_geometry<T>& g, rng::i_rng&,
                                  rng::i_rng::i_state&) override final {
    auto n = g.get_normal(h.primID);
    T dn = r.dir_x * n[0] + r.dir_y * n[1] + r.dir_z * n[2];
    util::triple<T> d = {r.dir_x - 2 * dn * n[0], r.dir_y - 2 * dn * n[1], r.dir_z - 2 * dn * n[2]};
    return {g.get_new_origin(r, h.primID), d};
  }
};
}  // namespace reflection

namespace ray {
template <typename T>
class i_direction {
 public:
  virtual ~i_direction() {}
  virtual util::triple<T> get(rng::i_rng&, rng::i_rng::i_state&, rng::i_rng::i_state&) const = 0;
  // Exponent n of the cos^n direction density about -z, which the GPU
  // samples itself. Direction classes that do not override it (NaN) are not
  // supported by the drop-in: run() reports an error for them.
  virtual T cos_exponent() const { return std::numeric_limits<T>::quiet_NaN(); }
};

template <typename T>
class power_cosine_direction_z : public i_direction<T> {
 public:
  explicit power_cosine_direction_z(T exponent) : exp(exponent) {}
  util::triple<T> get(rng::i_rng& rng, rng::i_rng::i_state& s1, rng::i_rng::i_state& s2) const override {
    T r1 = (T)rng.get(s1) / ((T)rng.max() + 1), r2 = (T)rng.get(s2) / ((T)rng.max() + 1);
    T tt = std::pow(r2, (T)2 / (exp + 1));
    util::triple<T> d = {std::cos((T)(2 * M_PI) * r1) * std::sqrt(1 - tt),
                         std::sin((T)(2 * M_PI) * r1) * std::sqrt(1 - tt), -std::sqrt(tt)};
    T l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    return {d[0] / l, d[1] / l, d[2] / l};
  }
  T cos_exponent() const override { return exp; }

 private:
  T exp;
};

template <typename T>
class cosine_direction_z : public power_cosine_direction_z<T> {
 public:
  cosine_direction_z() : power_cosine_direction_z<T>((T)1) {}
};
}  // namespace ray
}  // namespace rti
