// rti_cli: the reference's command-line tracer (src/rti/main/dev.cpp:51-288)
// on the MI355X library: a .vtp oriented-disc point cloud in, the flux trace
// with periodic x and y boundaries, a cosine source on the top of the
// bounding box and diffuse reflection, and a .vtp out with the cell data
// value (summed deposits), hitcnt, exposed-area, relative-error and
// variance-of-variance, string metadata, and the boundary box as a second
// .vtp file.
//
// Options (dev.cpp:51-72): -i/--infile (required), -o/--outfile,
// -r/--number-of-rays/--n-rays (default 128 * 1024), -s/--sticking-coefficient/
// --sticking-c/--sticking (default 0.8; reset to 1 outside (0, 1]),
// -m/--max-threads (accepted; the trace runs on the GPU), --discs,
// --triangles (rejected, as in the reference), --single-hit/--single
// (accepted, no effect, as in the reference). Extensions: --seed, --device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "rti/rti_c.h"
#include "vtkxml.hpp"

namespace {

struct options {
  std::string infile, outfile, rays, sticking, threads;
  bool triangles = false, discs = false, single_hit = false;
  uint32_t seed = 1;
  int device = 0;
};

const char* kUsage =
    "usage: rti_cli --infile|-i FILE [--outfile|-o FILE] [--number-of-rays|--n-rays|-r N]\n"
    "               [--sticking-coefficient|--sticking-c|--sticking|-s S] [--max-threads|-m N]\n"
    "               [--discs] [--triangles] [--single-hit|--single] [--seed N] [--device N]\n";

bool parse(int argc, char** argv, options& o) {
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto value = [&](std::string& dst) {
      if (i + 1 >= argc) return false;
      dst = argv[++i];
      return true;
    };
    if (a == "--infile" || a == "-i") { if (!value(o.infile)) return false; }
    else if (a == "--outfile" || a == "-o") { if (!value(o.outfile)) return false; }
    else if (a == "--number-of-rays" || a == "--n-rays" || a == "-r") { if (!value(o.rays)) return false; }
    else if (a == "--sticking-coefficient" || a == "--sticking-c" || a == "--sticking" || a == "-s") {
      if (!value(o.sticking)) return false;
    }
    else if (a == "--max-threads" || a == "-m") { if (!value(o.threads)) return false; }
    else if (a == "--triangles") o.triangles = true;
    else if (a == "--discs") o.discs = true;
    else if (a == "--single-hit" || a == "--single") o.single_hit = true;
    else if (a == "--seed") { std::string v; if (!value(v)) return false; o.seed = (uint32_t)std::stoul(v); }
    else if (a == "--device") { std::string v; if (!value(v)) return false; o.device = std::stoi(v); }
    else return false;
  }
  return !o.infile.empty();
}

std::string extension(const std::string& f) {
  size_t s = f.find_last_of("/\\");
  size_t d = f.find_last_of('.');
  if (d == std::string::npos || (s != std::string::npos && d < s)) return "";
  return f.substr(d);
}

std::string git_hash() {
  std::string out;
  std::unique_ptr<FILE, decltype(&pclose)> p(popen("git rev-parse HEAD 2>/dev/null", "r"), pclose);
  if (!p) return out;
  char buf[128];
  while (fgets(buf, sizeof buf, p.get())) out += buf;
  return out;
}

void check(int rc, rti_ctx* ctx, const char* what) {
  if (rc == RTI_OK) return;
  std::cerr << what << " failed: " << (ctx ? rti_last_error(ctx) : "") << " (" << rc << ")" << std::endl;
  std::exit(EXIT_FAILURE);
}

}  // namespace

int main(int argc, char** argv) {
  options o;
  if (!parse(argc, argv, o)) {
    std::cout << kUsage;
    return EXIT_FAILURE;
  }
  float sticking = 0.8f;  // dev.cpp:159-169
  try {
    if (!o.sticking.empty()) sticking = (float)std::stod(o.sticking);
  } catch (...) {
  }
  if (!(0 < sticking && sticking <= 1)) {
    std::cout << "Warning: sticking coefficient has been reset to sane value." << std::endl;
    sticking = 1.0f;
  }
  std::cout << "sticking coefficient == " << sticking << std::endl;
  if (extension(o.infile) == ".vtk") {  // dev.cpp:172-180
    std::cout << "Recognized \".vtk\" input file file extension. Please convert the input file to vtp (e.g., "
                 "using Paraview) and rerun "
              << argv[0] << " with the new file." << std::endl
              << "terminating" << std::endl;
    return EXIT_FAILURE;
  }
  if (o.triangles) {
    std::cerr << "Triangles in this version of " << argv[0] << " not supported " << std::endl;
    return EXIT_FAILURE;
  }
  uint64_t nrays = 128 * 1024ull;  // dev.cpp:217-222
  try {
    if (!o.rays.empty()) nrays = std::stoull(o.rays);
  } catch (...) {
  }

  rti_vtk::disc_cloud dc;
  try {
    dc = rti_vtk::discs_from(rti_vtk::read_vtp(o.infile));
  } catch (const std::exception& e) {
    std::cerr << "reading " << o.infile << ": " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  for (auto& w : dc.warnings) std::cerr << "Warning: " << w << ". File: " << o.infile << std::endl;
  const uint64_t n = dc.xyzr.size() / 4;
  // bounding box of the centres, top padded by the largest radius
  // (point_cloud_disc_geometry.hpp:83-87, 195-206); the source sits on its top
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f}, rmax = 0;
  for (uint64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], dc.xyzr[4 * i + k]);
      hi[k] = std::max(hi[k], dc.xyzr[4 * i + k]);
    }
    rmax = std::max(rmax, dc.xyzr[4 * i + 3]);
  }
  const float ztop = hi[2] + rmax;

  // child processes start before the GPU is initialised
  const std::string hash = o.outfile.empty() ? std::string() : git_hash();
  rti_ctx* ctx = nullptr;
  check(rti_ctx_create(&ctx, o.device), nullptr, "rti_ctx_create");
  check(rti_set_discs(ctx, dc.xyzr.data(), dc.nrm.data(), n), ctx, "rti_set_discs");
  rti_params p;
  rti_default_params(&p);
  p.n_rays = nrays;
  p.seed = o.seed;
  p.sticking = sticking;
  p.bc_x = RTI_BC_PERIODIC;  // dev.cpp:190-191
  p.bc_y = RTI_BC_PERIODIC;
  p.cos_exponent = 1.0f;     // cosine_direction_z (dev.cpp:212)
  p.diffuse_prob = 1.0f;     // reflection::diffuse (dev.cpp:245)
  p.src_z = ztop;
  p.moments = 1;             // for the variance of the variance
  check(rti_set_params(ctx, &p), ctx, "rti_set_params");
  const auto t0 = std::chrono::steady_clock::now();
  check(rti_run(ctx), ctx, "rti_run");
  const auto t1 = std::chrono::steady_clock::now();
  const long long ns = std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();

  std::vector<double> raw(n), rel(n), vov(n);
  std::vector<uint64_t> cnt(n);
  std::vector<float> area(n);
  check(rti_get_flux(ctx, nullptr, raw.data(), nullptr), ctx, "rti_get_flux");
  check(rti_get_counts(ctx, cnt.data()), ctx, "rti_get_counts");
  check(rti_get_areas(ctx, area.data()), ctx, "rti_get_areas");
  check(rti_get_error_stats(ctx, rel.data(), vov.data()), ctx, "rti_get_error_stats");
  rti_stats st;
  check(rti_get_stats(ctx, &st), ctx, "rti_get_stats");
  uint64_t hits = 0;
  for (uint64_t c : cnt) hits += c;
  std::cout << "trace result: discs " << n << ", rays " << st.n_rays << ", segments " << st.segments
            << ", surface hits " << st.surface_hits << ", deposits " << hits << ", misses " << st.misses
            << ", trace " << st.trace_ms << " ms, build " << st.build_ms << " ms, running time " << ns << " ns"
            << std::endl;
  std::cout << "after tracer" << std::endl;

  if (!o.outfile.empty()) {  // dev.cpp:254-285
    std::string out = o.outfile;
    if (extension(out) != ".vtp") {
      std::cout << "Appending .vtp to the given file name" << std::endl;
      out += ".vtp";
    }
    const std::string stem = out.substr(0, out.size() - 4);
    const std::string bbname = stem + ".bounding-box.vtp";
    std::string cmd;
    for (int i = 0; i < argc; ++i) cmd += std::string(" ") + argv[i];
    std::vector<rti_vtk::cell_array_out> arrays(5);
    arrays[0] = {"value", "Float64", 1, raw};
    arrays[1] = {"exposed-area", "Float64", 1, std::vector<double>(area.begin(), area.end())};
    arrays[2] = {"hitcnt", "UInt32", 1, std::vector<double>(n)};
    for (uint64_t i = 0; i < n; ++i) arrays[2].values[i] = (double)(uint32_t)cnt[i];
    arrays[3] = {"relative-error", "Float64", 1, rel};
    arrays[4] = {"variance-of-variance", "Float64", 1, vov};
    std::cout << "Writing output to " << out << std::endl;
    try {
      rti_vtk::write_point_cloud(out, dc.xyzr, dc.nrm, arrays,
                                 {{"running-time[ns]", std::to_string(ns)},
                                  {"git-hash", hash},
                                  {"cmd", cmd},
                                  {"geo-name", "point_cloud_disc_geometry"}});
      // boundary_x_y vertices and triangles (boundary_x_y.hpp:192-224)
      const double x0 = st.box[0], x1 = st.box[1], y0 = st.box[2], y1 = st.box[3], z0 = st.box[4], z1 = st.box[5];
      std::vector<double> v = {x1, y0, z0, x1, y0, z1, x1, y1, z0, x1, y1, z1,
                               x0, y0, z0, x0, y0, z1, x0, y1, z0, x0, y1, z1};
      std::vector<uint64_t> t = {0, 1, 2, 3, 2, 1, 4, 5, 0, 1, 0, 5, 6, 7, 4, 5, 4, 7, 2, 3, 6, 7, 6, 3};
      std::vector<double> nn;
      for (size_t k = 0; k < 8; ++k) {  // normalize((v1 - v0) x (v2 - v0)), util/utils.hpp:95-102
        const double* a = &v[3 * t[3 * k]];
        const double* b = &v[3 * t[3 * k + 1]];
        const double* c = &v[3 * t[3 * k + 2]];
        const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double w[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        const double l = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        for (int q = 0; q < 3; ++q) nn.push_back(cr[q] / l);
      }
      std::cout << "Writing bounding box to " << bbname << std::endl;
      rti_vtk::write_triangles(bbname, v, t, nn);
    } catch (const std::exception& e) {
      std::cerr << e.what() << std::endl;
      rti_ctx_destroy(ctx);
      return EXIT_FAILURE;
    }
  }
  rti_ctx_destroy(ctx);
  return EXIT_SUCCESS;
}
