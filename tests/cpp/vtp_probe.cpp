// Reads a .vtp point cloud with the command-line tool's reader and prints
// what it extracted: "n", then one "x y z r nx ny nz" line per disc. With a
// second argument, also writes the discs back out with the tool's writer
// (cell data "value" = 0.5 * index, "hitcnt" = index, string metadata).
#include <cstdio>
#include <exception>
#include <string>

#include "../../rti_amd/cli/vtkxml.hpp"

int main(int argc, char** argv) {
  if (argc != 2 && argc != 3) return 2;
  try {
    auto dc = rti_vtk::discs_from(rti_vtk::read_vtp(argv[1]));
    const size_t n = dc.xyzr.size() / 4;
    std::printf("%zu\n", n);
    for (size_t i = 0; i < n; ++i)
      std::printf("%.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", dc.xyzr[4 * i], dc.xyzr[4 * i + 1], dc.xyzr[4 * i + 2],
                  dc.xyzr[4 * i + 3], dc.nrm[3 * i], dc.nrm[3 * i + 1], dc.nrm[3 * i + 2]);
    if (argc == 3) {
      std::vector<rti_vtk::cell_array_out> arrays(2);
      arrays[0] = {"value", "Float64", 1, std::vector<double>(n)};
      arrays[1] = {"hitcnt", "UInt32", 1, std::vector<double>(n)};
      for (size_t i = 0; i < n; ++i) {
        arrays[0].values[i] = 0.5 * (double)i;
        arrays[1].values[i] = (double)i;
      }
      rti_vtk::write_point_cloud(argv[2], dc.xyzr, dc.nrm, arrays, {{"cmd", "probe <&> \"quoted\""}});
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
