// Test driver of the VTK writers (host/job_io.cpp): writes one vector and one scalar file in the
// format given on the command line from a small deterministic field.
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "job_io.h"

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  const std::string out = argv[1];
  const lpsio::VtkFormat fmt =
      std::string(argv[2]) == "binary" ? lpsio::VtkFormat::Binary : lpsio::VtkFormat::Ascii;
  const int n = 300;
  std::vector<double> nodes(2 * n), u(2 * n), th(n);
  for (int i = 0; i < n; i++) {
    nodes[2 * i] = 0.5 * (i % 20) - 5.0;
    nodes[2 * i + 1] = 0.5 * (i / 20) - 3.5;
    u[2 * i] = std::sin(0.1 * i) * 1e-3;
    u[2 * i + 1] = std::cos(0.37 * i) / 3.0;
    th[i] = std::exp(-0.01 * i) - 0.5;
  }
  const double exclude_r = 5.2;  // drops the corners
  bool ok = lpsio::write_vtk_vector(out + "-vec.vtk", nodes, u, "displacement", n, exclude_r, fmt);
  ok = ok && lpsio::write_vtk_scalar(out + "-sca.vtk", nodes, th, "dilatation", n, exclude_r, fmt);
  return ok ? 0 : 1;
}
