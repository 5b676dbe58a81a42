#include "job_io.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>

namespace lpsio {

static bool slurp(const std::string &path, std::string &buf) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf.resize(len > 0 ? (size_t)len : 0);
  size_t got = len > 0 ? std::fread(&buf[0], 1, (size_t)len, f) : 0;
  buf.resize(got);
  std::fclose(f);
  return true;
}

// strtod is correctly rounded, as the iostream extraction of the reference is.
template <class F>
static bool read_numbers(const std::string &path, int64_t count, F &&store) {
  std::string buf;
  if (!slurp(path, buf)) return false;
  const char *p = buf.c_str();
  int64_t i = 0;
  for (; i < count; i++) {
    char *end = nullptr;
    errno = 0;
    const double v = std::strtod(p, &end);
    if (end == p) break;  // no more numbers: the rest stays 0
    store(i, v);
    p = end;
  }
  for (; i < count; i++) store(i, 0.0);
  return true;
}

bool read_doubles(const std::string &path, int64_t count, std::vector<double> &out) {
  out.assign((size_t)count, 0.0);
  return read_numbers(path, count, [&](int64_t i, double v) { out[(size_t)i] = v; });
}

bool read_ints(const std::string &path, int64_t count, std::vector<int32_t> &out) {
  out.assign((size_t)count, 0);
  return read_numbers(path, count, [&](int64_t i, double v) { out[(size_t)i] = (int32_t)v; });
}

bool read_rounded_ints(const std::string &path, int64_t count, std::vector<int32_t> &out) {
  out.assign((size_t)count, 0);
  return read_numbers(path, count,
                      [&](int64_t i, double v) { out[(size_t)i] = static_cast<int>(v + 0.5); });
}

static bool excluded(const std::vector<double> &nodes, int32_t i, double exclude_r) {
  if (exclude_r <= 0.0) return false;
  const double x = nodes[2 * (size_t)i], y = nodes[2 * (size_t)i + 1];
  return std::sqrt(x * x + y * y) > exclude_r;
}

static void vtk_points(std::ofstream &file, const std::vector<double> &nodes, int32_t n,
                       double exclude_r, int32_t count) {
  file << "# vtk DataFile Version 3.0" << std::endl;
  file << "PD2_LPS_ABC output" << std::endl;
  file << "ASCII" << std::endl;
  file << "DATASET POLYDATA" << std::endl;
  file << "POINTS " << count << " double" << std::endl;
  for (int32_t i = 0; i < n; i++) {
    if (excluded(nodes, i, exclude_r)) continue;
    file << nodes[2 * (size_t)i] << " " << nodes[2 * (size_t)i + 1] << " 0.0" << std::endl;
  }
  file << "VERTICES " << count << " " << 2 * count << std::endl;
  for (int32_t i = 0; i < count; i++) file << "1 " << i << std::endl;
  file << "POINT_DATA " << count << std::endl;
}

static int32_t count_points(const std::vector<double> &nodes, int32_t n, double exclude_r) {
  int32_t count = 0;
  for (int32_t i = 0; i < n; i++)
    if (!excluded(nodes, i, exclude_r)) count++;
  return count;
}

bool write_vtk_vector(const std::string &path, const std::vector<double> &nodes,
                      const std::vector<double> &vec, const char *name, int32_t n,
                      double exclude_r) {
  const int32_t count = count_points(nodes, n, exclude_r);
  std::ofstream file(path.c_str());
  if (!file) return false;
  file.precision(12);
  vtk_points(file, nodes, n, exclude_r, count);
  file << "VECTORS " << name << " double" << std::endl;
  for (int32_t i = 0; i < n; i++) {
    if (excluded(nodes, i, exclude_r)) continue;
    file << vec[2 * (size_t)i] << " " << vec[2 * (size_t)i + 1] << " 0.0" << std::endl;
  }
  file << "SCALARS " << name << "_magnitude double 1" << std::endl;
  file << "LOOKUP_TABLE default" << std::endl;
  for (int32_t i = 0; i < n; i++) {
    if (excluded(nodes, i, exclude_r)) continue;
    const double a = vec[2 * (size_t)i], b = vec[2 * (size_t)i + 1];
    file << std::sqrt(a * a + b * b) << std::endl;
  }
  return true;
}

bool write_vtk_scalar(const std::string &path, const std::vector<double> &nodes,
                      const std::vector<double> &vec, const char *name, int32_t n,
                      double exclude_r) {
  const int32_t count = count_points(nodes, n, exclude_r);
  std::ofstream file(path.c_str());
  if (!file) return false;
  file.precision(12);
  vtk_points(file, nodes, n, exclude_r, count);
  file << "SCALARS " << name << " double 1" << std::endl;
  file << "LOOKUP_TABLE default" << std::endl;
  for (int32_t i = 0; i < n; i++) {
    if (excluded(nodes, i, exclude_r)) continue;
    file << vec[(size_t)i] << std::endl;
  }
  return true;
}

}  // namespace lpsio
