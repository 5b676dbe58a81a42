#include "job_io.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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

// Big-endian output of the BINARY encoding.
static void put_be(std::ofstream &file, double v) {
  uint64_t b;
  std::memcpy(&b, &v, 8);
  unsigned char c[8];
  for (int k = 0; k < 8; k++) c[k] = (unsigned char)(b >> (56 - 8 * k));
  file.write(reinterpret_cast<const char *>(c), 8);
}
static void put_be(std::ofstream &file, int32_t v) {
  const uint32_t b = (uint32_t)v;
  unsigned char c[4];
  for (int k = 0; k < 4; k++) c[k] = (unsigned char)(b >> (24 - 8 * k));
  file.write(reinterpret_cast<const char *>(c), 4);
}

static void vtk_points(std::ofstream &file, const std::vector<double> &nodes, int32_t n,
                       double exclude_r, int32_t count, VtkFormat fmt) {
  const bool bin = fmt == VtkFormat::Binary;
  file << "# vtk DataFile Version 3.0" << std::endl;
  file << "PD2_LPS_ABC output" << std::endl;
  file << (bin ? "BINARY" : "ASCII") << std::endl;
  file << "DATASET POLYDATA" << std::endl;
  file << "POINTS " << count << " double" << std::endl;
  for (int32_t i = 0; i < n; i++) {
    if (excluded(nodes, i, exclude_r)) continue;
    if (bin) {
      put_be(file, nodes[2 * (size_t)i]);
      put_be(file, nodes[2 * (size_t)i + 1]);
      put_be(file, 0.0);
    } else {
      file << nodes[2 * (size_t)i] << " " << nodes[2 * (size_t)i + 1] << " 0.0" << std::endl;
    }
  }
  if (bin) file << std::endl;
  file << "VERTICES " << count << " " << 2 * count << std::endl;
  for (int32_t i = 0; i < count; i++) {
    if (bin) {
      put_be(file, (int32_t)1);
      put_be(file, i);
    } else {
      file << "1 " << i << std::endl;
    }
  }
  if (bin) file << std::endl;
  file << "POINT_DATA " << count << std::endl;
}

static int32_t count_points(const std::vector<double> &nodes, int32_t n, double exclude_r) {
  int32_t count = 0;
  for (int32_t i = 0; i < n; i++)
    if (!excluded(nodes, i, exclude_r)) count++;
  return count;
}

static std::ofstream open_vtk(const std::string &path, VtkFormat fmt) {
  std::ofstream file(path.c_str(), fmt == VtkFormat::Binary ? std::ios::out | std::ios::binary
                                                            : std::ios::out);
  file.precision(12);
  return file;
}

bool write_vtk_vector(const std::string &path, const std::vector<double> &nodes,
                      const std::vector<double> &vec, const char *name, int32_t n,
                      double exclude_r, VtkFormat fmt) {
  const int32_t count = count_points(nodes, n, exclude_r);
  std::ofstream file = open_vtk(path, fmt);
  if (!file) return false;
  const bool bin = fmt == VtkFormat::Binary;
  vtk_points(file, nodes, n, exclude_r, count, fmt);
  file << "VECTORS " << name << " double" << std::endl;
  for (int32_t i = 0; i < n; i++) {
    if (excluded(nodes, i, exclude_r)) continue;
    if (bin) {
      put_be(file, vec[2 * (size_t)i]);
      put_be(file, vec[2 * (size_t)i + 1]);
      put_be(file, 0.0);
    } else {
      file << vec[2 * (size_t)i] << " " << vec[2 * (size_t)i + 1] << " 0.0" << std::endl;
    }
  }
  if (bin) file << std::endl;
  file << "SCALARS " << name << "_magnitude double 1" << std::endl;
  file << "LOOKUP_TABLE default" << std::endl;
  for (int32_t i = 0; i < n; i++) {
    if (excluded(nodes, i, exclude_r)) continue;
    const double a = vec[2 * (size_t)i], b = vec[2 * (size_t)i + 1];
    if (bin) put_be(file, std::sqrt(a * a + b * b));
    else file << std::sqrt(a * a + b * b) << std::endl;
  }
  if (bin) file << std::endl;
  return (bool)file;
}

bool write_vtk_scalar(const std::string &path, const std::vector<double> &nodes,
                      const std::vector<double> &vec, const char *name, int32_t n,
                      double exclude_r, VtkFormat fmt) {
  const int32_t count = count_points(nodes, n, exclude_r);
  std::ofstream file = open_vtk(path, fmt);
  if (!file) return false;
  const bool bin = fmt == VtkFormat::Binary;
  vtk_points(file, nodes, n, exclude_r, count, fmt);
  file << "SCALARS " << name << " double 1" << std::endl;
  file << "LOOKUP_TABLE default" << std::endl;
  for (int32_t i = 0; i < n; i++) {
    if (excluded(nodes, i, exclude_r)) continue;
    if (bin) put_be(file, vec[(size_t)i]);
    else file << vec[(size_t)i] << std::endl;
  }
  if (bin) file << std::endl;
  return (bool)file;
}

}  // namespace lpsio
