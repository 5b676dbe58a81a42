// Job files of the LPS-ABC solver: readers with the semantics of src/helpers/utils.h:24-80 and
// src/main.cpp:178-293, and the legacy-ASCII VTK writers of src/helpers/utils.h:94-224.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace lpsio {

// Reads `count` whitespace-separated numbers. Returns false if the file cannot be opened; values
// missing at the end of the file read as 0, as a failed iostream extraction leaves them.
bool read_doubles(const std::string &path, int64_t count, std::vector<double> &out);
bool read_ints(const std::string &path, int64_t count, std::vector<int32_t> &out);
// Index arrays written as floats: read as double, converted with (int)(x + 0.5) (utils.h:61, :76).
bool read_rounded_ints(const std::string &path, int64_t count, std::vector<int32_t> &out);

// Legacy VTK POLYDATA encodings: the reference's ASCII (precision 12), or BINARY (big-endian
// doubles and int32, the same values unrounded; no counterpart in the reference).
enum class VtkFormat { Ascii, Binary };

// src/helpers/utils.h:94-167 and :170-224 (points with r > exclude_r dropped when exclude_r > 0).
bool write_vtk_vector(const std::string &path, const std::vector<double> &nodes,
                      const std::vector<double> &vec, const char *name, int32_t n,
                      double exclude_r, VtkFormat fmt = VtkFormat::Ascii);
bool write_vtk_scalar(const std::string &path, const std::vector<double> &nodes,
                      const std::vector<double> &vec, const char *name, int32_t n,
                      double exclude_r, VtkFormat fmt = VtkFormat::Ascii);

}  // namespace lpsio
