// Harness around the reference's horizon build, compiled from the reference sources where they lie
// (src/helpers/kd_tree_interface.h + the vendored libkdtree++). TEST INFRASTRUCTURE ONLY: it writes
// golden neighbour sets for the tests; nothing in the product path links it.
#include <vector>
#include <stdexcept>
#include "kd_tree_interface.h"

namespace {
struct NodeMatrix {  // the (i, k) / size1() interface build_clouds reads
  const double *p;
  int n;
  int size1() const { return n; }
  double operator()(int i, int k) const { return p[2 * i + k]; }
};
struct IntMatrix {
  std::vector<int> v;
  int cols = 0;
  void resize(int r, int c) { v.assign((size_t)r * c, 0); cols = c; }
  int &operator()(int i, int j) { return v[(size_t)i * cols + j]; }
};
struct IntVector {
  std::vector<int> v;
  void resize(int n) { v.assign(n, 0); }
  int &operator[](int i) { return v[i]; }
};
}  // namespace

extern "C" int ref_build_clouds(int n, const double *nodes, double r, int max_len, int *clouds_out,
                                int *lengths_out) {
  NodeMatrix nm{nodes, n};
  IntMatrix cl;
  IntVector len;
  int maxlen = 0;
  try {
    build_clouds(0, max_len, r, 0.0, nm, cl, len, maxlen);
  } catch (const std::exception &) {
    return -1;
  }
  for (size_t k = 0; k < cl.v.size(); k++) clouds_out[k] = cl.v[k];
  for (int i = 0; i < n; i++) lengths_out[i] = len.v[i];
  return maxlen;
}
