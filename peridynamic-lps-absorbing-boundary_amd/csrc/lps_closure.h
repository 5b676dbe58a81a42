// Learned ABC closure: shared declarations of the MLP kernels (lps_mlp.hip) and of its coupling to
// the time loop (lps_closure.hip). The reference has no learned closure; its least-squares ABC
// (src/main.cpp:883-974) stays the update the closure corrects.
//
// The network: 1600 inputs -> 256 -> 256 -> 256 (ReLU) -> 4, BF16 operands on the matrix cores
// with FP32 accumulation, FP32 output layer.
//
// Inputs of one ABC row, in a canonical frame. The row's outward normal is mapped by a lattice
// symmetry (a signed permutation of x and y) into the sector 0 <= angle <= 45 degrees; every
// vector is expressed in that frame and the 49 lattice offsets |o|^2 <= 16 around the node (the
// row's cloud when it lies inside the domain) are visited in a fixed order:
//   k = 32 j + w, j < 49 the offset, w < 16 u, w >= 16 v; inside each half 8 history steps
//       (newest first) of (x, y): w = 16 p + 2 h + c.
//   j = 49: w = 0..3 the row features (nu', n'_x, n'_y, depth'), the rest 0.
// u values are divided by s_u (max |u| over the row's inputs), v values by s_v. Outputs are
// (du_x, du_y, dv_x, dv_y) in the canonical frame, in units of s_u sigma_u and s_v sigma_v, added
// to the least-squares U^{n+1} and V^{n+1} of the row's node.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

namespace lps {

constexpr int kClOffsets = 49;
constexpr int kClHist = 8;
constexpr int kClRec = 32;  // per history node: u (8 steps x 2), v (8 steps x 2), FP32
constexpr int kClIn = 32 * (kClOffsets + 1);  // 1600
constexpr int kClHid = 256;
constexpr int kClOut = 4;
constexpr int kClFeat = 4;
constexpr int kClRows = 32;  // rows per workgroup of the forward kernel

// FP32 master parameters, one flat array: W1 [256][1600], b1, W2 [256][256], b2, W3, b3,
// W4 [4][256], b4 (weights stored [out][in]).
constexpr int64_t kOffW1 = 0;
constexpr int64_t kOffB1 = kOffW1 + (int64_t)kClHid * kClIn;
constexpr int64_t kOffW2 = kOffB1 + kClHid;
constexpr int64_t kOffB2 = kOffW2 + (int64_t)kClHid * kClHid;
constexpr int64_t kOffW3 = kOffB2 + kClHid;
constexpr int64_t kOffB3 = kOffW3 + (int64_t)kClHid * kClHid;
constexpr int64_t kOffW4 = kOffB3 + kClHid;
constexpr int64_t kOffB4 = kOffW4 + (int64_t)kClOut * kClHid;
constexpr int64_t kClParams = kOffB4 + kClOut;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// What the forward kernels read of the network (device pointers).
struct MlpDev {
  const __bf16 *w1, *w2, *w3;     // BF16 copies of W1..W3, [out][in]
  const __bf16 *w1p, *w2p, *w3p;  // the same in matrix-core fragment order (pack_index)
  const float *p;                 // FP32 master parameters (biases and the output layer)
  float sigma_u, sigma_v;         // output scales
};

// Fragment order of a [256 out][K in] BF16 weight: K step k (16 inputs) holds the B operands of
// the 8 column tiles (32 outputs each) for the 64 lanes, 8 values per lane, contiguous (8 KB).
__device__ __host__ __forceinline__ int64_t pack_index(int64_t o, int64_t i) {
  const int64_t k = i >> 4, w = i & 15;
  return (((k * 8 + (o >> 5)) * 64 + (o & 31) + 32 * (w >> 3)) << 3) + (w & 7);
}

// Inputs of the ABC rows gathered from the per-node history (see the header comment).
struct GatherSrc {
  const float *hist;    // n_hist x 32
  const float *mx;      // n_hist x 2: max |u|, max |v| over the node's record
  const int32_t *hidx;  // n_rows x 49: history node of each offset, -1 when absent
  const int8_t *xf;     // n_rows: frame code, bit 0 swap x/y, bit 1 negate x', bit 2 negate y'
  const float *feat;    // n_rows x 4
};

// Canonical components of a physical vector (x, y) under frame code f.
__device__ __host__ __forceinline__ void to_frame(int f, float x, float y, float &cx, float &cy) {
  const float a = (f & 1) ? y : x, b = (f & 1) ? x : y;
  cx = (f & 2) ? -a : a;
  cy = (f & 4) ? -b : b;
}
// Physical components of a canonical vector.
__device__ __host__ __forceinline__ void from_frame(int f, double cx, double cy, double &x,
                                                    double &y) {
  const double a = (f & 2) ? -cx : cx, b = (f & 4) ? -cy : cy;
  x = (f & 1) ? b : a;
  y = (f & 1) ? a : b;
}

// 8 consecutive inputs k..k+7 (k % 8 == 0) of row r, scaled and rounded to BF16.
__device__ __forceinline__ bf16x8 gather8(const GatherSrc &g, int32_t r, int k, float inv_u,
                                          float inv_v) {
  bf16x8 out;
  const int j = k >> 5, w = k & 31;
  if (j == kClOffsets) {
#pragma unroll
    for (int e = 0; e < 8; e++) out[e] = (__bf16)0.0f;
    if (w == 0) {
#pragma unroll
      for (int e = 0; e < kClFeat; e++) out[e] = (__bf16)g.feat[(int64_t)r * kClFeat + e];
    }
    return out;
  }
  const int32_t h = g.hidx[(int64_t)r * kClOffsets + j];
  if (h < 0) {
#pragma unroll
    for (int e = 0; e < 8; e++) out[e] = (__bf16)0.0f;
    return out;
  }
  const float4 *q = (const float4 *)(g.hist + (int64_t)h * kClRec + w);
  const float4 a = q[0], b = q[1];
  const int f = g.xf[r];
  const float inv = w < 16 ? inv_u : inv_v;
  float cx, cy;
  to_frame(f, a.x, a.y, cx, cy); out[0] = (__bf16)(cx * inv); out[1] = (__bf16)(cy * inv);
  to_frame(f, a.z, a.w, cx, cy); out[2] = (__bf16)(cx * inv); out[3] = (__bf16)(cy * inv);
  to_frame(f, b.x, b.y, cx, cy); out[4] = (__bf16)(cx * inv); out[5] = (__bf16)(cy * inv);
  to_frame(f, b.z, b.w, cx, cy); out[6] = (__bf16)(cx * inv); out[7] = (__bf16)(cy * inv);
  return out;
}

// Row scales: s = max over the row's history nodes of mx; inv = 1 / s, or 0 when s < 1e-30.
__device__ __forceinline__ void row_scale(const GatherSrc &g, int32_t r, float &su, float &sv) {
  su = 0.0f;
  sv = 0.0f;
  for (int j = 0; j < kClOffsets; j++) {
    const int32_t h = g.hidx[(int64_t)r * kClOffsets + j];
    if (h >= 0) {
      su = fmaxf(su, g.mx[2 * (int64_t)h]);
      sv = fmaxf(sv, g.mx[2 * (int64_t)h + 1]);
    }
  }
}
__device__ __forceinline__ float safe_inv(float s) { return s >= 1e-30f ? 1.0f / s : 0.0f; }

// Forward of the network over n rows (lps_mlp.hip).
//  dataset inputs: x is n x 1600 BF16; h1..h3 (n x 256 BF16 activations) and y (n x 4) are
//    written when not null.
hipError_t mlp_forward_dataset(const MlpDev &m, const __bf16 *x, int32_t n, __bf16 *h1, __bf16 *h2,
                               __bf16 *h3, float *y, hipStream_t s);
//  gathered inputs: the correction of each row is added to U^{n+1}, V^{n+1} of its node:
//    slot[r] the node, gain[inst[r]] a factor per instance.
struct ApplyDst {
  const int32_t *slot;
  const int32_t *inst;
  const float *gain;
  const float *gate;  // per instance: no correction where s_u < gate
  double *ux1, *uy1, *vx1, *vy1;
  float *dbg;  // n x 8 when not null: the 4 outputs, s_u, s_v, and 0, 0 (tests)
  // n x 8 when not null: the network stores each row's 4 outputs, s_u, s_v, 0, 0 here instead of
  // adding its correction; mlp_apply_deferred adds them later (the same bits)
  float *defer = nullptr;
};
// wide_waves 4 or 8: 32 wide_waves rows per workgroup (every wavefront all 256 columns, weights
// staged in LDS), for large row counts; otherwise (0) 32 rows per workgroup. All give the same bits.
hipError_t mlp_forward_apply(const MlpDev &m, const GatherSrc &g, int32_t n, const ApplyDst &d,
                             int wide_waves, hipStream_t s);
// The corrections stored by mlp_forward_apply with d.defer set, added to the state.
hipError_t mlp_apply_deferred(const MlpDev &m, const GatherSrc &g, int32_t n, const ApplyDst &d,
                              hipStream_t s);
// Rows of gathered inputs written out (n x 1600 BF16) with their scales (n x 2).
hipError_t gather_rows(const GatherSrc &g, int32_t n, __bf16 *x, float *scale, hipStream_t s);

uint64_t splitmix64(uint64_t x);
__device__ __forceinline__ uint64_t dev_mix(uint64_t x) {  // splitmix64 on the device
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

}  // namespace lps

// The network, its optimizer state and its sample store on one device (lps_mlp.hip).
struct lps_closure {
  int32_t device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;  // params, grads, Adam moments
  __bf16 *w1 = nullptr, *w2 = nullptr, *w3 = nullptr;               // [out][in] BF16 copies
  __bf16 *w2t = nullptr, *w3t = nullptr;                            // [in][out] BF16 copies
  __bf16 *w1p = nullptr, *w2p = nullptr, *w3p = nullptr;            // fragment order
  float sigma_u = 1.0f, sigma_v = 1.0f;
  int64_t adam_t = 0;
  // samples: x (cap x 1600 BF16), aux (cap x 8: target 4, weight, s_u, s_v, 0)
  __bf16 *ds_x = nullptr;
  float *ds_aux = nullptr;
  unsigned long long *ds_count = nullptr;  // device counter (appends)
  int64_t ds_cap = 0;
  // training buffers for a batch of tb rows
  int32_t tb = 0;
  __bf16 *xb = nullptr, *xbt = nullptr, *h[3] = {nullptr, nullptr, nullptr};
  __bf16 *ht[3] = {nullptr, nullptr, nullptr}, *dz[3] = {nullptr, nullptr, nullptr};
  __bf16 *dzt[3] = {nullptr, nullptr, nullptr}, *dyt = nullptr;
  float *yb = nullptr, *tgt = nullptr, *wgt = nullptr, *dy = nullptr, *lrow = nullptr;
  float *loss_d = nullptr;
  // this device's copy of the staging area of lps_closure_forward_noalloc (its address from
  // hipGetSymbolAddress at creation; the array's own name is not a device address on the host)
  unsigned char *scratch = nullptr;
  std::vector<void *> allocs;

  lps::MlpDev dev() const {
    return lps::MlpDev{w1, w2, w3, w1p, w2p, w3p, p, sigma_u, sigma_v};
  }
  template <class T>
  hipError_t alloc(T **q, size_t count) {
    hipError_t e = hipMalloc((void **)q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) allocs.push_back(*q);
    return e;
  }
  int32_t fail(int32_t code, const std::string &msg) {
    err = msg;
    return code;
  }
  int32_t hip_fail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return -2;
  }
};
