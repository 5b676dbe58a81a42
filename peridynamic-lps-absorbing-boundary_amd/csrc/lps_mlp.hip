// Learned ABC closure, the network: forward (fused, matrix cores), backward, Adam, and the sample
// store, behind the lps_closure_* entry points of include/lps_amd.h. No counterpart in the
// reference (its ABC is the least-squares update alone, src/main.cpp:883-974).
//
// Matrix-core tiles: v_mfma_f32_32x32x16_bf16. Lane l holds A[l % 32][8 (l / 32) + e] and
// B[8 (l / 32) + e][l % 32], e < 8; accumulator element e of lane l is
// D[8 (e / 4) + 4 (l / 32) + e % 4][l % 32].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lps_amd.h"
#include "lps_closure.h"
#include "lps_internal.h"

namespace lps {

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int e = 0; e < 8; e++) z[e] = (__bf16)0.0f;
  return z;
}

// Accumulator element e of this lane -> row inside the 32 x 32 tile.
__device__ __forceinline__ int acc_row(int e, int lane) {
  return 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
}

constexpr int kHsPitch = kClHid + 8;  // LDS row pitch of the activations (BF16)

// bias + ReLU of a wave's 32 x 64 block (two 32 x 32 tiles) into LDS and, if out, global memory.
__device__ __forceinline__ void store_hidden(const f32x16 &c0, const f32x16 &c1, const float *bias,
                                             __bf16 (*hs)[kHsPitch], __bf16 *out, int r0, int n,
                                             int wv, int lane) {
  const int col0 = 64 * wv + (lane & 31), col1 = col0 + 32;
  const float b0 = bias[col0], b1 = bias[col1];
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const int m = acc_row(e, lane);
    const __bf16 v0 = (__bf16)fmaxf(c0[e] + b0, 0.0f), v1 = (__bf16)fmaxf(c1[e] + b1, 0.0f);
    hs[m][col0] = v0;
    hs[m][col1] = v1;
    if (out && r0 + m < n) {
      out[(int64_t)(r0 + m) * kClHid + col0] = v0;
      out[(int64_t)(r0 + m) * kClHid + col1] = v1;
    }
  }
}

// One hidden layer 256 -> 256 with the input in LDS: this wave's 64 output columns.
__device__ __forceinline__ void dense_lds(const __bf16 (*hs)[kHsPitch], const __bf16 *w, int wv,
                                          int lane, f32x16 &c0, f32x16 &c1) {
  const int ar = lane & 31, ak = 8 * (lane >> 5);
  const __bf16 *wa = w + (int64_t)(64 * wv + ar) * kClHid + ak;
  const __bf16 *wb = wa + 32 * kClHid;
#pragma unroll 4
  for (int k = 0; k < kClHid; k += 16) {
    const bf16x8 a = *(const bf16x8 *)&hs[ar][k + ak];
    c0 = mfma(a, *(const bf16x8 *)(wa + k), c0);
    c1 = mfma(a, *(const bf16x8 *)(wb + k), c1);
  }
}

// Correction of one row added to the state (the gathered path's epilogue).
__device__ __forceinline__ void apply_row(const MlpDev &m, const GatherSrc &g, const ApplyDst &d,
                                          int32_t row, const float *y, float su, float sv) {
  const int f = g.xf[row];
  const int32_t in = d.inst[row];
  const double gain = su >= d.gate[in] ? (double)d.gain[in] : 0.0;
  double dux, duy, dvx, dvy;
  from_frame(f, (double)y[0], (double)y[1], dux, duy);
  from_frame(f, (double)y[2], (double)y[3], dvx, dvy);
  const double ku = gain * m.sigma_u * (double)su, kv = gain * m.sigma_v * (double)sv;
  const int32_t i = d.slot[row];
  d.ux1[i] += ku * dux;
  d.uy1[i] += ku * duy;
  d.vx1[i] += kv * dvx;
  d.vy1[i] += kv * dvy;
  if (d.dbg) {
    float *o = d.dbg + (int64_t)row * 8;
    o[0] = y[0]; o[1] = y[1]; o[2] = y[2]; o[3] = y[3];
    o[4] = su; o[5] = sv; o[6] = 0.0f; o[7] = 0.0f;
  }
}

// A row's end in the gathered path: its correction added to the state, or with d.defer its
// outputs and scales stored, for k_apply_deferred to add.
__device__ __forceinline__ void finish_row(const MlpDev &m, const GatherSrc &g, const ApplyDst &d,
                                           int32_t row, const float *y, float su, float sv) {
  if (d.defer) {
    float4 *o = (float4 *)(d.defer + (int64_t)row * 8);
    o[0] = make_float4(y[0], y[1], y[2], y[3]);
    o[1] = make_float4(su, sv, 0.0f, 0.0f);
    return;
  }
  apply_row(m, g, d, row, y, su, sv);
}

// The stored corrections of n rows added to the state: the same arithmetic as the network's own
// epilogue, so a deferred step gives the bits of an undeferred one.
__global__ void __launch_bounds__(256) k_apply_deferred(MlpDev m, GatherSrc g, int32_t n,
                                                        ApplyDst d) {
  const int32_t row = blockIdx.x * 256 + threadIdx.x;
  if (row >= n) return;
  const float4 *o = (const float4 *)(d.defer + (int64_t)row * 8);
  const float4 a = o[0], b = o[1];
  const float y[4] = {a.x, a.y, a.z, a.w};
  apply_row(m, g, d, row, y, b.x, b.y);
}

// The network over 32 rows per workgroup (4 wavefronts, 64 output columns each). kGather: inputs
// gathered from the history (gather8) and the correction added to the state; else inputs from x.
template <bool kGather>
__global__ void __launch_bounds__(256) k_mlp_fwd(MlpDev m, const __bf16 *x, GatherSrc g, int32_t n,
                                                 __bf16 *h1o, __bf16 *h2o, __bf16 *h3o, float *yo,
                                                 ApplyDst d) {
  __shared__ __bf16 hs[kClRows][kHsPitch];
  __shared__ float sc[kClRows][4];  // 1/s_u, 1/s_v, s_u, s_v
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int r0 = blockIdx.x * kClRows;
  if constexpr (kGather) {
    const int r = t >> 3, j = t & 7;
    float su = 0.0f, sv = 0.0f;
    if (r0 + r < n) {
      for (int q = j; q < kClOffsets; q += 8) {
        const int32_t h = g.hidx[(int64_t)(r0 + r) * kClOffsets + q];
        if (h >= 0) {
          su = fmaxf(su, g.mx[2 * (int64_t)h]);
          sv = fmaxf(sv, g.mx[2 * (int64_t)h + 1]);
        }
      }
    }
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
      su = fmaxf(su, __shfl_xor(su, o));
      sv = fmaxf(sv, __shfl_xor(sv, o));
    }
    if (j == 0) {
      sc[r][0] = safe_inv(su);
      sc[r][1] = safe_inv(sv);
      sc[r][2] = su;
      sc[r][3] = sv;
    }
    __syncthreads();
  }
  // layer 1: 1600 -> 256, inputs from global memory
  {
    const int ar = lane & 31, ak = 8 * (lane >> 5);
    const int row = r0 + ar;
    const bool live = row < n;
    float iu = 0.0f, iv = 0.0f;
    if constexpr (kGather) {
      iu = sc[ar][0];
      iv = sc[ar][1];
    }
    const __bf16 *wa = m.w1 + (int64_t)(64 * wv + ar) * kClIn + ak;
    const __bf16 *wb = wa + 32 * kClIn;
    f32x16 c0 = {}, c1 = {};
#pragma unroll 2
    for (int k = 0; k < kClIn; k += 16) {
      bf16x8 a;
      if (!live) a = zero8();
      else if constexpr (kGather) a = gather8(g, row, k + ak, iu, iv);
      else a = *(const bf16x8 *)(x + (int64_t)row * kClIn + k + ak);
      c0 = mfma(a, *(const bf16x8 *)(wa + k), c0);
      c1 = mfma(a, *(const bf16x8 *)(wb + k), c1);
    }
    store_hidden(c0, c1, m.p + kOffB1, hs, h1o, r0, n, wv, lane);
  }
  __syncthreads();
  // layers 2 and 3: 256 -> 256 from LDS
  for (int layer = 0; layer < 2; layer++) {
    f32x16 c0 = {}, c1 = {};
    dense_lds(hs, layer == 0 ? m.w2 : m.w3, wv, lane, c0, c1);
    __syncthreads();
    store_hidden(c0, c1, m.p + (layer == 0 ? kOffB2 : kOffB3), hs, layer == 0 ? h2o : h3o, r0, n,
                 wv, lane);
    __syncthreads();
  }
  // output layer (FP32): 8 threads per row, 32 inputs each
  const int r = t >> 3, j = t & 7;
  const float *w4 = m.p + kOffW4;
  float y[kClOut] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int q = 0; q < 32; q++) {
    const int k = 32 * j + q;
    const float hv = (float)hs[r][k];
#pragma unroll
    for (int c = 0; c < kClOut; c++) y[c] += hv * w4[c * kClHid + k];
  }
#pragma unroll
  for (int c = 0; c < kClOut; c++) {
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) y[c] += __shfl_xor(y[c], o);
    y[c] += m.p[kOffB4 + c];
  }
  const int row = r0 + r;
  if (j != 0 || row >= n) return;
  if constexpr (!kGather) {
    if (yo) {
#pragma unroll
      for (int c = 0; c < kClOut; c++) yo[(int64_t)row * kClOut + c] = y[c];
    }
  } else {
    finish_row(m, g, d, row, y, sc[r][2], sc[r][3]);
  }
}

// The network over 32 kW rows per workgroup of kW wavefronts: wavefront w owns rows 32 w .. 32 w + 31
// and all 256 columns (8 accumulator tiles); each K step's weight fragments are staged once per
// workgroup in LDS (double buffered). The operands of K step k + kPf are loaded while step k
// computes. Same arithmetic order as k_mlp_fwd, so the same bits.
//  kW = 4: 128 rows, 86 KB of LDS, one wavefront per SIMD.
//  kW = 8: 256 rows, 152 KB of LDS (one workgroup per CU), two wavefronts per SIMD: half the
//    weight traffic per row, and each SIMD switches to the other wavefront while one waits.
constexpr int kPf = 4;

struct RawG {  // 8 history values (4 (x, y) pairs) of one K step, FP32
  float4 a, b;
};

template <bool kGather, int kW>
__global__ void __launch_bounds__(64 * kW) k_mlp_fwd_wide(MlpDev m, const __bf16 *x, GatherSrc g,
                                                          int32_t n, float *yo, ApplyDst d) {
  constexpr int kT = 64 * kW, kRows = 32 * kW;
  constexpr int kLd = 512 / kT;  // weight fragments (of 512 per K step) each thread stages
  __shared__ bf16x8 bs[2][8][64];
  __shared__ __bf16 hs[kRows][kHsPitch];
  __shared__ float sc[kRows][4];
  // layer 1 keeps the rows' history indices where the activations go later
  int32_t(*hi)[kClOffsets] = reinterpret_cast<int32_t(*)[kClOffsets]>(&hs[0][0]);
  static_assert(sizeof(int32_t) * kRows * kClOffsets <= sizeof(__bf16) * kRows * kHsPitch,
                "history indices fit in the activation buffer");
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  // XCD-contiguous row ranges: neighbouring rows share history records through one L2
  const int r0 = xcd_block(blockIdx.x, (n + kRows - 1) / kRows) * kRows;
  if (r0 >= n) return;  // the whole block: before any barrier
  const int ar = lane & 31, ak = 8 * (lane >> 5);
  const int lrow = 32 * wv + ar;
  const int32_t row = r0 + lrow;
  const bool live = row < n;
  if constexpr (kGather) {
    for (int q = t; q < kRows * kClOffsets; q += kT) {
      const int rr = q / kClOffsets, jj = q % kClOffsets;
      hi[rr][jj] = r0 + rr < n ? g.hidx[(int64_t)(r0 + rr) * kClOffsets + jj] : -1;
    }
    __syncthreads();
    const int r = t >> 1, j = t & 1;
    float su = 0.0f, sv = 0.0f;
    for (int q = j; q < kClOffsets; q += 2) {
      const int32_t h = hi[r][q];
      if (h >= 0) {
        su = fmaxf(su, g.mx[2 * (int64_t)h]);
        sv = fmaxf(sv, g.mx[2 * (int64_t)h + 1]);
      }
    }
    su = fmaxf(su, __shfl_xor(su, 1));
    sv = fmaxf(sv, __shfl_xor(sv, 1));
    if (j == 0) {
      sc[r][0] = safe_inv(su);
      sc[r][1] = safe_inv(sv);
      sc[r][2] = su;
      sc[r][3] = sv;
    }
    __syncthreads();
  }
  // the row's frame transform folded into its scales: to_frame swaps a pair's components (bit 0)
  // and negates the first (bit 1) and second (bit 2); (-c) * inv and c * (-inv) are the same
  // floats, so the inputs keep their bits. su_a / su_b: the first / second component's scale of
  // the u part, sv_a / sv_b of the v part.
  float su_a = 0.0f, su_b = 0.0f, sv_a = 0.0f, sv_b = 0.0f;
  bool swap = false;
  if constexpr (kGather) {
    const float iu = sc[lrow][0], iv = sc[lrow][1];
    const int f = live ? g.xf[row] : 0;
    swap = (f & 1) != 0;
    su_a = (f & 2) ? -iu : iu;
    su_b = (f & 4) ? -iu : iu;
    sv_a = (f & 2) ? -iv : iv;
    sv_b = (f & 4) ? -iv : iv;
  }
  // loads of layer-1 inputs: history record (gathered) or the input row
  auto issue_a = [&](int k, RawG &raw, bf16x8 &xr) {
    if constexpr (kGather) {
      raw.a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      raw.b = raw.a;
      const int j = k >> 1;
      if (live && j < kClOffsets) {
        const int32_t h = hi[lrow][j];
        if (h >= 0) {
          const float4 *q = (const float4 *)(g.hist + (int64_t)h * kClRec + 16 * (k & 1) + ak);
          raw.a = q[0];
          raw.b = q[1];
        }
      }
    } else {
      xr = live ? *(const bf16x8 *)(x + (int64_t)row * kClIn + 16 * k + ak) : zero8();
    }
  };
  auto make_a = [&](int k, const RawG &raw, const bf16x8 &xr) -> bf16x8 {
    if constexpr (kGather) {
      bf16x8 out;
      if ((k >> 1) == kClOffsets) {  // row features
#pragma unroll
        for (int e = 0; e < 8; e++) out[e] = (__bf16)0.0f;
        if (live && 16 * (k & 1) + ak == 0) {
#pragma unroll
          for (int e = 0; e < kClFeat; e++) out[e] = (__bf16)g.feat[(int64_t)row * kClFeat + e];
        }
        return out;
      }
      const float sa = (k & 1) ? sv_a : su_a, sb = (k & 1) ? sv_b : su_b;
      out[0] = (__bf16)((swap ? raw.a.y : raw.a.x) * sa);
      out[1] = (__bf16)((swap ? raw.a.x : raw.a.y) * sb);
      out[2] = (__bf16)((swap ? raw.a.w : raw.a.z) * sa);
      out[3] = (__bf16)((swap ? raw.a.z : raw.a.w) * sb);
      out[4] = (__bf16)((swap ? raw.b.y : raw.b.x) * sa);
      out[5] = (__bf16)((swap ? raw.b.x : raw.b.y) * sb);
      out[6] = (__bf16)((swap ? raw.b.w : raw.b.z) * sa);
      out[7] = (__bf16)((swap ? raw.b.z : raw.b.w) * sb);
      return out;
    } else {
      return xr;
    }
  };
  f32x16 acc[8];
  for (int layer = 0; layer < 3; layer++) {
    const __bf16 *wp = layer == 0 ? m.w1p : layer == 1 ? m.w2p : m.w3p;
    const int nk = (layer == 0 ? kClIn : kClHid) / 16;  // multiples of kPf
#pragma unroll
    for (int q = 0; q < 8; q++) acc[q] = f32x16{};
    RawG ra[kPf];
    bf16x8 xa[kPf], rb[kPf][kLd];
#pragma unroll
    for (int dd = 0; dd < kPf; dd++) {
      const bf16x8 *src = (const bf16x8 *)(wp + (int64_t)dd * 4096);
#pragma unroll
      for (int u = 0; u < kLd; u++) rb[dd][u] = src[t + kT * u];
      if (layer == 0) issue_a(dd, ra[dd], xa[dd]);
    }
    for (int k0 = 0; k0 < nk; k0 += kPf) {
#pragma unroll
      for (int dd = 0; dd < kPf; dd++) {
        const int k = k0 + dd;
#pragma unroll
        for (int u = 0; u < kLd; u++) {
          const int q = t + kT * u;
          bs[k & 1][q >> 6][q & 63] = rb[dd][u];
        }
        bf16x8 a;
        if (layer == 0) a = make_a(k, ra[dd], xa[dd]);
        else a = *(const bf16x8 *)&hs[lrow][16 * k + ak];
        if (k + kPf < nk) {
          const bf16x8 *src = (const bf16x8 *)(wp + (int64_t)(k + kPf) * 4096);
#pragma unroll
          for (int u = 0; u < kLd; u++) rb[dd][u] = src[t + kT * u];
          if (layer == 0) issue_a(k + kPf, ra[dd], xa[dd]);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; q++) acc[q] = mfma(a, bs[k & 1][q][lane], acc[q]);
      }
    }
    __syncthreads();  // every wavefront is done with hs (history indices or the layer's input)
    // bias + ReLU into LDS (this wavefront's own rows)
    const float *bias = m.p + (layer == 0 ? kOffB1 : layer == 1 ? kOffB2 : kOffB3);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int col = 32 * q + ar;
      const float b = bias[col];
#pragma unroll
      for (int e = 0; e < 16; e++)
        hs[32 * wv + acc_row(e, lane)][col] = (__bf16)fmaxf(acc[q][e] + b, 0.0f);
    }
    __syncthreads();
  }
  // output layer (FP32): 2 threads per row; thread j forms the partial sums of the 32-wide slices
  // 4 j .. 4 j + 3 and adds them in k_mlp_fwd's butterfly order
  const int r = t >> 1, j = t & 1;
  const float *w4 = m.p + kOffW4;
  float y[kClOut];
  {
    float ps[4][kClOut];
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
#pragma unroll
      for (int c = 0; c < kClOut; c++) ps[jj][c] = 0.0f;
      for (int q = 0; q < 32; q++) {
        const int k = 32 * (4 * j + jj) + q;
        const float hv = (float)hs[r][k];
#pragma unroll
        for (int c = 0; c < kClOut; c++) ps[jj][c] += hv * w4[c * kClHid + k];
      }
    }
#pragma unroll
    for (int c = 0; c < kClOut; c++) {
      y[c] = (ps[0][c] + ps[1][c]) + (ps[2][c] + ps[3][c]);
      y[c] += __shfl_xor(y[c], 1);
    }
  }
  const int32_t orow = r0 + r;
  if (j != 0 || orow >= n) return;
#pragma unroll
  for (int c = 0; c < kClOut; c++) y[c] += m.p[kOffB4 + c];
  if constexpr (!kGather) {
    if (yo) {
#pragma unroll
      for (int c = 0; c < kClOut; c++) yo[(int64_t)orow * kClOut + c] = y[c];
    }
  } else {
    finish_row(m, g, d, orow, y, sc[r][2], sc[r][3]);
  }
}

hipError_t mlp_forward_dataset(const MlpDev &m, const __bf16 *x, int32_t n, __bf16 *h1, __bf16 *h2,
                               __bf16 *h3, float *y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_fwd<false>, dim3((n + kClRows - 1) / kClRows), dim3(256), 0, s, m, x,
                     GatherSrc{}, n, h1, h2, h3, y, ApplyDst{});
  return hipGetLastError();
}

hipError_t mlp_forward_apply(const MlpDev &m, const GatherSrc &g, int32_t n, const ApplyDst &d,
                             int wide_waves, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (wide_waves == 4 || wide_waves == 8) {
    const int rows = 32 * wide_waves;  // blocks: a multiple of 8 (XCD-contiguous row ranges)
    const dim3 grid((n + 8 * rows - 1) / (8 * rows) * 8);
    if (wide_waves == 8)
      hipLaunchKernelGGL((k_mlp_fwd_wide<true, 8>), grid, dim3(512), 0, s, m, (const __bf16 *)nullptr,
                         g, n, (float *)nullptr, d);
    else
      hipLaunchKernelGGL((k_mlp_fwd_wide<true, 4>), grid, dim3(256), 0, s, m, (const __bf16 *)nullptr,
                         g, n, (float *)nullptr, d);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_mlp_fwd<true>, dim3((n + kClRows - 1) / kClRows), dim3(256), 0, s, m,
                     (const __bf16 *)nullptr, g, n, (__bf16 *)nullptr, (__bf16 *)nullptr,
                     (__bf16 *)nullptr, (float *)nullptr, d);
  return hipGetLastError();
}

hipError_t mlp_apply_deferred(const MlpDev &m, const GatherSrc &g, int32_t n, const ApplyDst &d,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (!d.defer) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_apply_deferred, dim3((n + 255) / 256), dim3(256), 0, s, m, g, n, d);
  return hipGetLastError();
}

// One wavefront per row: the row's 1600 inputs and its scales.
__global__ void __launch_bounds__(256) k_gather_rows(GatherSrc g, int32_t n, __bf16 *x,
                                                     float *scale) {
  const int lane = threadIdx.x & 63;
  const int32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  float su, sv;
  row_scale(g, r, su, sv);
  const float iu = safe_inv(su), iv = safe_inv(sv);
  for (int q = lane; q < kClIn / 8; q += 64)
    *(bf16x8 *)(x + (int64_t)r * kClIn + 8 * q) = gather8(g, r, 8 * q, iu, iv);
  if (lane == 0 && scale) {
    scale[2 * (int64_t)r] = su;
    scale[2 * (int64_t)r + 1] = sv;
  }
}

hipError_t gather_rows(const GatherSrc &g, int32_t n, __bf16 *x, float *scale, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_rows, dim3((n + 3) / 4), dim3(256), 0, s, g, n, x, scale);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Backward
// ---------------------------------------------------------------------------------------------

// C = A B^T over K: A is M x K, B is N x K (BF16, K contiguous), M and N multiples of 32, K of 16.
// Workgroup tile 64 x 64, one 32 x 32 tile per wavefront.
//  kMode 0: C (FP32, ldc) for rows m < m_lim.
//  kMode 1: O[m][n] = C if H[m][n] > 0 else 0, BF16 (the ReLU derivative of the layer whose
//           activation is H), both M x N.
template <int kMode>
__global__ void __launch_bounds__(256) k_gemm_nt(const __bf16 *A, int64_t lda, const __bf16 *B,
                                                 int64_t ldb, int M, int N, int K, float *C,
                                                 int64_t ldc, int m_lim, const __bf16 *H,
                                                 __bf16 *O) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 64 + 32 * (wv >> 1), n0 = blockIdx.x * 64 + 32 * (wv & 1);
  if (m0 >= M || n0 >= N) return;
  const __bf16 *pa = A + (int64_t)(m0 + (lane & 31)) * lda + 8 * (lane >> 5);
  const __bf16 *pb = B + (int64_t)(n0 + (lane & 31)) * ldb + 8 * (lane >> 5);
  f32x16 acc = {};
#pragma unroll 4
  for (int k = 0; k < K; k += 16) acc = mfma(*(const bf16x8 *)(pa + k), *(const bf16x8 *)(pb + k), acc);
  const int nn = n0 + (lane & 31);
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const int mm = m0 + acc_row(e, lane);
    if constexpr (kMode == 0) {
      if (mm < m_lim) C[(int64_t)mm * ldc + nn] = acc[e];
    } else {
      const int64_t q = (int64_t)mm * N + nn;
      O[q] = (__bf16)((float)H[q] > 0.0f ? acc[e] : 0.0f);
    }
  }
}

// out[c][r] = in[r][c]; rows and cols multiples of 64.
__global__ void __launch_bounds__(256) k_transpose(const __bf16 *in, int64_t rows, int64_t cols,
                                                   __bf16 *out) {
  __shared__ __bf16 tile[64][66];
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
  for (int i = t >> 6; i < 64; i += 4) tile[i][t & 63] = in[(r0 + i) * cols + c0 + (t & 63)];
  __syncthreads();
  for (int i = t >> 6; i < 64; i += 4) out[(c0 + i) * rows + r0 + (t & 63)] = tile[t & 63][i];
}

// Per row of the batch: loss term, dL/dy (FP32 n x 4 and BF16 transposed, rows 0..3 of 32 x n).
__global__ void __launch_bounds__(256) k_loss(const float *y, const float *tgt, const float *w,
                                              int32_t n, float inv_su, float inv_sv, float *dy,
                                              __bf16 *dyt, float *lrow) {
  const int32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n) return;
  const float wb = w[b];
  float l = 0.0f;
#pragma unroll
  for (int c = 0; c < kClOut; c++) {
    const float e = y[(int64_t)b * kClOut + c] - tgt[(int64_t)b * kClOut + c] * (c < 2 ? inv_su : inv_sv);
    l += wb * e * e;
    const float g = 2.0f * wb * e / (float)n;
    dy[(int64_t)b * kClOut + c] = g;
    dyt[(int64_t)c * n + b] = (__bf16)g;
  }
  lrow[b] = l / (float)n;
}

// dZ3[b][i] = (sum_c dy[b][c] W4[c][i]) where H3[b][i] > 0.
__global__ void __launch_bounds__(256) k_out_back(const float *dy, const float *w4,
                                                  const __bf16 *h3, int32_t n, __bf16 *dz3) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= (int64_t)n * kClHid) return;
  const int64_t b = q / kClHid;
  const int i = (int)(q % kClHid);
  float v = 0.0f;
#pragma unroll
  for (int c = 0; c < kClOut; c++) v += dy[b * kClOut + c] * w4[c * kClHid + i];
  dz3[q] = (__bf16)((float)h3[q] > 0.0f ? v : 0.0f);
}

// out[r] = sum over the row of in (rows x len, BF16), one wavefront per row.
__global__ void __launch_bounds__(256) k_rowsum(const __bf16 *in, int32_t rows, int32_t len,
                                                float *out) {
  const int lane = threadIdx.x & 63;
  const int32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= rows) return;
  float s = 0.0f;
  for (int32_t k = lane; k < len; k += 64) s += (float)in[(int64_t)r * len + k];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) out[r] = s;
}

__global__ void __launch_bounds__(1024) k_sum(const float *in, int32_t n, float *out) {
  __shared__ float part[1024];
  float s = 0.0f;
  for (int32_t k = threadIdx.x; k < n; k += 1024) s += in[k];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

// Adam update (when lr != 0) and the BF16 copies of W1..W3 ([out][in], and [in][out] for W2, W3).
__global__ void __launch_bounds__(256) k_adam(float *p, const float *g, float *m1, float *m2,
                                              float lr, float c1, float c2, __bf16 *w1, __bf16 *w2,
                                              __bf16 *w3, __bf16 *w2t, __bf16 *w3t, __bf16 *w1p,
                                              __bf16 *w2p, __bf16 *w3p) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= kClParams) return;
  float w = p[q];
  if (lr != 0.0f) {
    const float gq = g[q];
    const float a = 0.9f * m1[q] + 0.1f * gq;
    const float b = 0.999f * m2[q] + 0.001f * gq * gq;
    m1[q] = a;
    m2[q] = b;
    w -= lr * (a * c1) / (sqrtf(b * c2) + 1e-8f);
    p[q] = w;
  }
  if (q < kOffB1) {
    w1[q] = (__bf16)w;
    w1p[pack_index(q / kClIn, q % kClIn)] = (__bf16)w;
  } else if (q >= kOffW2 && q < kOffB2) {
    const int64_t j = q - kOffW2;
    w2[j] = (__bf16)w;
    w2t[(j % kClHid) * kClHid + j / kClHid] = (__bf16)w;
    w2p[pack_index(j / kClHid, j % kClHid)] = (__bf16)w;
  } else if (q >= kOffW3 && q < kOffB3) {
    const int64_t j = q - kOffW3;
    w3[j] = (__bf16)w;
    w3t[(j % kClHid) * kClHid + j / kClHid] = (__bf16)w;
    w3p[pack_index(j / kClHid, j % kClHid)] = (__bf16)w;
  }
}

// Batch rows drawn with replacement from samples [first, first + count).
// noise > 0: Gaussian noise of that standard deviation added to the history inputs (the first
// 49 x 32), so the network learns to be insensitive to perturbed histories.
__global__ void __launch_bounds__(256) k_draw(const __bf16 *dsx, const float *dsaux, int64_t first,
                                              int64_t count, uint64_t seed, int32_t n, __bf16 *xb,
                                              float *tgt, float *w, float noise) {
  const int lane = threadIdx.x & 63;
  const int32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const uint64_t hi = dev_mix(seed ^ (uint64_t)i * 0x2545F4914F6CDD1Dull);
  const int64_t src = first + (int64_t)(hi % (uint64_t)count);
  const bf16x8 *a = (const bf16x8 *)(dsx + src * kClIn);
  bf16x8 *b = (bf16x8 *)(xb + (int64_t)i * kClIn);
  for (int q = lane; q < kClIn / 8; q += 64) {
    bf16x8 v = a[q];
    if (noise > 0.0f && q < kClOffsets * 4) {
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        const uint64_t h = dev_mix(hi ^ ((uint64_t)(8 * q + e) << 40));
        const float u1 = ((float)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
        const float u2 = (float)((h >> 16) & 0xFFFFFF) * (1.0f / 16777216.0f);
        const float rr = noise * sqrtf(-2.0f * logf(u1));
        v[e] = (__bf16)((float)v[e] + rr * cosf(6.2831853f * u2));
        v[e + 1] = (__bf16)((float)v[e + 1] + rr * sinf(6.2831853f * u2));
      }
    }
    b[q] = v;
  }
  if (lane < kClOut) tgt[(int64_t)i * kClOut + lane] = dsaux[src * 8 + lane];
  if (lane == kClOut) w[i] = dsaux[src * 8 + kClOut];
}

// Validation: per row w |f - t/sigma|^2, w |t/sigma|^2, w (FP64 out).
__global__ void __launch_bounds__(256) k_eval_rows(const float *y, const float *aux, int32_t n,
                                                   float inv_su, float inv_sv, double *out3) {
  const int32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n) return;
  const double wb = aux[(int64_t)b * 8 + kClOut];
  double l = 0.0, base = 0.0;
  for (int c = 0; c < kClOut; c++) {
    const double t = (double)aux[(int64_t)b * 8 + c] * (c < 2 ? inv_su : inv_sv);
    const double e = (double)y[(int64_t)b * kClOut + c] - t;
    l += e * e;
    base += t * t;
  }
  out3[3 * (int64_t)b] = wb * l;
  out3[3 * (int64_t)b + 1] = wb * base;
  out3[3 * (int64_t)b + 2] = wb;
}

static dim3 blocks(int64_t n, int per) { return dim3((unsigned)((n + per - 1) / per)); }

}  // namespace lps

using namespace lps;

namespace {

#define CLO_TRY(clo, call)                                    \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return (clo)->hip_fail(e_, #call);  \
  } while (0)

double normal_from(uint64_t &state) {  // Box-Muller on splitmix64 draws
  state = splitmix64(state);
  const double u1 = ((double)(state >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  state = splitmix64(state);
  const double u2 = (double)(state >> 11) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

int32_t repack(lps_closure *c, float lr) {
  const float c1 = c->adam_t > 0 ? 1.0f / (1.0f - std::pow(0.9f, (float)c->adam_t)) : 1.0f;
  const float c2 = c->adam_t > 0 ? 1.0f / (1.0f - std::pow(0.999f, (float)c->adam_t)) : 1.0f;
  hipLaunchKernelGGL(k_adam, blocks(kClParams, 256), dim3(256), 0, c->stream, c->p, c->g, c->m1,
                     c->m2, lr, c1, c2, c->w1, c->w2, c->w3, c->w2t, c->w3t, c->w1p, c->w2p,
                     c->w3p);
  CLO_TRY(c, hipGetLastError());
  return LPS_OK;
}

int32_t ensure_batch(lps_closure *c, int32_t n) {
  if (n <= 0 || n % 64) return c->fail(LPS_ERR_ARG, "batch size must be a positive multiple of 64");
  if (c->tb >= n) return LPS_OK;
  for (void *q : {(void *)c->xb, (void *)c->xbt, (void *)c->h[0], (void *)c->h[1], (void *)c->h[2],
                  (void *)c->ht[0], (void *)c->ht[1], (void *)c->ht[2], (void *)c->dz[0],
                  (void *)c->dz[1], (void *)c->dz[2], (void *)c->dzt[0], (void *)c->dzt[1],
                  (void *)c->dzt[2], (void *)c->dyt, (void *)c->yb, (void *)c->tgt, (void *)c->wgt,
                  (void *)c->dy, (void *)c->lrow}) {
    if (!q) continue;
    (void)hipFree(q);
    c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), q), c->allocs.end());
  }
  const size_t N = (size_t)n;
  CLO_TRY(c, c->alloc(&c->xb, N * kClIn));
  CLO_TRY(c, c->alloc(&c->xbt, N * kClIn));
  for (int l = 0; l < 3; l++) {
    CLO_TRY(c, c->alloc(&c->h[l], N * kClHid));
    CLO_TRY(c, c->alloc(&c->ht[l], N * kClHid));
    CLO_TRY(c, c->alloc(&c->dz[l], N * kClHid));
    CLO_TRY(c, c->alloc(&c->dzt[l], N * kClHid));
  }
  CLO_TRY(c, c->alloc(&c->dyt, N * 32));
  CLO_TRY(c, hipMemsetAsync(c->dyt, 0, sizeof(__bf16) * N * 32, c->stream));
  CLO_TRY(c, c->alloc(&c->yb, N * kClOut));
  CLO_TRY(c, c->alloc(&c->tgt, N * kClOut));
  CLO_TRY(c, c->alloc(&c->wgt, N));
  CLO_TRY(c, c->alloc(&c->dy, N * kClOut));
  CLO_TRY(c, c->alloc(&c->lrow, N));
  c->tb = n;
  return LPS_OK;
}

hipError_t gemm_nt(const __bf16 *A, int64_t lda, const __bf16 *B, int64_t ldb, int M, int N, int K,
                   float *C, int64_t ldc, int m_lim, hipStream_t s) {
  hipLaunchKernelGGL(k_gemm_nt<0>, dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, s, A, lda, B,
                     ldb, M, N, K, C, ldc, m_lim, (const __bf16 *)nullptr, (__bf16 *)nullptr);
  return hipGetLastError();
}
hipError_t gemm_nt_mask(const __bf16 *A, int64_t lda, const __bf16 *B, int64_t ldb, int M, int N,
                        int K, const __bf16 *H, __bf16 *O, hipStream_t s) {
  hipLaunchKernelGGL(k_gemm_nt<1>, dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, s, A, lda, B,
                     ldb, M, N, K, (float *)nullptr, (int64_t)0, 0, H, O);
  return hipGetLastError();
}
hipError_t transpose(const __bf16 *in, int64_t rows, int64_t cols, __bf16 *out, hipStream_t s) {
  hipLaunchKernelGGL(k_transpose, dim3((unsigned)(cols / 64), (unsigned)(rows / 64)), dim3(256), 0,
                     s, in, rows, cols, out);
  return hipGetLastError();
}

// Forward, loss and gradient of the batch in c->xb / c->tgt / c->wgt (n rows), the loss into
// *loss_dst (device), then the Adam step.
int32_t train_step(lps_closure *c, int32_t n, float lr, float *loss_dst) {
  const hipStream_t s = c->stream;
  const MlpDev m = c->dev();
  CLO_TRY(c, mlp_forward_dataset(m, c->xb, n, c->h[0], c->h[1], c->h[2], c->yb, s));
  hipLaunchKernelGGL(k_loss, blocks(n, 256), dim3(256), 0, s, c->yb, c->tgt, c->wgt, n,
                     1.0f / c->sigma_u, 1.0f / c->sigma_v, c->dy, c->dyt, c->lrow);
  CLO_TRY(c, hipGetLastError());
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, s, c->lrow, n, loss_dst);
  CLO_TRY(c, hipGetLastError());
  hipLaunchKernelGGL(k_out_back, blocks((int64_t)n * kClHid, 256), dim3(256), 0, s, c->dy,
                     c->p + kOffW4, c->h[2], n, c->dz[2]);
  CLO_TRY(c, hipGetLastError());
  // transposed activations and input (K = batch for the weight gradients)
  for (int l = 0; l < 3; l++) CLO_TRY(c, transpose(c->h[l], n, kClHid, c->ht[l], s));
  CLO_TRY(c, transpose(c->xb, n, kClIn, c->xbt, s));
  // output layer: dW4 = dY^T H3, db4
  CLO_TRY(c, gemm_nt(c->dyt, n, c->ht[2], n, 32, kClHid, n, c->g + kOffW4, kClHid, kClOut, s));
  hipLaunchKernelGGL(k_rowsum, dim3(1), dim3(256), 0, s, c->dyt, kClOut, n, c->g + kOffB4);
  CLO_TRY(c, hipGetLastError());
  // hidden layers, last to first: dW = dZ^T H_prev, db = sum dZ, dZ_prev = (dZ W) masked
  const int64_t offW[3] = {kOffW1, kOffW2, kOffW3}, offB[3] = {kOffB1, kOffB2, kOffB3};
  const __bf16 *wt[3] = {nullptr, c->w2t, c->w3t};
  for (int l = 2; l >= 0; l--) {
    CLO_TRY(c, transpose(c->dz[l], n, kClHid, c->dzt[l], s));
    const __bf16 *prev_t = l > 0 ? c->ht[l - 1] : c->xbt;
    const int in = l > 0 ? kClHid : kClIn;
    CLO_TRY(c, gemm_nt(c->dzt[l], n, prev_t, n, kClHid, in, n, c->g + offW[l], in, kClHid, s));
    hipLaunchKernelGGL(k_rowsum, blocks(kClHid, 4), dim3(256), 0, s, c->dzt[l], kClHid, n,
                       c->g + offB[l]);
    CLO_TRY(c, hipGetLastError());
    if (l > 0)
      CLO_TRY(c, gemm_nt_mask(c->dz[l], kClHid, wt[l], kClHid, n, kClHid, kClHid, c->h[l - 1],
                              c->dz[l - 1], s));
  }
  if (lr != 0.0f) c->adam_t++;
  return repack(c, lr);
}

}  // namespace

extern "C" {

int64_t lps_closure_param_count(void) { return kClParams; }

const char *lps_closure_error(const lps_closure *clo) { return clo ? clo->err.c_str() : ""; }

void lps_closure_destroy(lps_closure *clo) {
  if (!clo) return;
  (void)hipSetDevice(clo->device);
  if (clo->stream) (void)hipStreamSynchronize(clo->stream);
  for (void *q : clo->allocs) (void)hipFree(q);
  if (clo->stream) (void)hipStreamDestroy(clo->stream);
  delete clo;
}

// Staging area of lps_closure_forward_noalloc: the inputs (BF16) and outputs (FP32) of at most
// kNoAllocRows rows. A file-scope device array: every GPU the library runs on holds its own copy
// until the process ends, and all calls on one GPU share it (calls that overlap on one GPU race).
// Nothing in it needs to start at zero: a call writes every byte it reads.
constexpr int64_t kNoAllocRows = 2048;
constexpr size_t kNoAllocXBytes = (size_t)kNoAllocRows * kClIn * sizeof(__bf16);
constexpr size_t kNoAllocBytes = kNoAllocXBytes + (size_t)kNoAllocRows * kClOut * sizeof(float);
static_assert(kNoAllocXBytes % 256 == 0, "output block alignment");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

int32_t lps_closure_create(int32_t device, uint64_t seed, lps_closure **out) {
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return LPS_ERR_HIP;
  if (device < 0 || device >= ndev) return LPS_ERR_ARG;
  lps_closure *c = new lps_closure();
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = c->alloc(&c->p, kClParams);
  if (e == hipSuccess) e = c->alloc(&c->g, kClParams);
  if (e == hipSuccess) e = c->alloc(&c->m1, kClParams);
  if (e == hipSuccess) e = c->alloc(&c->m2, kClParams);
  if (e == hipSuccess) e = c->alloc(&c->w1, (size_t)kClHid * kClIn);
  if (e == hipSuccess) e = c->alloc(&c->w2, (size_t)kClHid * kClHid);
  if (e == hipSuccess) e = c->alloc(&c->w3, (size_t)kClHid * kClHid);
  if (e == hipSuccess) e = c->alloc(&c->w2t, (size_t)kClHid * kClHid);
  if (e == hipSuccess) e = c->alloc(&c->w3t, (size_t)kClHid * kClHid);
  if (e == hipSuccess) e = c->alloc(&c->w1p, (size_t)kClHid * kClIn);
  if (e == hipSuccess) e = c->alloc(&c->w2p, (size_t)kClHid * kClHid);
  if (e == hipSuccess) e = c->alloc(&c->w3p, (size_t)kClHid * kClHid);
  if (e == hipSuccess) e = c->alloc(&c->ds_count, 1);
  if (e == hipSuccess) e = c->alloc(&c->loss_d, 1);
  if (e == hipSuccess) e = hipMemsetAsync(c->ds_count, 0, sizeof(unsigned long long), c->stream);
  // this device's copy of the no-alloc staging area (the device is current here)
  if (e == hipSuccess) e = hipGetSymbolAddress((void **)&c->scratch, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    lps_closure_destroy(c);
    return LPS_ERR_HIP;
  }
  // He-normal hidden layers, small output layer, zero biases
  std::vector<float> p(kClParams, 0.0f);
  uint64_t st = seed ^ 0xC105u;
  auto fill = [&](int64_t off, int64_t count, double sd) {
    for (int64_t q = 0; q < count; q++) p[off + q] = (float)(sd * normal_from(st));
  };
  fill(kOffW1, (int64_t)kClHid * kClIn, std::sqrt(2.0 / kClIn));
  fill(kOffW2, (int64_t)kClHid * kClHid, std::sqrt(2.0 / kClHid));
  fill(kOffW3, (int64_t)kClHid * kClHid, std::sqrt(2.0 / kClHid));
  fill(kOffW4, (int64_t)kClOut * kClHid, 1e-3 / std::sqrt((double)kClHid));
  int32_t rc = lps_closure_set_params(c, p.data());
  if (rc != LPS_OK) {
    lps_closure_destroy(c);
    return rc;
  }
  *out = c;
  return LPS_OK;
}

int32_t lps_closure_get_params(lps_closure *c, float *params) {
  CLO_TRY(c, hipSetDevice(c->device));
  CLO_TRY(c, hipMemcpyAsync(params, c->p, sizeof(float) * kClParams, hipMemcpyDeviceToHost, c->stream));
  CLO_TRY(c, hipStreamSynchronize(c->stream));
  return LPS_OK;
}

int32_t lps_closure_set_params(lps_closure *c, const float *params) {
  CLO_TRY(c, hipSetDevice(c->device));
  CLO_TRY(c, hipMemcpyAsync(c->p, params, sizeof(float) * kClParams, hipMemcpyHostToDevice, c->stream));
  CLO_TRY(c, hipMemsetAsync(c->g, 0, sizeof(float) * kClParams, c->stream));
  CLO_TRY(c, hipMemsetAsync(c->m1, 0, sizeof(float) * kClParams, c->stream));
  CLO_TRY(c, hipMemsetAsync(c->m2, 0, sizeof(float) * kClParams, c->stream));
  c->adam_t = 0;
  int32_t rc = repack(c, 0.0f);
  if (rc != LPS_OK) return rc;
  CLO_TRY(c, hipStreamSynchronize(c->stream));
  return LPS_OK;
}

int32_t lps_closure_set_scales(lps_closure *c, float sigma_u, float sigma_v) {
  if (!(sigma_u > 0.0f) || !(sigma_v > 0.0f)) return c->fail(LPS_ERR_ARG, "scales must be positive");
  c->sigma_u = sigma_u;
  c->sigma_v = sigma_v;
  return LPS_OK;
}

int32_t lps_closure_get_scales(lps_closure *c, float *sigma2) {
  sigma2[0] = c->sigma_u;
  sigma2[1] = c->sigma_v;
  return LPS_OK;
}

int32_t lps_closure_forward(lps_closure *c, const uint16_t *x, int64_t n, float *out) {
  if (n < 0 || n > INT32_MAX / kClIn) return c->fail(LPS_ERR_ARG, "row count out of range");
  if (n == 0) return LPS_OK;
  CLO_TRY(c, hipSetDevice(c->device));
  __bf16 *dx = nullptr;
  float *dy = nullptr;
  hipError_t e = hipMalloc((void **)&dx, sizeof(__bf16) * n * kClIn);
  if (e == hipSuccess) e = hipMalloc((void **)&dy, sizeof(float) * n * kClOut);
  if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sizeof(__bf16) * n * kClIn, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = mlp_forward_dataset(c->dev(), dx, (int32_t)n, nullptr, nullptr, nullptr, dy, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dy, sizeof(float) * n * kClOut, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dx);
  (void)hipFree(dy);
  if (e != hipSuccess) return c->hip_fail(e, "lps_closure_forward");
  return LPS_OK;
}

int32_t lps_closure_forward_noalloc(lps_closure *c, const uint16_t *x, int64_t n, float *out) {
  if (n < 0 || n > kNoAllocRows)
    return c->fail(LPS_ERR_ARG, "lps_closure_forward_noalloc: n must be 0 .. " +
                                    std::to_string(kNoAllocRows) +
                                    " rows (the size of its staging area)");
  if (n == 0) return LPS_OK;
  if (!x || !out) return c->fail(LPS_ERR_ARG, "lps_closure_forward_noalloc: x and out are required");
  if (!c->scratch) return c->fail(LPS_ERR_HIP, "lps_closure_forward_noalloc: no staging area");
  CLO_TRY(c, hipSetDevice(c->device));
  __bf16 *dx = (__bf16 *)c->scratch;
  float *dy = (float *)(c->scratch + kNoAllocXBytes);
  hipError_t e = hipMemcpyAsync(dx, x, sizeof(__bf16) * n * kClIn, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = mlp_forward_dataset(c->dev(), dx, (int32_t)n, nullptr, nullptr, nullptr, dy, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dy, sizeof(float) * n * kClOut, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return c->hip_fail(e, "lps_closure_forward_noalloc");
  return LPS_OK;
}

int32_t lps_closure_train_batch(lps_closure *c, const uint16_t *x, const float *target,
                                const float *weight, int32_t n, float lr, float *loss,
                                float *grads) {
  CLO_TRY(c, hipSetDevice(c->device));
  int32_t rc = ensure_batch(c, n);
  if (rc != LPS_OK) return rc;
  CLO_TRY(c, hipMemcpyAsync(c->xb, x, sizeof(__bf16) * n * kClIn, hipMemcpyHostToDevice, c->stream));
  CLO_TRY(c, hipMemcpyAsync(c->tgt, target, sizeof(float) * n * kClOut, hipMemcpyHostToDevice, c->stream));
  CLO_TRY(c, hipMemcpyAsync(c->wgt, weight, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  rc = train_step(c, n, lr, c->loss_d);
  if (rc != LPS_OK) return rc;
  if (loss) CLO_TRY(c, hipMemcpyAsync(loss, c->loss_d, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  if (grads) CLO_TRY(c, hipMemcpyAsync(grads, c->g, sizeof(float) * kClParams, hipMemcpyDeviceToHost, c->stream));
  CLO_TRY(c, hipStreamSynchronize(c->stream));
  return LPS_OK;
}

int32_t lps_closure_dataset_reserve(lps_closure *c, int64_t capacity) {
  if (capacity < 0) return c->fail(LPS_ERR_ARG, "capacity must be >= 0");
  CLO_TRY(c, hipSetDevice(c->device));
  CLO_TRY(c, hipStreamSynchronize(c->stream));
  if (capacity <= c->ds_cap) return LPS_OK;
  __bf16 *nx = nullptr;
  float *na = nullptr;
  CLO_TRY(c, hipMalloc((void **)&nx, sizeof(__bf16) * std::max<int64_t>(capacity, 1) * kClIn));
  hipError_t e = hipMalloc((void **)&na, sizeof(float) * std::max<int64_t>(capacity, 1) * 8);
  if (e != hipSuccess) {
    (void)hipFree(nx);
    return c->hip_fail(e, "hipMalloc (samples)");
  }
  const int64_t keep = std::min<int64_t>(lps_closure_dataset_count(c), c->ds_cap);
  if (keep > 0) {
    CLO_TRY(c, hipMemcpyAsync(nx, c->ds_x, sizeof(__bf16) * keep * kClIn, hipMemcpyDeviceToDevice, c->stream));
    CLO_TRY(c, hipMemcpyAsync(na, c->ds_aux, sizeof(float) * keep * 8, hipMemcpyDeviceToDevice, c->stream));
    CLO_TRY(c, hipStreamSynchronize(c->stream));
  }
  if (c->ds_x) (void)hipFree(c->ds_x);
  if (c->ds_aux) (void)hipFree(c->ds_aux);
  c->ds_x = nx;
  c->ds_aux = na;
  c->ds_cap = capacity;
  return LPS_OK;
}

int64_t lps_closure_dataset_count(lps_closure *c) {
  unsigned long long n = 0;
  if (hipSetDevice(c->device) != hipSuccess) return LPS_ERR_HIP;
  if (hipMemcpyAsync(&n, c->ds_count, sizeof(n), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return LPS_ERR_HIP;
  return (int64_t)n;
}

int32_t lps_closure_dataset_clear(lps_closure *c) {
  CLO_TRY(c, hipSetDevice(c->device));
  CLO_TRY(c, hipMemsetAsync(c->ds_count, 0, sizeof(unsigned long long), c->stream));
  CLO_TRY(c, hipStreamSynchronize(c->stream));
  return LPS_OK;
}

int32_t lps_closure_dataset_append(lps_closure *c, const uint16_t *x, const float *aux, int64_t n) {
  const int64_t have = lps_closure_dataset_count(c);
  if (have < 0) return c->fail(LPS_ERR_HIP, "reading the sample count failed");
  if (n < 0 || have + n > c->ds_cap) return c->fail(LPS_ERR_ARG, "samples beyond the capacity");
  if (n == 0) return LPS_OK;
  CLO_TRY(c, hipMemcpyAsync(c->ds_x + have * kClIn, x, sizeof(__bf16) * n * kClIn, hipMemcpyHostToDevice, c->stream));
  CLO_TRY(c, hipMemcpyAsync(c->ds_aux + have * 8, aux, sizeof(float) * n * 8, hipMemcpyHostToDevice, c->stream));
  const unsigned long long nn = (unsigned long long)(have + n);
  CLO_TRY(c, hipMemcpyAsync(c->ds_count, &nn, sizeof(nn), hipMemcpyHostToDevice, c->stream));
  CLO_TRY(c, hipStreamSynchronize(c->stream));
  return LPS_OK;
}

int32_t lps_closure_dataset_read(lps_closure *c, int64_t first, int64_t n, uint16_t *x, float *aux) {
  const int64_t have = std::min<int64_t>(lps_closure_dataset_count(c), c->ds_cap);
  if (first < 0 || n < 0 || first + n > have) return c->fail(LPS_ERR_ARG, "sample range out of bounds");
  if (n == 0) return LPS_OK;
  if (x) CLO_TRY(c, hipMemcpyAsync(x, c->ds_x + first * kClIn, sizeof(__bf16) * n * kClIn, hipMemcpyDeviceToHost, c->stream));
  if (aux) CLO_TRY(c, hipMemcpyAsync(aux, c->ds_aux + first * 8, sizeof(float) * n * 8, hipMemcpyDeviceToHost, c->stream));
  CLO_TRY(c, hipStreamSynchronize(c->stream));
  return LPS_OK;
}

int32_t lps_closure_train(lps_closure *c, int32_t nsteps, int32_t batch, float lr0, float lr1,
                          int64_t first, int64_t count, uint64_t seed, float input_noise,
                          float *loss_curve) {
  const int64_t have = std::min<int64_t>(lps_closure_dataset_count(c), c->ds_cap);
  if (nsteps < 0 || count <= 0 || first < 0 || first + count > have)
    return c->fail(LPS_ERR_ARG, "training range out of bounds");
  CLO_TRY(c, hipSetDevice(c->device));
  int32_t rc = ensure_batch(c, batch);
  if (rc != LPS_OK) return rc;
  float *curve = nullptr;
  CLO_TRY(c, hipMalloc((void **)&curve, sizeof(float) * std::max(nsteps, 1)));
  for (int32_t it = 0; it < nsteps && rc == LPS_OK; it++) {
    const double f = nsteps > 1 ? (double)it / (nsteps - 1) : 0.0;
    const float lr = (float)(lr1 + 0.5 * (lr0 - lr1) * (1.0 + std::cos(3.141592653589793 * f)));
    const uint64_t sd = splitmix64(seed ^ splitmix64((uint64_t)it + 1));
    hipLaunchKernelGGL(k_draw, blocks(batch, 4), dim3(256), 0, c->stream, c->ds_x, c->ds_aux, first,
                       count, sd, batch, c->xb, c->tgt, c->wgt, input_noise);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = c->hip_fail(e, "k_draw");
    else rc = train_step(c, batch, lr, curve + it);
  }
  if (rc == LPS_OK && loss_curve && nsteps > 0) {
    hipError_t e = hipMemcpyAsync(loss_curve, curve, sizeof(float) * nsteps, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) rc = c->hip_fail(e, "copy of the loss curve");
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  (void)hipFree(curve);
  if (rc == LPS_OK && e != hipSuccess) rc = c->hip_fail(e, "lps_closure_train");
  return rc;
}

int32_t lps_closure_evaluate(lps_closure *c, int64_t first, int64_t count, double *out3) {
  const int64_t have = std::min<int64_t>(lps_closure_dataset_count(c), c->ds_cap);
  if (count < 0 || first < 0 || first + count > have) return c->fail(LPS_ERR_ARG, "range out of bounds");
  out3[0] = out3[1] = out3[2] = 0.0;
  CLO_TRY(c, hipSetDevice(c->device));
  const int32_t chunk = 65536;
  float *y = nullptr;
  double *rows = nullptr;
  CLO_TRY(c, hipMalloc((void **)&y, sizeof(float) * chunk * kClOut));
  hipError_t e = hipMalloc((void **)&rows, sizeof(double) * chunk * 3);
  std::vector<double> h((size_t)chunk * 3);
  for (int64_t q = first; q < first + count && e == hipSuccess; q += chunk) {
    const int32_t n = (int32_t)std::min<int64_t>(chunk, first + count - q);
    e = mlp_forward_dataset(c->dev(), c->ds_x + q * kClIn, n, nullptr, nullptr, nullptr, y, c->stream);
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(k_eval_rows, blocks(n, 256), dim3(256), 0, c->stream, y, c->ds_aux + q * 8,
                       n, 1.0f / c->sigma_u, 1.0f / c->sigma_v, rows);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), rows, sizeof(double) * n * 3, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    for (int32_t r = 0; r < n && e == hipSuccess; r++)
      for (int k = 0; k < 3; k++) out3[k] += h[3 * (size_t)r + k];
  }
  (void)hipFree(y);
  (void)hipFree(rows);
  if (e != hipSuccess) return c->hip_fail(e, "lps_closure_evaluate");
  return LPS_OK;
}

}  // extern "C"
