// Learned ABC closure, coupling to the time loop: the per-node history of u and v around the ABC
// rows, the correction applied inside the step after the least-squares G_u / G_v update
// (src/main.cpp:883-974 is the update it corrects), and teacher-forced sample collection from a
// larger, reflection-free domain. No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/lps_amd.h"
#include "lps_closure.h"
#include "lps_ctx.h"

namespace lps {

// The closure attached to a context (or, with clo == nullptr, the history alone).
struct ClosureAttach {
  lps_closure *clo = nullptr;
  int32_t n_rows = 0, n_hist = 0, n_inst = 1;
  int32_t *hist_slot = nullptr;  // n_hist: slot of each history node
  int32_t *hidx = nullptr;       // n_rows x 49
  int8_t *xf = nullptr;          // n_rows
  float *feat = nullptr;         // n_rows x 4
  int32_t *row_slot = nullptr, *row_inst = nullptr;
  float *gain = nullptr;                            // n_inst
  float *gate = nullptr;                            // n_inst: no correction where s_u < gate
  float *hist[2] = {nullptr, nullptr}, *mx[2] = {nullptr, nullptr};  // by step parity
  float *dbg = nullptr;                             // n_rows x 8, last applied correction
  float *defer = nullptr;  // n_rows x 8: the network's outputs and scales of the step, to be applied
  // per-instance gains: as set by the caller, and as in force (the guard zeroes them)
  std::vector<float> gain_user, gain_now;
  std::vector<double> e_prev;      // energy at the previous output record
  std::vector<int64_t> off_step;   // step of the fallback to least squares, -1 none
  double guard_tol = 1e-3;         // <= 0: no guard
  int64_t wide_min = 16384;        // row count from which the wide forward kernel runs
  int wide_waves = 8;              // its wavefronts per workgroup, 4 or 8
  std::vector<void *> allocs;
  template <class T>
  hipError_t alloc(T **q, size_t count) {
    hipError_t e = hipMalloc((void **)q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) allocs.push_back(*q);
    return e;
  }
  template <class T>
  hipError_t upload(T **q, const std::vector<T> &h, hipStream_t s) {
    hipError_t e = alloc(q, h.size());
    if (e == hipSuccess && !h.empty())
      e = hipMemcpyAsync(*q, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s);
    return e;
  }
  ~ClosureAttach() {
    for (void *q : allocs) (void)hipFree(q);
  }
  GatherSrc src(int par) const { return GatherSrc{hist[par], mx[par], hidx, xf, feat}; }
};

// History shift of step n: record = (U^n, V^n, the 7 steps before), and its max |u|, max |v|.
// U^n of an interior node (kind 1) comes from the prediction ring (rux, ruy).
__global__ void __launch_bounds__(256) k_hist_shift(int32_t n_hist, const int32_t *slot,
                                                    const int8_t *kind, const double *rux,
                                                    const double *ruy, const double *ux,
                                                    const double *uy, const double *vx,
                                                    const double *vy, const float *prev,
                                                    float *cur, float *mx) {
  const int32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n_hist) return;
  const int32_t i = slot[b];
  if (kind[i] == 1) {
    ux = rux;
    uy = ruy;
  }
  const float4 *p = (const float4 *)(prev + (int64_t)b * kClRec);
  float4 *q = (float4 *)(cur + (int64_t)b * kClRec);
  float mu = 0.0f, mv = 0.0f;
#pragma unroll
  for (int half = 0; half < 2; half++) {
    const float nx = (float)(half == 0 ? ux[i] : vx[i]), ny = (float)(half == 0 ? uy[i] : vy[i]);
    float4 in[4], out[4];
#pragma unroll
    for (int k = 0; k < 4; k++) in[k] = p[4 * half + k];
    // shift by one (x, y) pair: out pair h = in pair h - 1, out pair 0 = current
    out[0] = make_float4(nx, ny, in[0].x, in[0].y);
#pragma unroll
    for (int k = 1; k < 4; k++) out[k] = make_float4(in[k - 1].z, in[k - 1].w, in[k].x, in[k].y);
    float m = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      q[4 * half + k] = out[k];
      m = fmaxf(m, fmaxf(fmaxf(fabsf(out[k].x), fabsf(out[k].y)), fmaxf(fabsf(out[k].z), fabsf(out[k].w))));
    }
    if (half == 0) mu = m;
    else mv = m;
  }
  mx[2 * (int64_t)b] = mu;
  mx[2 * (int64_t)b + 1] = mv;
}

hipError_t closure_step_hook(lps_ctx *c, const StepState &st, hipStream_t s, bool defer) {
  ClosureAttach *A = c->clo;
  const int par = (int)(c->step & 1);
  if (A->n_hist > 0) {
    hipLaunchKernelGGL(k_hist_shift, dim3((A->n_hist + 255) / 256), dim3(256), 0, s, A->n_hist,
                       A->hist_slot, c->kind, c->ring_u(c->step, 0).x, c->ring_u(c->step, 0).y,
                       st.ux, st.uy, st.vx, st.vy, A->hist[par ^ 1], A->hist[par],
                       A->mx[par]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (!A->clo || A->n_rows == 0) return hipSuccess;
  const ApplyDst d{A->row_slot, A->row_inst, A->gain, A->gate, st.ux1, st.uy1, st.vx1, st.vy1,
                   A->dbg, defer ? A->defer : nullptr};
  // timing mode 3: the step's two timing events around the network kernel, on its stream
  const unsigned evf = c->capturing ? hipEventRecordExternal : 0;
  hipError_t e = hipSuccess;
  if (c->clo_ev) e = hipEventRecordWithFlags(c->clo_ev[1], s, evf);
  if (e == hipSuccess)
    e = mlp_forward_apply(A->clo->dev(), A->src(par), A->n_rows, d,
                          A->n_rows >= A->wide_min ? A->wide_waves : 0, s);
  if (e == hipSuccess && c->clo_ev) e = hipEventRecordWithFlags(c->clo_ev[2], s, evf);
  return e;
}

hipError_t closure_apply_hook(lps_ctx *c, const StepState &st, hipStream_t s) {
  ClosureAttach *A = c->clo;
  if (!A->clo || A->n_rows == 0) return hipSuccess;
  const ApplyDst d{A->row_slot, A->row_inst, A->gain, A->gate, st.ux1, st.uy1, st.vx1, st.vy1,
                   A->dbg, A->defer};
  return mlp_apply_deferred(A->clo->dev(), A->src((int)(c->step & 1)), A->n_rows, d, s);
}

hipError_t closure_reset_hook(lps_ctx *c) {
  ClosureAttach *A = c->clo;
  A->gain_now = A->gain_user;
  A->e_prev.assign(A->n_inst, -1.0);
  A->off_step.assign(A->n_inst, -1);
  if (!A->gain_now.empty()) {
    hipError_t e = hipMemcpyAsync(A->gain, A->gain_now.data(), sizeof(float) * A->n_inst,
                                  hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return e;
  }
  const size_t hb = sizeof(float) * std::max<int64_t>(A->n_hist, 1);
  for (int q = 0; q < 2; q++) {
    hipError_t e = hipMemsetAsync(A->hist[q], 0, hb * kClRec, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(A->mx[q], 0, hb * 2, c->stream);
    if (e != hipSuccess) return e;
  }
  if (A->dbg) return hipMemsetAsync(A->dbg, 0, sizeof(float) * 8 * std::max(A->n_rows, 1), c->stream);
  return hipSuccess;
}

int32_t closure_record_hook(lps_ctx *c, const lps_diag *d, int32_t count) {
  ClosureAttach *A = c->clo;
  if (!A->clo || count != A->n_inst) return LPS_OK;
  bool changed = false;
  for (int32_t i = 0; i < count; i++) {
    const double E0 = c->batch ? c->E0_inst[i] : c->E0;
    if (A->guard_tol > 0.0 && A->gain_now[i] != 0.0f && A->e_prev[i] >= 0.0 && E0 > 0.0 &&
        !(d[i].energy <= A->e_prev[i] + A->guard_tol * E0)) {
      A->gain_now[i] = 0.0f;
      A->off_step[i] = c->step;
      changed = true;
    }
    A->e_prev[i] = d[i].energy;
  }
  if (changed) {
    CTX_TRY(c, hipMemcpyAsync(A->gain, A->gain_now.data(), sizeof(float) * A->n_inst,
                              hipMemcpyHostToDevice, c->stream));
    CTX_TRY(c, hipStreamSynchronize(c->stream));
  }
  return LPS_OK;
}

void closure_free(lps_ctx *c) {
  delete c->clo;
  c->clo = nullptr;
}

// ---------------------------------------------------------------------------------------------
// Row tables
// ---------------------------------------------------------------------------------------------

// The 49 lattice offsets |o|^2 <= 16, x-major.
static std::vector<std::pair<int, int>> canon_offsets() {
  std::vector<std::pair<int, int>> o;
  for (int x = -4; x <= 4; x++)
    for (int y = -4; y <= 4; y++)
      if (x * x + y * y <= 16) o.push_back({x, y});
  return o;
}

// Frame code of a row whose outward normal is (nx, ny): the canonical normal is
// (max(|nx|, |ny|), min(|nx|, |ny|)).
static int frame_code(double nx, double ny) {
  const double ax = std::fabs(nx), ay = std::fabs(ny);
  if (ax >= ay) return (nx < 0 ? 2 : 0) | (ny < 0 ? 4 : 0);
  return 1 | (ny < 0 ? 2 : 0) | (nx < 0 ? 4 : 0);
}

struct Grid {  // lattice position -> slot over a bounding box
  int32_t x0 = 0, y0 = 0, nx = 0, ny = 0;
  std::vector<int32_t> v;
  void init(const std::vector<int32_t> &ix, const std::vector<int32_t> &iy) {
    if (ix.empty()) return;
    int32_t x1 = ix[0], y1 = iy[0];
    x0 = ix[0];
    y0 = iy[0];
    for (size_t k = 0; k < ix.size(); k++) {
      x0 = std::min(x0, ix[k]); x1 = std::max(x1, ix[k]);
      y0 = std::min(y0, iy[k]); y1 = std::max(y1, iy[k]);
    }
    nx = x1 - x0 + 1;
    ny = y1 - y0 + 1;
    v.assign((size_t)nx * ny, -1);
  }
  void set(int32_t ix, int32_t iy, int32_t s) { v[(size_t)(ix - x0) * ny + (iy - y0)] = s; }
  int32_t get(int32_t ix, int32_t iy) const {
    ix -= x0;
    iy -= y0;
    if (ix < 0 || iy < 0 || ix >= nx || iy >= ny) return -1;
    return v[(size_t)ix * ny + iy];
  }
};

struct RowIn {
  int32_t ix, iy;
  double x, y, nu, R;
};

struct Tables {
  std::vector<int32_t> hist_slot, hidx;
  std::vector<int8_t> xf;
  std::vector<float> feat;
};

// present(ix, iy) -> bool, slot_of(ix, iy) -> slot of the context that keeps the history.
template <class Present, class SlotOf>
static bool build_rows(const std::vector<RowIn> &rows, double delta, Present present,
                       SlotOf slot_of, Tables &t) {
  static const std::vector<std::pair<int, int>> off = canon_offsets();
  std::unordered_map<int32_t, int32_t> hmap;
  const size_t n = rows.size();
  t.hidx.assign(n * kClOffsets, -1);
  t.xf.assign(n, 0);
  t.feat.assign(n * kClFeat, 0.0f);
  for (size_t r = 0; r < n; r++) {
    const RowIn &q = rows[r];
    const double rr = std::sqrt(q.x * q.x + q.y * q.y);
    const double nx = rr > 0 ? q.x / rr : 1.0, ny = rr > 0 ? q.y / rr : 0.0;
    const int f = frame_code(nx, ny);
    t.xf[r] = (int8_t)f;
    t.feat[r * kClFeat + 0] = (float)((q.nu - 0.3) / 0.15);
    t.feat[r * kClFeat + 1] = (float)std::max(std::fabs(nx), std::fabs(ny));
    t.feat[r * kClFeat + 2] = (float)std::min(std::fabs(nx), std::fabs(ny));
    t.feat[r * kClFeat + 3] = (float)((q.R - rr) / delta);
    for (int j = 0; j < kClOffsets; j++) {
      double px, py;
      from_frame(f, (double)off[j].first, (double)off[j].second, px, py);
      const int32_t jx = q.ix + (int32_t)std::lround(px), jy = q.iy + (int32_t)std::lround(py);
      if (!present(jx, jy)) continue;
      const int32_t s = slot_of(jx, jy);
      if (s < 0) return false;
      auto it = hmap.find(s);
      int32_t h;
      if (it == hmap.end()) {
        h = (int32_t)t.hist_slot.size();
        hmap.emplace(s, h);
        t.hist_slot.push_back(s);
      } else {
        h = it->second;
      }
      t.hidx[r * kClOffsets + j] = h;
    }
  }
  return true;
}

template <class T>
static int32_t download(lps_ctx *c, std::vector<T> &h, const T *d, size_t n) {
  h.resize(n);
  if (n == 0) return LPS_OK;
  CTX_TRY(c, hipMemcpyAsync(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  return LPS_OK;
}

// Device side of a set of tables; clo may be null (history only).
static int32_t make_attach(lps_ctx *c, const Tables &t, lps_closure *clo,
                           const std::vector<int32_t> &row_slot,
                           const std::vector<int32_t> &row_inst, int32_t n_inst,
                           ClosureAttach **out) {
  ClosureAttach *A = new ClosureAttach();
  A->clo = clo;
  A->n_rows = (int32_t)row_slot.size();
  A->n_hist = (int32_t)t.hist_slot.size();
  A->n_inst = n_inst;
  // LPS_CLOSURE_WIDE_MIN: row count from which the 128-row kernel runs (both give the same bits)
  if (const char *w = std::getenv("LPS_CLOSURE_WIDE_MIN")) A->wide_min = std::atoll(w);
  if (const char *w = std::getenv("LPS_CLOSURE_WIDE_WAVES")) A->wide_waves = std::atoi(w) == 4 ? 4 : 8;
  A->gain_user.assign(std::max(n_inst, 1), 1.0f);
  A->gain_now = A->gain_user;
  A->e_prev.assign(std::max(n_inst, 1), -1.0);
  A->off_step.assign(std::max(n_inst, 1), -1);
  const hipStream_t s = c->stream;
  hipError_t e = A->upload(&A->hist_slot, t.hist_slot, s);
  if (e == hipSuccess) e = A->upload(&A->hidx, t.hidx, s);
  if (e == hipSuccess) e = A->upload(&A->xf, t.xf, s);
  if (e == hipSuccess) e = A->upload(&A->feat, t.feat, s);
  if (e == hipSuccess) e = A->upload(&A->row_slot, row_slot, s);
  if (e == hipSuccess) e = A->upload(&A->row_inst, row_inst, s);
  if (e == hipSuccess) e = A->upload(&A->gain, std::vector<float>(std::max(n_inst, 1), 1.0f), s);
  if (e == hipSuccess) e = A->upload(&A->gate, std::vector<float>(std::max(n_inst, 1), 0.0f), s);
  for (int q = 0; q < 2 && e == hipSuccess; q++) {
    e = A->alloc(&A->hist[q], (size_t)std::max(A->n_hist, 1) * kClRec);
    if (e == hipSuccess) e = A->alloc(&A->mx[q], (size_t)std::max(A->n_hist, 1) * 2);
  }
  if (e == hipSuccess) e = A->alloc(&A->dbg, (size_t)std::max(A->n_rows, 1) * 8);
  if (e == hipSuccess && A->clo) e = A->alloc(&A->defer, (size_t)std::max(A->n_rows, 1) * 8);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete A;
    return c->hip_fail(e, "closure tables");
  }
  *out = A;
  return LPS_OK;
}

// Lattice index of a coordinate on an origin-centred lattice.
static inline int32_t lat(double x, double d) { return (int32_t)std::lround(x / d); }

// ---------------------------------------------------------------------------------------------
// Teacher-forced samples
// ---------------------------------------------------------------------------------------------
struct CollectArgs {
  GatherSrc g;
  int32_t n_rows;
  // trunc's G_u, G_v rows in the solver's layout (StepArgs): row r has members off[r] .. off[r+1],
  // its operator block at gofs[r] under the lattice transform xf[r], member map pmap
  const int32_t *off, *gofs;
  const int8_t *xf;
  const uint8_t *pmap;
  const int32_t *cloud;             // slots of the state below (-1 skip)
  const double *gu, *gv;
  const double *ux, *uy, *uxm, *uym, *vx, *vy, *vxm, *vym;  // U^n, U^{n-1}, V^n, V^{n-1}
  // the target: U^{n+1}, V^{n+1} of the reflection-free run at the row's node (slot tgt_slot[r])
  const int32_t *tgt_slot;
  const double *ux1, *uy1, *vx1, *vy1;
  uint64_t seed, keep_thr;
  int64_t step;  // n + 1
  float amp_min;
  double y_max;
  __bf16 *dsx;
  float *dsaux;
  unsigned long long *count;
  int64_t cap;
};

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__global__ void __launch_bounds__(256) k_collect(CollectArgs a) {
  const int lane = threadIdx.x & 63;
  const int32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= a.n_rows) return;
  const uint64_t h = dev_mix(a.seed ^ dev_mix(((uint64_t)a.step << 32) ^ (uint64_t)r));
  if (h > a.keep_thr) return;
  float su, sv;
  row_scale(a.g, r, su, sv);
  if (!(su >= a.amp_min) || su < 1e-30f) return;
  // least-squares update on the truth state, as the solver's G_u / G_v rows
  const int64_t o = a.off[r], gb = a.gofs[r];
  const int32_t L = a.off[r + 1] - a.off[r];
  const int xf = a.xf[r];
  double Ux = 0.0, Uy = 0.0, Vx = 0.0, Vy = 0.0;
  for (int base = 0; base < L; base += 64) {
    const int jj = base + lane;
    const int32_t k = jj < L ? a.cloud[o + jj] : -1;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
    if (k >= 0) {
      const int m = a.pmap[o + jj];
      const double *bu = a.gu + 8 * gb + 4 * m, *bv = a.gv + 8 * gb + 4 * m;
      double4 g0, g1, h0, h1;
      lattice_map(xf, *(const double4 *)bu, *(const double4 *)(bu + 4 * L), g0, g1);
      lattice_map(xf, *(const double4 *)bv, *(const double4 *)(bv + 4 * L), h0, h1);
      const double ux = a.ux[k], uy = a.uy[k], uxm = a.uxm[k], uym = a.uym[k];
      const double vx = a.vx[k], vy = a.vy[k], vxm = a.vxm[k], vym = a.vym[k];
      t0 = g0.x * ux + g0.y * uy + g0.z * uxm + g0.w * uym;
      t1 = g1.x * ux + g1.y * uy + g1.z * uxm + g1.w * uym;
      t2 = h0.x * vx + h0.y * vy + h0.z * vxm + h0.w * vym;
      t3 = h1.x * vx + h1.y * vy + h1.z * vxm + h1.w * vym;
    }
    Ux += wsum(t0);
    Uy += wsum(t1);
    Vx += wsum(t2);
    Vy += wsum(t3);
  }
  // targets in the canonical frame, units of s_u / s_v (every lane holds the sums)
  const float iu = safe_inv(su), iv = safe_inv(sv);
  const int32_t i = a.tgt_slot[r];
  const double dux = a.ux1[i] - Ux, duy = a.uy1[i] - Uy;
  const double dvx = a.vx1[i] - Vx, dvy = a.vy1[i] - Vy;
  const int f = a.g.xf[r];
  const double pu = (f & 1) ? duy : dux, qu = (f & 1) ? dux : duy;
  const double pv = (f & 1) ? dvy : dvx, qv = (f & 1) ? dvx : dvy;
  const double y0 = ((f & 2) ? -pu : pu) * iu, y1 = ((f & 4) ? -qu : qu) * iu;
  const double y2 = ((f & 2) ? -pv : pv) * iv, y3 = ((f & 4) ? -qv : qv) * iv;
  if (!(fmax(fmax(fabs(y0), fabs(y1)), fmax(fabs(y2), fabs(y3))) <= a.y_max)) return;
  unsigned long long idx = 0;
  if (lane == 0) idx = atomicAdd(a.count, 1ull);
  idx = ((unsigned long long)__shfl((unsigned)(idx >> 32), 0) << 32) | (unsigned)__shfl((unsigned)idx, 0);
  if ((int64_t)idx >= a.cap) return;
  for (int q = lane; q < kClIn / 8; q += 64)
    *(bf16x8 *)(a.dsx + (int64_t)idx * kClIn + 8 * q) = gather8(a.g, r, 8 * q, iu, iv);
  if (lane == 0) {
    float *o = a.dsaux + (int64_t)idx * 8;
    o[0] = (float)y0;
    o[1] = (float)y1;
    o[2] = (float)y2;
    o[3] = (float)y3;
    o[4] = 1.0f;
    o[5] = su;
    o[6] = sv;
    o[7] = (float)r;
  }
}

}  // namespace lps

using namespace lps;

extern "C" {

int32_t lps_closure_attach(lps_ctx *c, lps_closure *clo) {
  if (!clo) return c->fail(LPS_ERR_ARG, "closure is NULL");
  if (!c->fast) return c->fail(LPS_ERR_ARG, "the learned closure needs the fast (tile) path");
  if (clo->device != c->opt.device) return c->fail(LPS_ERR_ARG, "closure and context on different devices");
  CTX_TRY(c, hipSetDevice(c->opt.device));
  const int32_t B = c->n_inst;
  std::vector<double> xs, ys;
  std::vector<int8_t> kind;
  std::vector<int32_t> aslot;
  int32_t rc;
  if ((rc = download(c, xs, c->x, c->n_slots)) || (rc = download(c, ys, c->y, c->n_slots)) ||
      (rc = download(c, kind, c->kind, c->n_slots)) || (rc = download(c, aslot, c->abc_slot, c->n_abc)))
    return rc;
  std::vector<RowIn> rows;
  std::vector<int32_t> row_slot, row_inst;
  std::vector<Grid> grids(B);
  int32_t a0 = 0;
  for (int32_t i = 0; i < B; i++) {
    const int64_t s0 = c->batch ? c->slot_off[i] : 0, s1 = c->batch ? c->slot_off[i + 1] : c->n_slots;
    const int32_t na = c->batch ? c->abc_n[i] : c->n_abc;
    const InstParams &P = c->inst_h[i];
    const double nu = P.lam / (2.0 * (P.lam + P.mu));
    std::vector<int32_t> ix, iy;
    double R = 0.0;
    for (int64_t k = s0; k < s1; k++)
      if (kind[k]) {
        ix.push_back(lat(xs[k], c->dx));
        iy.push_back(lat(ys[k], c->dy));
        R = std::max(R, std::sqrt(xs[k] * xs[k] + ys[k] * ys[k]));
      }
    Grid &G = grids[i];
    G.init(ix, iy);
    for (int64_t k = s0; k < s1; k++)
      if (kind[k]) G.set(lat(xs[k], c->dx), lat(ys[k], c->dy), (int32_t)k);
    for (int32_t r = a0; r < a0 + na; r++) {
      const int32_t k = aslot[r];
      rows.push_back(RowIn{lat(xs[k], c->dx), lat(ys[k], c->dy), xs[k], ys[k], nu, R});
      row_slot.push_back(k);
      row_inst.push_back(i);
    }
    a0 += na;
  }
  // rows of instance i only see instance i's lattice
  Tables t;
  size_t cur = 0;
  int32_t cur_inst = 0;
  auto present = [&](int32_t x, int32_t y) { return grids[cur_inst].get(x, y) >= 0; };
  auto slot_of = [&](int32_t x, int32_t y) { return grids[cur_inst].get(x, y); };
  // build instance by instance, keeping one history index space
  {
    Tables all;
    std::unordered_map<int32_t, int32_t> hglob;
    for (int32_t i = 0; i < B; i++) {
      std::vector<RowIn> sub;
      for (size_t r = cur; r < rows.size() && row_inst[r] == i; r++) sub.push_back(rows[r]);
      cur_inst = i;
      Tables ti;
      if (!build_rows(sub, c->delta, present, slot_of, ti))
        return c->fail(LPS_ERR_ARG, "closure tables: a lattice position without a slot");
      for (size_t r = 0; r < sub.size(); r++) {
        for (int j = 0; j < kClOffsets; j++) {
          const int32_t h = ti.hidx[r * kClOffsets + j];
          int32_t g = -1;
          if (h >= 0) {
            const int32_t s = ti.hist_slot[h];
            auto it = hglob.find(s);
            if (it == hglob.end()) {
              g = (int32_t)all.hist_slot.size();
              hglob.emplace(s, g);
              all.hist_slot.push_back(s);
            } else {
              g = it->second;
            }
          }
          all.hidx.push_back(g);
        }
        all.xf.push_back(ti.xf[r]);
        for (int e = 0; e < kClFeat; e++) all.feat.push_back(ti.feat[r * kClFeat + e]);
      }
      cur += sub.size();
    }
    t = std::move(all);
  }
  ClosureAttach *A = nullptr;
  if ((rc = make_attach(c, t, clo, row_slot, row_inst, B, &A))) return rc;
  closure_free(c);
  c->clo = A;
  ctx_drop_graphs(c);
  if (closure_reset_hook(c) != hipSuccess) return c->fail(LPS_ERR_HIP, "clearing the closure history");
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  return LPS_OK;
}

int32_t lps_closure_detach(lps_ctx *c) {
  if (c->timing_closure) {  // its events would no longer be recorded
    int32_t rc = lps_set_timing(c, 0);
    if (rc != LPS_OK) return rc;
  }
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  CTX_TRY(c, hipStreamSynchronize(c->stream2));
  closure_free(c);
  ctx_drop_graphs(c);
  return LPS_OK;
}

int32_t lps_closure_set_gain(lps_ctx *c, const float *gain) {
  ClosureAttach *A = c->clo;
  if (!A) return c->fail(LPS_ERR_ARG, "no closure attached");
  A->gain_user.assign(gain, gain + A->n_inst);
  A->gain_now = A->gain_user;
  A->off_step.assign(A->n_inst, -1);
  CTX_TRY(c, hipMemcpyAsync(A->gain, A->gain_now.data(), sizeof(float) * A->n_inst, hipMemcpyHostToDevice, c->stream));
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  return LPS_OK;
}

int32_t lps_closure_set_gate(lps_ctx *c, const float *min_su) {
  ClosureAttach *A = c->clo;
  if (!A) return c->fail(LPS_ERR_ARG, "no closure attached");
  CTX_TRY(c, hipMemcpyAsync(A->gate, min_su, sizeof(float) * A->n_inst, hipMemcpyHostToDevice, c->stream));
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  return LPS_OK;
}

int32_t lps_closure_guard(lps_ctx *c, double tol, int64_t *fallback_step) {
  ClosureAttach *A = c->clo;
  if (!A) return c->fail(LPS_ERR_ARG, "no closure attached");
  if (tol >= 0.0) A->guard_tol = tol;
  if (fallback_step)
    for (int32_t i = 0; i < A->n_inst; i++) fallback_step[i] = A->off_step[i];
  return LPS_OK;
}

int32_t lps_closure_inspect(lps_ctx *c, uint16_t *x, float *scale2, float *applied8) {
  ClosureAttach *A = c->clo;
  if (!A) return c->fail(LPS_ERR_ARG, "no closure attached");
  if (c->step < 1) return c->fail(LPS_ERR_ARG, "no step completed since the attachment");
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  CTX_TRY(c, hipStreamSynchronize(c->stream2));
  const int par = (int)((c->step - 1) & 1);
  const int32_t n = A->n_rows;
  if (n == 0) return LPS_OK;
  __bf16 *dx = nullptr;
  float *ds = nullptr;
  hipError_t e = hipMalloc((void **)&dx, sizeof(__bf16) * n * kClIn);
  if (e == hipSuccess) e = hipMalloc((void **)&ds, sizeof(float) * n * 2);
  if (e == hipSuccess) e = gather_rows(A->src(par), n, dx, ds, c->stream);
  if (e == hipSuccess && x)
    e = hipMemcpyAsync(x, dx, sizeof(__bf16) * n * kClIn, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && scale2)
    e = hipMemcpyAsync(scale2, ds, sizeof(float) * n * 2, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && applied8)
    e = hipMemcpyAsync(applied8, A->dbg, sizeof(float) * n * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dx);
  (void)hipFree(ds);
  if (e != hipSuccess) return c->hip_fail(e, "lps_closure_inspect");
  return LPS_OK;
}

int32_t lps_closure_collect(lps_closure *clo, lps_ctx *c, const lps_problem *p, int32_t nsteps,
                            int32_t first_step, int32_t sample_every, double keep, double amp_min,
                            double y_max, uint64_t seed) {
  if (!clo || !p) return c->fail(LPS_ERR_ARG, "closure and truncated problem are required");
  if (!c->fast || c->batch) return c->fail(LPS_ERR_ARG, "the truth run needs a single fast-path context");
  if (clo->device != c->opt.device) return c->fail(LPS_ERR_ARG, "closure and context on different devices");
  if (clo->ds_cap <= 0) return c->fail(LPS_ERR_ARG, "reserve the sample store first");
  if (nsteps < 0 || sample_every <= 0 || !(y_max > 0.0)) return c->fail(LPS_ERR_ARG, "bad sampling arguments");
  const int32_t N = p->node_no, C = p->constrains_no, M = p->max_cloud_size;
  if (N <= 0 || C <= 0 || M <= 0 || M > 64 || !p->nodes || !p->inf_nodes || !p->iclouds ||
      !p->ci_nodes_u || !p->ci_nodes_v)
    return c->fail(LPS_ERR_ARG, "the truncated problem needs nodes and ABC rows (cloud <= 64)");
  if (p->dx != c->dx || p->dy != c->dy) return c->fail(LPS_ERR_ARG, "truncated and truth lattices differ");
  CTX_TRY(c, hipSetDevice(c->opt.device));
  // truth lattice
  std::vector<double> xs, ys;
  std::vector<int8_t> kind;
  int32_t rc;
  if ((rc = download(c, xs, c->x, c->n_slots)) || (rc = download(c, ys, c->y, c->n_slots)) ||
      (rc = download(c, kind, c->kind, c->n_slots)))
    return rc;
  std::vector<int32_t> ix, iy;
  for (int32_t k = 0; k < c->n_slots; k++)
    if (kind[k]) { ix.push_back(lat(xs[k], c->dx)); iy.push_back(lat(ys[k], c->dy)); }
  Grid truth;
  truth.init(ix, iy);
  for (int32_t k = 0; k < c->n_slots; k++)
    if (kind[k]) truth.set(lat(xs[k], c->dx), lat(ys[k], c->dy), k);
  // truncated lattice
  std::vector<int32_t> tx(N), ty(N);
  double R = 0.0;
  for (int32_t i = 0; i < N; i++) {
    const double x = p->nodes[2 * (int64_t)i], y = p->nodes[2 * (int64_t)i + 1];
    tx[i] = lat(x, p->dx);
    ty[i] = lat(y, p->dy);
    R = std::max(R, std::sqrt(x * x + y * y));
  }
  Grid trunc;
  trunc.init(tx, ty);
  for (int32_t i = 0; i < N; i++) trunc.set(tx[i], ty[i], i);
  std::vector<RowIn> rows;
  std::vector<int32_t> row_slot, cloud((size_t)C * M, -1);
  for (int32_t a = 0; a < C; a++) {
    const int32_t node = p->inf_nodes[a];
    if (node < 0 || node >= N) return c->fail(LPS_ERR_ARG, "infNodesIndex entry out of range");
    const double x = p->nodes[2 * (int64_t)node], y = p->nodes[2 * (int64_t)node + 1];
    rows.push_back(RowIn{tx[node], ty[node], x, y, p->nu, R});
    const int32_t s = truth.get(tx[node], ty[node]);
    if (s < 0) return c->fail(LPS_ERR_ARG, "an ABC node of the truncated problem is not in the truth domain");
    row_slot.push_back(s);
    for (int32_t j = 0; j < M; j++) {
      const int32_t k = p->iclouds[(int64_t)a * M + j];
      if (k < 0 || k >= N) continue;
      const int32_t sk = truth.get(tx[k], ty[k]);
      if (sk < 0) return c->fail(LPS_ERR_ARG, "a cloud node of the truncated problem is not in the truth domain");
      cloud[(size_t)a * M + j] = sk;
    }
  }
  Tables t;
  auto present = [&](int32_t x, int32_t y) { return trunc.get(x, y) >= 0; };
  auto slot_of = [&](int32_t x, int32_t y) { return truth.get(x, y); };
  if (!build_rows(rows, c->delta, present, slot_of, t))
    return c->fail(LPS_ERR_ARG, "a node of the truncated problem is not in the truth domain");
  ClosureAttach *A = nullptr;
  if ((rc = make_attach(c, t, nullptr, row_slot, std::vector<int32_t>(C, 0), 1, &A))) return rc;
  int32_t *dcloud = nullptr, *doff = nullptr;
  int8_t *dxf = nullptr;
  uint8_t *dpmap = nullptr;
  double *dgu = nullptr, *dgv = nullptr;
  const size_t gl = (size_t)C * 8 * M;
  // the problem's rows as they are: M members each, operators of their own (identity map)
  std::vector<int32_t> off(C + 1);
  for (int32_t a = 0; a <= C; a++) off[a] = a * M;
  std::vector<uint8_t> pmap((size_t)C * M);
  for (size_t q = 0; q < pmap.size(); q++) pmap[q] = (uint8_t)(q % M);
  hipError_t e = A->upload(&dcloud, cloud, c->stream);
  if (e == hipSuccess) e = A->upload(&doff, off, c->stream);
  if (e == hipSuccess) e = A->upload(&dxf, std::vector<int8_t>(std::max(C, 1), 0), c->stream);
  if (e == hipSuccess) e = A->upload(&dpmap, pmap, c->stream);
  if (e == hipSuccess) e = A->alloc(&dgu, gl);
  if (e == hipSuccess) e = A->alloc(&dgv, gl);
  if (e == hipSuccess) e = hipMemcpyAsync(dgu, p->ci_nodes_u, sizeof(double) * gl, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dgv, p->ci_nodes_v, sizeof(double) * gl, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    delete A;
    return c->hip_fail(e, "closure collection tables");
  }
  // the history rides on the truth run's steps
  ClosureAttach *saved = c->clo;
  const int32_t saved_graph = c->opt.use_graph;
  c->clo = A;
  c->opt.use_graph = 0;
  ctx_drop_graphs(c);
  rc = closure_reset_hook(c) == hipSuccess ? LPS_OK : c->fail(LPS_ERR_HIP, "clearing the history");
  CollectArgs ca{};
  ca.n_rows = C;
  ca.off = doff;
  ca.gofs = doff;
  ca.xf = dxf;
  ca.pmap = dpmap;
  ca.tgt_slot = A->row_slot;
  ca.cloud = dcloud;
  ca.gu = dgu;
  ca.gv = dgv;
  ca.seed = seed;
  ca.keep_thr = keep >= 1.0 ? ~0ull : (uint64_t)(std::max(keep, 0.0) * 18446744073709551615.0);
  ca.amp_min = (float)amp_min;
  ca.y_max = y_max;
  ca.dsx = clo->ds_x;
  ca.dsaux = clo->ds_aux;
  ca.count = clo->ds_count;
  ca.cap = clo->ds_cap;
  for (int32_t it = 0; it < nsteps && rc == LPS_OK; it++) {
    const int64_t n = c->step;  // the step about to run
    rc = lps_step(c, 1);
    if (rc != LPS_OK) break;
    const int64_t s1 = n + 1;
    if (s1 < first_step || s1 % sample_every != 0) continue;
    const int par = (int)(n & 1);
    ca.g = A->src(par);
    ca.step = s1;
    // U^{s1}, U^{s1-1}, U^{s1-2} of every node
    if ((rc = ctx_sync_u(c, 3)) != LPS_OK) break;
    const lps_ctx::Dev2 &U1 = c->U[s1 % 3], &U0 = c->U[n % 3], &Um = c->U[(n + 2) % 3];
    const lps_ctx::Dev2 &V1 = c->V[s1 % 3], &V0 = c->V[n % 3], &Vm = c->V[(n + 2) % 3];
    ca.ux1 = U1.x; ca.uy1 = U1.y; ca.vx1 = V1.x; ca.vy1 = V1.y;
    ca.ux = U0.x; ca.uy = U0.y; ca.uxm = Um.x; ca.uym = Um.y;
    ca.vx = V0.x; ca.vy = V0.y; ca.vxm = Vm.x; ca.vym = Vm.y;
    hipLaunchKernelGGL(k_collect, dim3((C + 3) / 4), dim3(256), 0, c->stream, ca);
    e = hipGetLastError();
    if (e != hipSuccess) rc = c->hip_fail(e, "k_collect");
  }
  e = hipStreamSynchronize(c->stream);
  if (rc == LPS_OK && e != hipSuccess) rc = c->hip_fail(e, "lps_closure_collect");
  c->clo = saved;
  c->opt.use_graph = saved_graph;
  ctx_drop_graphs(c);
  delete A;
  return rc;
}

int32_t lps_closure_collect_closed(lps_closure *clo, lps_ctx *truth, lps_ctx *c, int32_t nsteps,
                                   int32_t first_step, int32_t sample_every, double keep,
                                   double amp_min, double y_max, uint64_t seed) {
  if (!clo || !truth) return c->fail(LPS_ERR_ARG, "closure and truth context are required");
  ClosureAttach *A = c->clo;
  if (!A || !A->clo || c->batch || !c->fast)
    return c->fail(LPS_ERR_ARG, "the truncated run needs a single fast-path context with a closure attached");
  if (!truth->fast || truth->batch || truth->n_abc != 0)
    return c->fail(LPS_ERR_ARG, "the truth run needs a single fast-path context without ABC rows");
  if (clo->device != c->opt.device || truth->opt.device != c->opt.device)
    return c->fail(LPS_ERR_ARG, "closure and contexts on different devices");
  if (clo->ds_cap <= 0) return c->fail(LPS_ERR_ARG, "reserve the sample store first");
  if (nsteps < 0 || sample_every <= 0 || !(y_max > 0.0)) return c->fail(LPS_ERR_ARG, "bad sampling arguments");
  if (truth->step != c->step) return c->fail(LPS_ERR_ARG, "truth and truncated runs are at different steps");
  if (truth->dx != c->dx || truth->dy != c->dy || truth->dt != c->dt)
    return c->fail(LPS_ERR_ARG, "truth and truncated runs differ in lattice or time step");
  if (c->M > 64) return c->fail(LPS_ERR_ARG, "clouds of more than 64 nodes");
  CTX_TRY(c, hipSetDevice(c->opt.device));
  // the truth slot of every ABC row's node
  std::vector<double> xs, ys, txs, tys;
  std::vector<int8_t> tkind;
  std::vector<int32_t> aslot;
  int32_t rc;
  if ((rc = download(c, xs, c->x, c->n_slots)) || (rc = download(c, ys, c->y, c->n_slots)) ||
      (rc = download(c, aslot, c->abc_slot, c->n_abc)))
    return rc;
  if ((rc = download(truth, txs, truth->x, truth->n_slots)) ||
      (rc = download(truth, tys, truth->y, truth->n_slots)) ||
      (rc = download(truth, tkind, truth->kind, truth->n_slots)))
    return c->fail(rc, "truth context: " + truth->err);
  std::vector<int32_t> ix, iy;
  for (int32_t k = 0; k < truth->n_slots; k++)
    if (tkind[k]) { ix.push_back(lat(txs[k], c->dx)); iy.push_back(lat(tys[k], c->dy)); }
  Grid tg;
  tg.init(ix, iy);
  for (int32_t k = 0; k < truth->n_slots; k++)
    if (tkind[k]) tg.set(lat(txs[k], c->dx), lat(tys[k], c->dy), k);
  std::vector<int32_t> tslot(c->n_abc);
  for (int32_t r = 0; r < c->n_abc; r++) {
    const int32_t k = aslot[r];
    tslot[r] = tg.get(lat(xs[k], c->dx), lat(ys[k], c->dy));
    if (tslot[r] < 0) return c->fail(LPS_ERR_ARG, "an ABC node of the truncated run is not in the truth domain");
  }
  int32_t *dtslot = nullptr;
  CTX_TRY(c, hipMalloc((void **)&dtslot, sizeof(int32_t) * std::max(c->n_abc, 1)));
  hipError_t e = hipMemcpyAsync(dtslot, tslot.data(), sizeof(int32_t) * c->n_abc, hipMemcpyHostToDevice, c->stream);
  CollectArgs ca{};
  ca.n_rows = c->n_abc;
  ca.off = c->abc_off;
  ca.gofs = c->abc_gofs;
  ca.xf = c->abc_xf;
  ca.pmap = c->abc_pmap;
  ca.cloud = c->cloud;
  ca.gu = c->gu;
  ca.gv = c->gv;
  ca.tgt_slot = dtslot;
  ca.seed = seed;
  ca.keep_thr = keep >= 1.0 ? ~0ull : (uint64_t)(std::max(keep, 0.0) * 18446744073709551615.0);
  ca.amp_min = (float)amp_min;
  ca.y_max = y_max;
  ca.dsx = clo->ds_x;
  ca.dsaux = clo->ds_aux;
  ca.count = clo->ds_count;
  ca.cap = clo->ds_cap;
  rc = e == hipSuccess ? LPS_OK : c->hip_fail(e, "lps_closure_collect_closed");
  for (int32_t it = 0; it < nsteps && rc == LPS_OK; it++) {
    const int64_t n = c->step;  // the step about to run in both
    if ((rc = lps_step(truth, 1)) != LPS_OK) {
      rc = c->fail(rc, "truth step: " + truth->err);
      break;
    }
    if ((rc = lps_step(c, 1)) != LPS_OK) break;
    const int64_t s1 = n + 1;
    if (s1 < first_step || s1 % sample_every != 0) continue;
    // truth's U^{s1} (interior nodes from its prediction ring) and V^{s1}, complete before this
    // context's stream reads them
    if ((rc = ctx_sync_u(truth, 1)) != LPS_OK) {
      rc = c->fail(rc, "truth: " + truth->err);
      break;
    }
    if ((e = hipStreamSynchronize(truth->stream)) != hipSuccess) { rc = c->hip_fail(e, "truth stream"); break; }
    if ((rc = ctx_sync_u(c, 3)) != LPS_OK) break;
    // the closure's inputs of step n: the history it gathered from
    ca.g = A->src((int)(n & 1));
    ca.step = s1;
    const lps_ctx::Dev2 &U0 = c->U[n % 3], &Um = c->U[(n + 2) % 3];
    const lps_ctx::Dev2 &V0 = c->V[n % 3], &Vm = c->V[(n + 2) % 3];
    ca.ux = U0.x; ca.uy = U0.y; ca.uxm = Um.x; ca.uym = Um.y;
    ca.vx = V0.x; ca.vy = V0.y; ca.vxm = Vm.x; ca.vym = Vm.y;
    const lps_ctx::Dev2 &T1 = truth->U[s1 % 3], &TV1 = truth->V[s1 % 3];
    ca.ux1 = T1.x; ca.uy1 = T1.y; ca.vx1 = TV1.x; ca.vy1 = TV1.y;
    if ((e = hipStreamSynchronize(c->stream2)) != hipSuccess) { rc = c->hip_fail(e, "closure stream"); break; }
    hipLaunchKernelGGL(k_collect, dim3((c->n_abc + 3) / 4), dim3(256), 0, c->stream, ca);
    e = hipGetLastError();
    // done before truth's next steps overwrite what it read
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = c->hip_fail(e, "k_collect");
  }
  e = hipStreamSynchronize(c->stream);
  if (rc == LPS_OK && e != hipSuccess) rc = c->hip_fail(e, "lps_closure_collect_closed");
  (void)hipFree(dtslot);
  return rc;
}

}  // extern "C"
