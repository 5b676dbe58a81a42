// Least-squares ABC operators on the GPU: the per-row solves of the reference's preprocessor
// (preprocessor/main.cpp:415-637 for G_u / G_v, :657-785 for G_theta, the LS solve :368-402) for
// every ABC row at once.
//
// Per row and per operator, the system A x = b is built on the device (one workgroup per row):
//   rows    4 per wave mode (cos/sin of the P and S parts), in the mode order given;
//   columns (time level tl, cloud slot k, component c) at tl*2*Mc + 2*k + c for G_u / G_v and
//           tl*Mc + k for G_theta, Mc = max_cloud. Slots k >= the row's cloud size are zero
//           columns; the reference's column order (tl*2*nC + 2*k + c) is a permutation of these,
//           and neither a permutation nor zero columns change the minimum-norm solution.
// It is solved like LAPACK dgelsd with rcond: the SVD A = U S V' (rocSOLVER dgesvd, batched),
// singular values s_i <= rcond * s_1 dropped, x = V S^+ U' b. The solution is scattered into the
// .ciNodesU / .ciNodesV / .ciNodesTheta layouts (preprocessor/main.cpp:621-636, 778-784).
//
// These systems are numerically rank-deficient (about 20 of up to 196 singular values lie above
// 1e-14 s_1, the smallest kept ones near 1e-13 s_1), so any backward-stable solver determines the
// solution along the smallest kept directions only to about eps s_1 / s_r: LAPACK's dgelsd and
// dgelss differ on them by 1.7e-3 (median over rows), up to 0.19 on single rows. Results are
// compared with the CPU builder through the least-squares residual, the kept rank, the operators'
// norms and the runs they give.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/lps_amd.h"
#include "lps_internal.h"

namespace {

enum Kind { kU = 0, kV = 1, kTheta = 2 };

struct RowsDev {
  int32_t count, max_cloud;
  const double *xy, *tau, *cloud_xy;
  const int32_t *cloud_len;
};

struct ModesDev {
  int32_t count;
  const double *a1, *a2, *om_p, *om_s, *g_p, *g_s, *s_r, *s_i;
};

// A (column-major, lda = m) and b (m x nrhs, column-major) of rows first .. first + nb - 1, one
// workgroup per row. m = 4 modes; n = 4 Mc (G_u, G_v) or 2 Mc (G_theta).
__global__ void __launch_bounds__(256) k_abc_system(RowsDev R, ModesDev Md, double dt,
                                                    double th_coeff, int32_t kind, int32_t first,
                                                    int32_t nb, double *A, double *B) {
#pragma clang fp contract(off)
  const int32_t p = blockIdx.x;
  if (p >= nb) return;
  const int32_t r = first + p;
  const int32_t Mc = R.max_cloud, nm = Md.count, m = 4 * nm;
  const int32_t n = kind == kTheta ? 2 * Mc : 4 * Mc;
  const int32_t nrhs = kind == kTheta ? 1 : 2;
  double *Ap = A + (int64_t)p * m * n;
  double *Bp = B + (int64_t)p * m * nrhs;
  const double ax = R.xy[2 * r], ay = R.xy[2 * r + 1];
  const double cosT = cos(R.tau[r]), sinT = sin(R.tau[r]);
  const int32_t nC = R.cloud_len[r];
  // columns: one (mode, slot, time level) per iteration
  for (int32_t q = threadIdx.x; q < nm * Mc * 2; q += blockDim.x) {
    const int32_t md = q / (Mc * 2), k = (q / 2) % Mc, tl = q % 2;
    double *a0 = Ap + 4 * md;  // rows 4 md .. 4 md + 3
    if (k >= nC) {
      if (kind == kTheta) {
        const int64_t col = (int64_t)(tl * Mc + k) * m;
        for (int b = 0; b < 4; b++) a0[col + b] = 0.0;
      } else {
        const int64_t col = (int64_t)(tl * 2 * Mc + 2 * k) * m;
        for (int b = 0; b < 4; b++) {
          a0[col + b] = 0.0;
          a0[col + m + b] = 0.0;
        }
      }
      continue;
    }
    const double *cxy = R.cloud_xy + ((int64_t)r * Mc + k) * 2;
    const double dxg = cxy[0] - ax, dyg = cxy[1] - ay;
    const double xl = cosT * dxg + sinT * dyg;
    const double yl = -sinT * dxg + cosT * dyg;
    const double phase = Md.a1[md] * xl + Md.a2[md] * yl;
    const double tv = -(double)tl * dt;
    const double omP = Md.om_p[md], omS = Md.om_s[md];
    const double pP = phase - omP * tv, pS = phase - omS * tv;
    const double cP = cos(pP), sP = sin(pP), cS = cos(pS), sS = sin(pS);
    if (kind == kTheta) {
      const double *gp = Md.g_p + 2 * md, *gs = Md.g_s + 2 * md;
      const double *sr = Md.s_r + 2 * md, *si = Md.s_i + 2 * md;
      const double thPr = th_coeff * (gp[0] * sr[0] + gp[1] * sr[1]);
      const double thPi = th_coeff * (gp[0] * si[0] + gp[1] * si[1]);
      const double thSr = th_coeff * (gs[0] * sr[0] + gs[1] * sr[1]);
      const double thSi = th_coeff * (gs[0] * si[0] + gs[1] * si[1]);
      const int64_t col = (int64_t)(tl * Mc + k) * m;
      a0[col + 0] = thPr * cP - thPi * sP;
      a0[col + 1] = thSr * cS - thSi * sS;
      a0[col + 2] = thPr * sP + thPi * cP;
      a0[col + 3] = thSr * sS + thSi * cS;
    } else {
      const double gPx = Md.g_p[2 * md], gPy = Md.g_p[2 * md + 1];
      const double gSx = Md.g_s[2 * md], gSy = Md.g_s[2 * md + 1];
      double lx[4], ly[4];
      if (kind == kU) {
        lx[0] = gPx * cP; ly[0] = gPy * cP;
        lx[1] = gSx * cS; ly[1] = gSy * cS;
        lx[2] = gPx * sP; ly[2] = gPy * sP;
        lx[3] = gSx * sS; ly[3] = gSy * sS;
      } else {
        lx[0] = gPx * omP * sP; ly[0] = gPy * omP * sP;
        lx[1] = gSx * omS * sS; ly[1] = gSy * omS * sS;
        lx[2] = -gPx * omP * cP; ly[2] = -gPy * omP * cP;
        lx[3] = -gSx * omS * cS; ly[3] = -gSy * omS * cS;
      }
      const int64_t col = (int64_t)(tl * 2 * Mc + 2 * k) * m;
      for (int b = 0; b < 4; b++) {
        a0[col + b] = cosT * lx[b] - sinT * ly[b];
        a0[col + m + b] = sinT * lx[b] + cosT * ly[b];
      }
    }
  }
  // right-hand sides: one mode per iteration
  for (int32_t md = threadIdx.x; md < nm; md += blockDim.x) {
    const double omP = Md.om_p[md], omS = Md.om_s[md];
    const double cPd = cos(-omP * dt), sPd = sin(-omP * dt);
    const double cSd = cos(-omS * dt), sSd = sin(-omS * dt);
    double *b0 = Bp + 4 * md;
    if (kind == kTheta) {
      const double *gp = Md.g_p + 2 * md, *gs = Md.g_s + 2 * md;
      const double *sr = Md.s_r + 2 * md, *si = Md.s_i + 2 * md;
      const double thPr = th_coeff * (gp[0] * sr[0] + gp[1] * sr[1]);
      const double thPi = th_coeff * (gp[0] * si[0] + gp[1] * si[1]);
      const double thSr = th_coeff * (gs[0] * sr[0] + gs[1] * sr[1]);
      const double thSi = th_coeff * (gs[0] * si[0] + gs[1] * si[1]);
      b0[0] = thPr * cPd - thPi * sPd;
      b0[1] = thSr * cSd - thSi * sSd;
      b0[2] = thPr * sPd + thPi * cPd;
      b0[3] = thSr * sSd + thSi * cSd;
    } else {
      for (int b = 0; b < 4; b++) {
        const bool S = b % 2 == 1;
        const double gx = S ? Md.g_s[2 * md] : Md.g_p[2 * md];
        const double gy = S ? Md.g_s[2 * md + 1] : Md.g_p[2 * md + 1];
        const double om = S ? omS : omP, cd = S ? cSd : cPd, sd = S ? sSd : sPd;
        double lx, ly;
        if (kind == kU) {
          lx = b < 2 ? gx * cd : gx * sd;
          ly = b < 2 ? gy * cd : gy * sd;
        } else {
          lx = b < 2 ? gx * om * sd : -gx * om * cd;
          ly = b < 2 ? gy * om * sd : -gy * om * cd;
        }
        b0[b] = cosT * lx - sinT * ly;
        b0[m + b] = sinT * lx + cosT * ly;
      }
    }
  }
}

// x = V S^+ U' b for each problem of a chunk: s_i <= rcond s_1 dropped (dgelsd's rule). U: m x k
// (ldu = m), VT: k x n (ldv = k), B: m x nrhs, X: n x nrhs. One workgroup per problem; rank[p]
// receives the number of singular values kept, or -1 when the SVD did not converge.
__global__ void __launch_bounds__(256) k_pinv_solve(int32_t m, int32_t n, int32_t nrhs, int32_t nb,
                                                    const double *S, const double *U,
                                                    const double *VT, const double *B,
                                                    const int32_t *info, double rcond, double *X,
                                                    int32_t *rank) {
  extern __shared__ double c[];  // k x nrhs
  const int32_t p = blockIdx.x;
  if (p >= nb) return;
  const int32_t k = min(m, n);
  const double *Sp = S + (int64_t)p * k, *Up = U + (int64_t)p * m * k;
  const double *Vp = VT + (int64_t)p * k * n, *Bp = B + (int64_t)p * m * nrhs;
  double *Xp = X + (int64_t)p * n * nrhs;
  const double thr = rcond * Sp[0];
  for (int32_t t = threadIdx.x; t < k * nrhs; t += blockDim.x) {
    const int32_t i = t % k, j = t / k;
    double v = 0.0;
    if (Sp[i] > thr) {
      for (int32_t row = 0; row < m; row++) v = fma(Up[(int64_t)i * m + row], Bp[(int64_t)j * m + row], v);
      v /= Sp[i];
    }
    c[t] = v;
  }
  __syncthreads();
  for (int32_t t = threadIdx.x; t < n * nrhs; t += blockDim.x) {
    const int32_t col = t % n, j = t / n;
    double v = 0.0;
    for (int32_t i = 0; i < k; i++) v = fma(Vp[(int64_t)col * k + i], c[j * k + i], v);
    Xp[(int64_t)j * n + col] = v;
  }
  if (threadIdx.x == 0 && rank) {
    int32_t q = 0;
    for (int32_t i = 0; i < k; i++) q += Sp[i] > thr;
    rank[p] = info[p] == 0 ? q : -1;
  }
}

// Solutions of a chunk of rows into the operator layouts.
__global__ void k_abc_scatter(int32_t kind, int32_t first, int32_t nb, int32_t Mc,
                              const int32_t *cloud_len, const double *X, double *G) {
  const int32_t p = blockIdx.x;
  if (p >= nb) return;
  const int32_t r = first + p, nC = cloud_len[r];
  if (kind == kTheta) {
    const double *x = X + (int64_t)p * 2 * Mc;
    for (int32_t t = threadIdx.x; t < 2 * Mc; t += blockDim.x) {
      const int32_t k = t / 2, tl = t % 2;
      G[(int64_t)r * 2 * Mc + t] = k < nC ? x[tl * Mc + k] : 0.0;
    }
  } else {
    const int32_t n = 4 * Mc;
    const double *x = X + (int64_t)p * n * 2;
    for (int32_t t = threadIdx.x; t < 2 * n; t += blockDim.x) {
      const int32_t rr = t / n, d = t % n;  // G row 2 r + rr, column d = 4 k + 2 tl + c
      const int32_t k = d / 4, tl = (d / 2) % 2, cc = d % 2;
      G[(int64_t)(2 * r + rr) * n + d] = k < nC ? x[(int64_t)rr * n + tl * 2 * Mc + 2 * k + cc] : 0.0;
    }
  }
}

struct Dev {
  std::vector<void *> ptrs;
  ~Dev() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t get(T **p, size_t count) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
  template <class T>
  hipError_t put(T **p, const T *h, size_t count) {
    hipError_t e = get(p, count);
    if (e == hipSuccess && count) e = hipMemcpy(*p, h, count * sizeof(T), hipMemcpyHostToDevice);
    return e;
  }
};

struct Handle {
  rocblas_handle h = nullptr;
  hipStream_t s = nullptr;
  ~Handle() {
    if (h) rocblas_destroy_handle(h);
    if (s) (void)hipStreamDestroy(s);
  }
};

int32_t fail(const std::string &msg) {
  lps::g_create_error = msg;
  return LPS_ERR_ARG;
}
int32_t hip_fail(hipError_t e, const char *what) {
  lps::g_create_error = std::string(what) + ": " + hipGetErrorString(e);
  return LPS_ERR_HIP;
}
#define ABC_TRY(call)                                \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

// Workspace of the batched SVD + solve for chunks of up to cap problems of size m x n.
struct Solver {
  int32_t m = 0, n = 0, k = 0, nrhs = 0, cap = 0;
  double *A = nullptr, *B = nullptr, *S = nullptr, *U = nullptr, *VT = nullptr, *E = nullptr,
         *X = nullptr;
  int32_t *info = nullptr, *rank = nullptr;
  hipError_t init(Dev &d, int32_t m_, int32_t n_, int32_t nrhs_, int32_t cap_) {
    m = m_; n = n_; nrhs = nrhs_; cap = cap_; k = std::min(m, n);
    hipError_t e;
    if ((e = d.get(&A, (size_t)cap * m * n)) != hipSuccess) return e;
    if ((e = d.get(&B, (size_t)cap * m * nrhs)) != hipSuccess) return e;
    if ((e = d.get(&S, (size_t)cap * k)) != hipSuccess) return e;
    if ((e = d.get(&E, (size_t)cap * k)) != hipSuccess) return e;
    if ((e = d.get(&U, (size_t)cap * m * k)) != hipSuccess) return e;
    if ((e = d.get(&VT, (size_t)cap * k * n)) != hipSuccess) return e;
    if ((e = d.get(&X, (size_t)cap * n * nrhs)) != hipSuccess) return e;
    if ((e = d.get(&info, cap)) != hipSuccess) return e;
    return d.get(&rank, cap);
  }
  // SVD and minimum-norm solution of the nb problems in A, B (A is destroyed).
  int32_t solve(const Handle &H, int32_t nb, double rcond) {
    const rocblas_status st = rocsolver_dgesvd_strided_batched(
        H.h, rocblas_svect_singular, rocblas_svect_singular, m, n, A, m, (rocblas_stride)m * n, S,
        k, U, m, (rocblas_stride)m * k, VT, k, (rocblas_stride)k * n, E, k, rocblas_outofplace,
        info, nb);
    if (st != rocblas_status_success)
      return fail("rocsolver_dgesvd_strided_batched failed: " +
                  std::string(rocblas_status_to_string(st)));
    hipLaunchKernelGGL(k_pinv_solve, dim3(nb), dim3(256), sizeof(double) * k * nrhs, H.s, m, n,
                       nrhs, nb, S, U, VT, B, info, rcond, X, rank);
    ABC_TRY(hipGetLastError());
    return LPS_OK;
  }
};

// Problems per chunk: about 1.5 GB of workspace.
int32_t chunk_of(int32_t m, int32_t n, int32_t nrhs) {
  const double per = 8.0 * ((double)m * n * 2 + (double)n * n + (double)m * nrhs + n * nrhs + 2 * n);
  return (int32_t)std::max(1.0, std::min(4096.0, 1.5e9 / per));
}

int32_t check_rows(const lps_abc_rows *rows, const lps_abc_modes *modes) {
  if (!rows || !modes) return fail("rows and modes are required");
  if (rows->count < 0 || rows->max_cloud <= 0 || modes->count <= 0)
    return fail("rows->count >= 0, rows->max_cloud > 0 and modes->count > 0 are required");
  if (rows->count > 0 && (!rows->xy || !rows->tau || !rows->cloud_len || !rows->cloud_xy))
    return fail("rows->xy, tau, cloud_len and cloud_xy are required");
  if (!modes->a1 || !modes->a2 || !modes->om_p || !modes->om_s || !modes->g_p || !modes->g_s ||
      !modes->s_r || !modes->s_i)
    return fail("every mode array is required");
  for (int32_t r = 0; r < rows->count; r++)
    if (rows->cloud_len[r] < 1 || rows->cloud_len[r] > rows->max_cloud)
      return fail("cloud_len[" + std::to_string(r) + "] is outside 1 .. max_cloud");
  return LPS_OK;
}

// Uploads rows and modes; the device copies live as long as d.
hipError_t upload(Dev &d, const lps_abc_rows *rows, const lps_abc_modes *modes, RowsDev *R,
                  ModesDev *M) {
  const size_t C = rows->count, Mc = rows->max_cloud, nm = modes->count;
  R->count = rows->count;
  R->max_cloud = rows->max_cloud;
  double *xy, *tau, *cxy, *a1, *a2, *op, *os, *gp, *gs, *sr, *si;
  int32_t *cl;
  hipError_t e;
  if ((e = d.put(&xy, rows->xy, 2 * C)) != hipSuccess) return e;
  if ((e = d.put(&tau, rows->tau, C)) != hipSuccess) return e;
  if ((e = d.put(&cxy, rows->cloud_xy, 2 * C * Mc)) != hipSuccess) return e;
  if ((e = d.put(&cl, rows->cloud_len, C)) != hipSuccess) return e;
  if ((e = d.put(&a1, modes->a1, nm)) != hipSuccess) return e;
  if ((e = d.put(&a2, modes->a2, nm)) != hipSuccess) return e;
  if ((e = d.put(&op, modes->om_p, nm)) != hipSuccess) return e;
  if ((e = d.put(&os, modes->om_s, nm)) != hipSuccess) return e;
  if ((e = d.put(&gp, modes->g_p, 2 * nm)) != hipSuccess) return e;
  if ((e = d.put(&gs, modes->g_s, 2 * nm)) != hipSuccess) return e;
  if ((e = d.put(&sr, modes->s_r, 2 * nm)) != hipSuccess) return e;
  if ((e = d.put(&si, modes->s_i, 2 * nm)) != hipSuccess) return e;
  R->xy = xy; R->tau = tau; R->cloud_xy = cxy; R->cloud_len = cl;
  M->count = modes->count;
  M->a1 = a1; M->a2 = a2; M->om_p = op; M->om_s = os; M->g_p = gp; M->g_s = gs; M->s_r = sr;
  M->s_i = si;
  return hipSuccess;
}

int32_t open_handle(int32_t device, Handle &H) {
  ABC_TRY(hipSetDevice(device));
  ABC_TRY(hipStreamCreate(&H.s));
  if (rocblas_create_handle(&H.h) != rocblas_status_success || !H.h)
    return fail("rocblas_create_handle failed");
  if (rocblas_set_stream(H.h, H.s) != rocblas_status_success) return fail("rocblas_set_stream failed");
  return LPS_OK;
}

}  // namespace

extern "C" {

int32_t lps_abc_system(int32_t device, const lps_abc_rows *rows, const lps_abc_modes *modes,
                       double dt, double theta_coeff, int32_t kind, int32_t first, int32_t count,
                       double *A, double *B) {
  int32_t rc = check_rows(rows, modes);
  if (rc != LPS_OK) return rc;
  if (kind < 0 || kind > 2) return fail("kind must be 0 (G_u), 1 (G_v) or 2 (G_theta)");
  if (first < 0 || count < 0 || first + count > rows->count || !A || !B)
    return fail("rows first .. first + count - 1 must exist, and A and B be set");
  Handle H;
  if ((rc = open_handle(device, H)) != LPS_OK) return rc;
  Dev d;
  RowsDev R;
  ModesDev M;
  ABC_TRY(upload(d, rows, modes, &R, &M));
  const int32_t m = 4 * modes->count, n = (kind == kTheta ? 2 : 4) * rows->max_cloud;
  const int32_t nrhs = kind == kTheta ? 1 : 2;
  double *dA, *dB;
  ABC_TRY(d.get(&dA, (size_t)count * m * n));
  ABC_TRY(d.get(&dB, (size_t)count * m * nrhs));
  if (count > 0)
    hipLaunchKernelGGL(k_abc_system, dim3(count), dim3(256), 0, H.s, R, M, dt, theta_coeff, kind,
                       first, count, dA, dB);
  ABC_TRY(hipGetLastError());
  ABC_TRY(hipStreamSynchronize(H.s));
  ABC_TRY(hipMemcpy(A, dA, sizeof(double) * count * m * n, hipMemcpyDeviceToHost));
  ABC_TRY(hipMemcpy(B, dB, sizeof(double) * count * m * nrhs, hipMemcpyDeviceToHost));
  return LPS_OK;
}

// dgelsd's threshold (LAPACK dlalsd): an rcond <= 0 or >= 1 means the machine epsilon 2^-53;
// singular values s_i <= rcond * s_1 are dropped. NaN stays NaN (the callers refuse it).
static double lapack_rcond(double rcond) {
  if (rcond != rcond) return rcond;
  return (rcond <= 0.0 || rcond >= 1.0) ? 0x1p-53 : rcond;
}

int32_t lps_abc_operators(int32_t device, const lps_abc_rows *rows, const lps_abc_modes *modes,
                          double dt, double theta_coeff, double rcond, double *gu, double *gv,
                          double *gth, int32_t *rank) {
  rcond = lapack_rcond(rcond);
  if (rcond != rcond) return fail("rcond is NaN");
  int32_t rc = check_rows(rows, modes);
  if (rc != LPS_OK) return rc;
  if (!gu || !gv) return fail("gu and gv are required");
  Handle H;
  if ((rc = open_handle(device, H)) != LPS_OK) return rc;
  Dev d;
  RowsDev R;
  ModesDev M;
  ABC_TRY(upload(d, rows, modes, &R, &M));
  const int32_t C = rows->count, Mc = rows->max_cloud, m = 4 * modes->count;
  double *G[3];
  ABC_TRY(d.get(&G[0], (size_t)2 * C * 4 * Mc));
  ABC_TRY(d.get(&G[1], (size_t)2 * C * 4 * Mc));
  ABC_TRY(d.get(&G[2], (size_t)C * 2 * Mc));
  std::vector<int32_t> rk((size_t)3 * C, 0), part;
  for (int32_t kind = 0; kind < 3; kind++) {
    if (kind == kTheta && !gth) continue;
    const int32_t n = (kind == kTheta ? 2 : 4) * Mc, nrhs = kind == kTheta ? 1 : 2;
    const int32_t cap = std::min(std::max(C, 1), chunk_of(m, n, nrhs));
    Dev w;
    Solver sv;
    ABC_TRY(sv.init(w, m, n, nrhs, cap));
    for (int32_t first = 0; first < C; first += cap) {
      const int32_t nb = std::min(cap, C - first);
      hipLaunchKernelGGL(k_abc_system, dim3(nb), dim3(256), 0, H.s, R, M, dt, theta_coeff, kind,
                         first, nb, sv.A, sv.B);
      ABC_TRY(hipGetLastError());
      if ((rc = sv.solve(H, nb, rcond)) != LPS_OK) return rc;
      hipLaunchKernelGGL(k_abc_scatter, dim3(nb), dim3(256), 0, H.s, kind, first, nb, Mc,
                         R.cloud_len, sv.X, G[kind]);
      ABC_TRY(hipGetLastError());
      part.resize(nb);
      ABC_TRY(hipMemcpyAsync(part.data(), sv.rank, sizeof(int32_t) * nb, hipMemcpyDeviceToHost,
                             H.s));
      ABC_TRY(hipStreamSynchronize(H.s));
      for (int32_t p = 0; p < nb; p++) rk[(size_t)(first + p) * 3 + kind] = part[p];
    }
  }
  ABC_TRY(hipStreamSynchronize(H.s));
  ABC_TRY(hipMemcpy(gu, G[0], sizeof(double) * 2 * C * 4 * Mc, hipMemcpyDeviceToHost));
  ABC_TRY(hipMemcpy(gv, G[1], sizeof(double) * 2 * C * 4 * Mc, hipMemcpyDeviceToHost));
  if (gth) ABC_TRY(hipMemcpy(gth, G[2], sizeof(double) * C * 2 * Mc, hipMemcpyDeviceToHost));
  if (rank) std::copy(rk.begin(), rk.end(), rank);
  return LPS_OK;
}

int32_t lps_lstsq_batched(int32_t device, int32_t m, int32_t n, int32_t nrhs, int32_t batch,
                          const double *A, const double *B, double rcond, double *X,
                          int32_t *rank) {
  if (m <= 0 || n <= 0 || nrhs <= 0 || batch < 0 || !A || !B || !X)
    return fail("m, n, nrhs > 0, batch >= 0 and A, B, X are required");
  rcond = lapack_rcond(rcond);
  if (rcond != rcond) return fail("rcond is NaN");
  Handle H;
  int32_t rc = open_handle(device, H);
  if (rc != LPS_OK) return rc;
  const int32_t cap = std::min(std::max(batch, 1), chunk_of(m, n, nrhs));
  Dev w;
  Solver sv;
  ABC_TRY(sv.init(w, m, n, nrhs, cap));
  for (int32_t first = 0; first < batch; first += cap) {
    const int32_t nb = std::min(cap, batch - first);
    ABC_TRY(hipMemcpyAsync(sv.A, A + (size_t)first * m * n, sizeof(double) * nb * m * n,
                           hipMemcpyHostToDevice, H.s));
    ABC_TRY(hipMemcpyAsync(sv.B, B + (size_t)first * m * nrhs, sizeof(double) * nb * m * nrhs,
                           hipMemcpyHostToDevice, H.s));
    if ((rc = sv.solve(H, nb, rcond)) != LPS_OK) return rc;
    ABC_TRY(hipMemcpyAsync(X + (size_t)first * n * nrhs, sv.X, sizeof(double) * nb * n * nrhs,
                           hipMemcpyDeviceToHost, H.s));
    if (rank)
      ABC_TRY(hipMemcpyAsync(rank + first, sv.rank, sizeof(int32_t) * nb, hipMemcpyDeviceToHost,
                             H.s));
    ABC_TRY(hipStreamSynchronize(H.s));
  }
  return LPS_OK;
}

}  // extern "C"
