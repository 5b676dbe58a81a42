// C ABI of the LPS-ABC time loop: context setup, step loop, output block, fetch.
//
// Setup follows src/main.cpp:178-684, the time loop :718-1164 (see lps_kernels.hip for the
// passes). Buffer rotation (src/main.cpp:1145-1163) is index arithmetic on fixed buffers:
// at step s (s completed steps) U^n lives in U[s % 3], U^{n-1} in U[(s + 2) % 3] and U^{n+1} is
// written to U[(s + 1) % 3]; V likewise; A and the prediction alternate between two buffers; the
// dilatation history uses three (theta of step s is written to TH[(s + 1) % 3], ThetaN is
// TH[s % 3], ThetaNm1 is TH[(s + 2) % 3]). The pattern repeats every 6 steps, so a captured
// 30-step graph replays from any step that is a multiple of 30.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/lps_amd.h"
#include "lps_internal.h"

using namespace lps;

namespace {
thread_local std::string g_create_error;

struct Dev2 {
  double *x = nullptr, *y = nullptr;
};
}  // namespace

struct lps_ctx {
  lps_options opt{};
  hipStream_t stream = nullptr;
  std::string err;
  int32_t n = 0, n_int = 0, n_abc = 0, M = 0;
  double E = 0, nu = 0, rho = 0, dt = 0, delta = 0, dx = 0, dy = 0;
  double lam = 0, mu = 0, kappa = 0, m_ref = 0, exclude_r = -1, E0 = -1;
  bool fast = false;
  int64_t step = 0;
  std::vector<int32_t> perm_h;  // new -> original
  HorizonBuild hb;
  std::vector<void *> allocs;
  // device arrays
  double *x = nullptr, *y = nullptr, *m = nullptr, *bcoef = nullptr, *c2m = nullptr;
  int32_t *lat = nullptr;
  Geo *geo = nullptr;
  int32_t geo_k = 0;
  int8_t *is_bc = nullptr;
  int32_t *cloud = nullptr;
  double *gu = nullptr, *gv = nullptr, *gth = nullptr;
  Dev2 U[3], V[3], A[2], P[2], IC, ICD;
  double *TH[3] = {nullptr, nullptr, nullptr};
  double *acoef = nullptr;
  DiagPartial *partials = nullptr, *dout = nullptr, *hout = nullptr;
  int32_t n_partials = 0;
  // graph of 30 steps from a step that is a multiple of 30
  hipGraphExec_t graph30 = nullptr;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;   // 3 per step
  int32_t ev_used = 0;
  double t_theta = 0, t_force = 0, t_step = 0;
  int64_t t_count = 0;

  template <class T>
  hipError_t alloc(T **p, size_t count) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) allocs.push_back(*p);
    return e;
  }
  int32_t fail(int32_t code, const std::string &msg) {
    err = msg;
    return code;
  }
  int32_t hip_fail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return LPS_ERR_HIP;
  }
  StepArgs args() const;
  StepState state(int64_t s) const;
};

#define CTX_TRY(ctx, call)                                   \
  do {                                                       \
    hipError_t e_ = (call);                                  \
    if (e_ != hipSuccess) return (ctx)->hip_fail(e_, #call); \
  } while (0)

StepArgs lps_ctx::args() const {
  StepArgs a{};
  a.n = n; a.n_int = n_int; a.n_abc = n_abc; a.M = M;
  a.dt = dt; a.rho = rho; a.lam = lam; a.mu = mu; a.kappa = kappa;
  a.delta = delta; a.dx = dx; a.dy = dy;
  a.x = x; a.y = y; a.lat = lat; a.geo = geo; a.geo_k = geo_k;
  a.row_ptr = hb.row_ptr; a.col = hb.col;
  a.sell_off = hb.sell_off; a.sell_w = hb.sell_w; a.sell_col = hb.sell_col;
  a.m = m; a.bcoef = bcoef; a.c2m = c2m; a.is_bc = is_bc;
  a.cloud = cloud; a.gu = gu; a.gv = gv; a.gth = gth;
  a.theta_method = opt.theta_method;
  a.abc_on = opt.abc_enabled;
  return a;
}

StepState lps_ctx::state(int64_t s) const {
  StepState st{};
  const Dev2 &Pc = P[s % 2], &Pn = P[(s + 1) % 2];
  const Dev2 &Uc = U[s % 3], &Um = U[(s + 2) % 3], &Un = U[(s + 1) % 3];
  const Dev2 &Vc = V[s % 3], &Vm = V[(s + 2) % 3], &Vn = V[(s + 1) % 3];
  const Dev2 &Ac = A[s % 2], &An = A[(s + 1) % 2];
  st.px = Pc.x; st.py = Pc.y;
  st.ux = Uc.x; st.uy = Uc.y; st.uxm = Um.x; st.uym = Um.y;
  st.vx = Vc.x; st.vy = Vc.y; st.vxm = Vm.x; st.vym = Vm.y;
  st.ax = Ac.x; st.ay = Ac.y;
  st.theta = TH[(s + 1) % 3];
  st.acoef = acoef;
  st.thn = TH[s % 3];
  st.thnm1 = TH[(s + 2) % 3];
  st.ux1 = Un.x; st.uy1 = Un.y; st.vx1 = Vn.x; st.vy1 = Vn.y; st.ax1 = An.x; st.ay1 = An.y;
  st.pnx = Pn.x; st.pny = Pn.y;
  return st;
}

static double beta_host(double xi, double R, double dx) {
  if (xi <= R - dx / 2.0) return 1.0;
  else if (xi <= R + dx / 2.0) return (R + dx / 2.0 - xi) / dx;
  else return 0.0;
}

extern "C" {

static int32_t reset_state(lps_ctx *c);
static int32_t flush_events(lps_ctx *c);

void lps_default_options(lps_options *opt) {
  opt->theta_method = LPS_THETA_GTHETA;
  opt->abc_enabled = 1;
  opt->exact = 0;
  opt->use_graph = 1;
  opt->device = 0;
  opt->horizon_radius = 0.0;
}

const char *lps_global_error(void) { return g_create_error.c_str(); }
const char *lps_last_error(const lps_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }
int64_t lps_step_index(const lps_ctx *ctx) { return ctx->step; }
double lps_m_ref(const lps_ctx *ctx) { return ctx->m_ref; }

void lps_destroy(lps_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->graph30) (void)hipGraphExecDestroy(ctx->graph30);
  for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
  for (void *p : ctx->allocs) (void)hipFree(p);
  free_horizons(&ctx->hb);
  if (ctx->hout) (void)hipHostFree(ctx->hout);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

static int32_t create_impl(lps_ctx *c, const lps_problem *p) {
  const int32_t N = p->node_no, C = p->constrains_no, M = p->max_cloud_size;
  if (N <= 0 || !p->nodes || !p->bctypes || !p->icvalues || !p->icdvalues)
    return c->fail(LPS_ERR_ARG, "node_no must be positive and nodes/bctypes/icvalues/icdvalues set");
  if (C < 0 || (C > 0 && M <= 0)) return c->fail(LPS_ERR_ARG, "inconsistent constrains_no / max_cloud_size");
  const bool abc_on = c->opt.abc_enabled && C > 0;
  if (abc_on && (!p->inf_nodes || !p->iclouds || !p->ci_nodes_u || !p->ci_nodes_v))
    return c->fail(LPS_ERR_ARG, "ABC rows need inf_nodes, iclouds, ci_nodes_u and ci_nodes_v");
  if (abc_on && c->opt.theta_method == LPS_THETA_GTHETA && !p->ci_nodes_theta)
    return c->fail(LPS_ERR_ARG, "the G_theta method needs ci_nodes_theta");
  if (c->opt.theta_method != LPS_THETA_GTHETA && c->opt.theta_method != LPS_THETA_IDW)
    return c->fail(LPS_ERR_ARG, "theta_method must be LPS_THETA_GTHETA or LPS_THETA_IDW");
  c->n = N;
  c->M = abc_on ? M : 0;
  c->E = p->E; c->nu = p->nu; c->rho = p->rho; c->dt = p->dt;
  c->delta = p->delta; c->dx = p->dx; c->dy = p->dy;
  // Lame constants and their FATAL consistency check (src/main.cpp:215-217, 310-322)
  c->lam = c->E * c->nu / ((1.0 + c->nu) * (1.0 - 2.0 * c->nu));
  c->mu = c->E / (2.0 * (1.0 + c->nu));
  c->kappa = c->lam + c->mu;
  {
    const double E_check = c->mu * (3.0 * c->lam + 2.0 * c->mu) / (c->lam + c->mu);
    const double Nu_check = c->lam / (2.0 * (c->lam + c->mu));
    if (!(std::abs(E_check - c->E) / c->E <= 1e-10) || !(std::abs(Nu_check - c->nu) <= 1e-10))
      return c->fail(LPS_ERR_CONSISTENCY, "Lame parameter consistency check failed");
  }
  // ABC rows: validate, pick the active ones (bctypes[2 node] == 1), in row order
  std::vector<int32_t> active_rows;
  std::vector<int8_t> seg(N, 0);
  std::vector<int32_t> seg_rank(N, 0);
  if (abc_on) {
    std::unordered_set<int32_t> seen;
    for (int32_t a = 0; a < C; a++) {
      const int32_t node = p->inf_nodes[a];
      if (node < 0 || node >= N) return c->fail(LPS_ERR_ARG, "infNodesIndex entry out of range");
      if (!seen.insert(node).second) return c->fail(LPS_ERR_ARG, "infNodesIndex lists a node twice");
      for (int32_t j = 0; j < M; j++) {
        const int32_t k = p->iclouds[(int64_t)a * M + j];
        if (k < -1 || k >= N) return c->fail(LPS_ERR_ARG, "iClouds entry out of range");
      }
      if (p->bctypes[2 * (int64_t)node] == 1) {
        seg[node] = 1;
        seg_rank[node] = (int32_t)active_rows.size();
        active_rows.push_back(a);
      }
    }
  }
  c->n_abc = (int32_t)active_rows.size();
  c->n_int = N - c->n_abc;
  const hipStream_t s = c->stream;

  // coordinates (original order) -> device, horizon build
  double *xo, *yo;
  int8_t *dseg;
  int32_t *drank;
  CTX_TRY(c, c->alloc(&xo, N));
  CTX_TRY(c, c->alloc(&yo, N));
  CTX_TRY(c, c->alloc(&dseg, N));
  CTX_TRY(c, c->alloc(&drank, N));
  {
    std::vector<double> hx(N), hy(N);
    for (int32_t i = 0; i < N; i++) {
      hx[i] = p->nodes[2 * (int64_t)i];
      hy[i] = p->nodes[2 * (int64_t)i + 1];
    }
    CTX_TRY(c, hipMemcpyAsync(xo, hx.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    CTX_TRY(c, hipMemcpyAsync(yo, hy.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    CTX_TRY(c, hipMemcpyAsync(dseg, seg.data(), N, hipMemcpyHostToDevice, s));
    CTX_TRY(c, hipMemcpyAsync(drank, seg_rank.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
    CTX_TRY(c, hipStreamSynchronize(s));
  }
  const double search_r = c->opt.horizon_radius > 0.0 ? c->opt.horizon_radius
                                                      : c->delta + c->dx / 2.0 + 1e-5;  // src/main.cpp:352
  bool overflow = false;
  CTX_TRY(c, build_horizons(N, xo, yo, dseg, drank, search_r, kMaxHorizon, &c->hb, &overflow, s));
  if (overflow) return c->fail(LPS_ERR_HORIZON, "found_node_no > max_node_no; increase max_node_no.");
  c->perm_h.resize(N);
  CTX_TRY(c, hipMemcpyAsync(c->perm_h.data(), c->hb.perm, sizeof(int32_t) * N, hipMemcpyDeviceToHost, s));
  CTX_TRY(c, hipStreamSynchronize(s));
  std::vector<int32_t> iperm(N);
  for (int32_t k = 0; k < N; k++) iperm[c->perm_h[k]] = k;

  // permuted coordinates, bc flags
  std::vector<double> hx(N), hy(N);
  std::vector<int8_t> hbc(N);
  for (int32_t k = 0; k < N; k++) {
    const int32_t o = c->perm_h[k];
    hx[k] = p->nodes[2 * (int64_t)o];
    hy[k] = p->nodes[2 * (int64_t)o + 1];
    hbc[k] = p->bctypes[2 * (int64_t)o] == 1;
  }
  CTX_TRY(c, c->alloc(&c->x, N));
  CTX_TRY(c, c->alloc(&c->y, N));
  CTX_TRY(c, c->alloc(&c->is_bc, N));
  CTX_TRY(c, hipMemcpyAsync(c->x, hx.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
  CTX_TRY(c, hipMemcpyAsync(c->y, hy.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
  CTX_TRY(c, hipMemcpyAsync(c->is_bc, hbc.data(), N, hipMemcpyHostToDevice, s));

  // m_i on the device, m_ref and the ABC override on the host (src/main.cpp:408-480)
  CTX_TRY(c, c->alloc(&c->m, N));
  {
    StepArgs a = c->args();
    CTX_TRY(c, launch_weighted_volume(a, c->m, s));
  }
  std::vector<double> hm(N);
  CTX_TRY(c, hipMemcpyAsync(hm.data(), c->m, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  CTX_TRY(c, hipStreamSynchronize(s));
  double m_max = 0.0;
  for (int32_t k = 0; k < N; k++)
    if (hm[k] > 0.0) m_max = std::max(m_max, hm[k]);
  c->m_ref = m_max;
  if (abc_on)
    for (int32_t a = 0; a < C; a++) hm[iperm[p->inf_nodes[a]]] = c->m_ref;
  CTX_TRY(c, hipMemcpyAsync(c->m, hm.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
  // LPS constants and the lambda_PD identity (src/main.cpp:485-504)
  {
    const double AlphaLPS = 8.0 * c->mu / c->m_ref;
    const double LambdaPD = 2.0 * (c->lam - c->mu) / c->m_ref;
    const double lpd_check = 2.0 * c->kappa / c->m_ref - AlphaLPS / 2.0;
    if (std::abs(lpd_check - LambdaPD) > 1e-15 * std::abs(LambdaPD + 1e-30))
      return c->fail(LPS_ERR_CONSISTENCY, "lambda_PD consistency check failed");
  }
  // output / energy exclusion radius (src/main.cpp:518-530)
  {
    double domainRadius = 0.0;
    for (int32_t i = 0; i < N; i++) {
      const double x = p->nodes[2 * (int64_t)i], y = p->nodes[2 * (int64_t)i + 1];
      const double r = std::sqrt(x * x + y * y);
      if (r > domainRadius) domainRadius = r;
    }
    c->exclude_r = abc_on ? domainRadius - c->delta - c->dx / 100.0 : -1.0;
  }

  // fast path: lattice coordinates and the offset table
  c->fast = !c->opt.exact;
  if (c->fast) {
    double xmin = hx[0], ymin = hy[0];
    for (int32_t k = 1; k < N; k++) { xmin = std::min(xmin, hx[k]); ymin = std::min(ymin, hy[k]); }
    std::vector<int32_t> hl(N);
    for (int32_t k = 0; k < N && c->fast; k++) {
      const double fx = (hx[k] - xmin) / c->dx, fy = (hy[k] - ymin) / c->dy;
      const long long ix = std::llround(fx), iy = std::llround(fy);
      if (ix < 0 || iy < 0 || ix > 32767 || iy > 65535 ||
          std::abs(xmin + (double)ix * c->dx - hx[k]) > 1e-9 * c->dx ||
          std::abs(ymin + (double)iy * c->dy - hy[k]) > 1e-9 * c->dy)
        c->fast = false;
      else
        hl[k] = (int32_t)((ix << 16) | iy);
    }
    const int32_t K = (int32_t)std::ceil(search_r / std::min(c->dx, c->dy));
    if (K > kMaxGeoK) c->fast = false;
    if (c->fast) {
      c->geo_k = K;
      const int32_t W = 2 * K + 1;
      std::vector<Geo> g(W * W);
      for (int32_t a = -K; a <= K; a++)
        for (int32_t b = -K; b <= K; b++) {
          Geo &e = g[(a + K) * W + (b + K)];
          const double xi_x = a * c->dx, xi_y = b * c->dy;
          const double xn = std::sqrt(xi_x * xi_x + xi_y * xi_y);
          if (xn == 0.0) { e.hx = e.hy = e.g1 = e.g2 = 0.0; continue; }
          const double bt = beta_host(xn, c->delta, c->dx);
          e.hx = xi_x / xn;
          e.hy = xi_y / xn;
          e.g1 = xn * bt * c->dx * c->dy;
          e.g2 = bt * c->dx * c->dy;
        }
      std::vector<double> hb(N), hc(N);
      for (int32_t k = 0; k < N; k++) {
        hb[k] = hm[k] > 0.0 ? 8.0 * c->mu / hm[k] : -1.0;
        hc[k] = hm[k] > 0.0 ? 2.0 / hm[k] : 0.0;
      }
      CTX_TRY(c, c->alloc(&c->lat, N));
      CTX_TRY(c, c->alloc(&c->geo, g.size()));
      CTX_TRY(c, c->alloc(&c->bcoef, N));
      CTX_TRY(c, c->alloc(&c->c2m, N));
      CTX_TRY(c, c->alloc(&c->acoef, N));
      CTX_TRY(c, hipMemcpyAsync(c->lat, hl.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
      CTX_TRY(c, hipMemcpyAsync(c->geo, g.data(), sizeof(Geo) * g.size(), hipMemcpyHostToDevice, s));
      CTX_TRY(c, hipMemcpyAsync(c->bcoef, hb.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
      CTX_TRY(c, hipMemcpyAsync(c->c2m, hc.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
      CTX_TRY(c, hipStreamSynchronize(s));
    }
  }

  // ABC rows on the device, active rows only, permuted cloud ids
  if (c->n_abc > 0) {
    const int32_t R = c->n_abc;
    std::vector<int32_t> hcl((size_t)R * M);
    std::vector<double> hgu((size_t)R * 8 * M), hgv((size_t)R * 8 * M), hgt;
    if (c->opt.theta_method == LPS_THETA_GTHETA) hgt.resize((size_t)R * 2 * M);
    for (int32_t r = 0; r < R; r++) {
      const int64_t a = active_rows[r];
      for (int32_t j = 0; j < M; j++) {
        const int32_t k = p->iclouds[a * M + j];
        hcl[(size_t)r * M + j] = (k < 0) ? -1 : iperm[k];
      }
      std::memcpy(&hgu[(size_t)r * 8 * M], p->ci_nodes_u + a * 8 * M, sizeof(double) * 8 * M);
      std::memcpy(&hgv[(size_t)r * 8 * M], p->ci_nodes_v + a * 8 * M, sizeof(double) * 8 * M);
      if (!hgt.empty())
        std::memcpy(&hgt[(size_t)r * 2 * M], p->ci_nodes_theta + a * 2 * M, sizeof(double) * 2 * M);
    }
    CTX_TRY(c, c->alloc(&c->cloud, hcl.size()));
    CTX_TRY(c, c->alloc(&c->gu, hgu.size()));
    CTX_TRY(c, c->alloc(&c->gv, hgv.size()));
    CTX_TRY(c, hipMemcpyAsync(c->cloud, hcl.data(), sizeof(int32_t) * hcl.size(), hipMemcpyHostToDevice, s));
    CTX_TRY(c, hipMemcpyAsync(c->gu, hgu.data(), sizeof(double) * hgu.size(), hipMemcpyHostToDevice, s));
    CTX_TRY(c, hipMemcpyAsync(c->gv, hgv.data(), sizeof(double) * hgv.size(), hipMemcpyHostToDevice, s));
    if (!hgt.empty()) {
      CTX_TRY(c, c->alloc(&c->gth, hgt.size()));
      CTX_TRY(c, hipMemcpyAsync(c->gth, hgt.data(), sizeof(double) * hgt.size(), hipMemcpyHostToDevice, s));
    }
    CTX_TRY(c, hipStreamSynchronize(s));
  }

  // state buffers; the initial conditions stay on the device for lps_reset
  for (int b = 0; b < 3; b++) {
    CTX_TRY(c, c->alloc(&c->U[b].x, N)); CTX_TRY(c, c->alloc(&c->U[b].y, N));
    CTX_TRY(c, c->alloc(&c->V[b].x, N)); CTX_TRY(c, c->alloc(&c->V[b].y, N));
    CTX_TRY(c, c->alloc(&c->TH[b], N));
  }
  for (int b = 0; b < 2; b++) {
    CTX_TRY(c, c->alloc(&c->A[b].x, N)); CTX_TRY(c, c->alloc(&c->A[b].y, N));
    CTX_TRY(c, c->alloc(&c->P[b].x, N)); CTX_TRY(c, c->alloc(&c->P[b].y, N));
  }
  CTX_TRY(c, c->alloc(&c->IC.x, N)); CTX_TRY(c, c->alloc(&c->IC.y, N));
  CTX_TRY(c, c->alloc(&c->ICD.x, N)); CTX_TRY(c, c->alloc(&c->ICD.y, N));
  {
    std::vector<double> ux(N), uy(N), vx(N), vy(N);
    for (int32_t k = 0; k < N; k++) {
      const int64_t o = c->perm_h[k];
      ux[k] = p->icvalues[2 * o]; uy[k] = p->icvalues[2 * o + 1];
      vx[k] = p->icdvalues[2 * o]; vy[k] = p->icdvalues[2 * o + 1];
    }
    CTX_TRY(c, hipMemcpyAsync(c->IC.x, ux.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    CTX_TRY(c, hipMemcpyAsync(c->IC.y, uy.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    CTX_TRY(c, hipMemcpyAsync(c->ICD.x, vx.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    CTX_TRY(c, hipMemcpyAsync(c->ICD.y, vy.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    CTX_TRY(c, hipStreamSynchronize(s));
  }
  // output block buffers
  c->n_partials = diag_partials(N);
  CTX_TRY(c, c->alloc(&c->partials, c->n_partials));
  CTX_TRY(c, c->alloc(&c->dout, 1));
  CTX_TRY(c, hipHostMalloc((void **)&c->hout, sizeof(DiagPartial)));
  return reset_state(c);
}

// Initial state (src/main.cpp:624-634), theta history (:664-682) and the prediction of step 1.
static int32_t reset_state(lps_ctx *c) {
  const int32_t N = c->n;
  const hipStream_t s = c->stream;
  const size_t bytes = sizeof(double) * N;
  for (int b = 0; b < 3; b++) {
    CTX_TRY(c, hipMemsetAsync(c->U[b].x, 0, bytes, s));
    CTX_TRY(c, hipMemsetAsync(c->U[b].y, 0, bytes, s));
    CTX_TRY(c, hipMemsetAsync(c->V[b].x, 0, bytes, s));
    CTX_TRY(c, hipMemsetAsync(c->V[b].y, 0, bytes, s));
    CTX_TRY(c, hipMemsetAsync(c->TH[b], 0, bytes, s));
  }
  for (int b = 0; b < 2; b++) {
    CTX_TRY(c, hipMemsetAsync(c->A[b].x, 0, bytes, s));
    CTX_TRY(c, hipMemsetAsync(c->A[b].y, 0, bytes, s));
  }
  CTX_TRY(c, hipMemcpyAsync(c->U[0].x, c->IC.x, bytes, hipMemcpyDeviceToDevice, s));
  CTX_TRY(c, hipMemcpyAsync(c->U[0].y, c->IC.y, bytes, hipMemcpyDeviceToDevice, s));
  CTX_TRY(c, hipMemcpyAsync(c->V[0].x, c->ICD.x, bytes, hipMemcpyDeviceToDevice, s));
  CTX_TRY(c, hipMemcpyAsync(c->V[0].y, c->ICD.y, bytes, hipMemcpyDeviceToDevice, s));
  StepArgs a = c->args();
  // theta history from the IC, m already overridden (src/main.cpp:664-682)
  if (c->opt.theta_method == LPS_THETA_GTHETA)
    CTX_TRY(c, launch_theta_field(a, c->U[0].x, c->U[0].y, c->TH[0], s));
  // prediction of step 1 (later ones come from the force pass)
  CTX_TRY(c, launch_predict(a, c->U[0].x, c->U[0].y, c->V[0].x, c->V[0].y, c->A[0].x, c->A[0].y,
                            c->P[0].x, c->P[0].y, s));
  CTX_TRY(c, hipStreamSynchronize(s));
  c->step = 0;
  c->E0 = -1.0;
  return LPS_OK;
}

int32_t lps_reset(lps_ctx *c) {
  int32_t rc = flush_events(c);
  if (rc != LPS_OK) return rc;
  return reset_state(c);
}

int32_t lps_create(const lps_problem *problem, const lps_options *options, lps_ctx **out) {
  *out = nullptr;
  if (!problem) {
    g_create_error = "problem is NULL";
    return LPS_ERR_ARG;
  }
  lps_ctx *c = new lps_ctx();
  if (options) c->opt = *options;
  else lps_default_options(&c->opt);
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    g_create_error = "no HIP device available";
    delete c;
    return LPS_ERR_HIP;
  }
  if (c->opt.device < 0 || c->opt.device >= ndev) {
    g_create_error = "device ordinal out of range";
    delete c;
    return LPS_ERR_ARG;
  }
  e = hipSetDevice(c->opt.device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_create_error = std::string("HIP setup failed: ") + hipGetErrorString(e);
    delete c;
    return LPS_ERR_HIP;
  }
  const int32_t rc = create_impl(c, problem);
  if (rc != LPS_OK) {
    g_create_error = c->err;
    lps_destroy(c);
    return rc;
  }
  *out = c;
  return LPS_OK;
}

static hipError_t step_once(lps_ctx *c, bool record_events) {
  const StepArgs a = c->args();
  const StepState st = c->state(c->step);
  const bool exact = !c->fast;
  hipEvent_t *ev = nullptr;
  if (record_events) {
    ev = &c->ev[3 * (size_t)c->ev_used];
    hipError_t e = hipEventRecord(ev[0], c->stream);
    if (e != hipSuccess) return e;
  }
  hipError_t e = launch_theta(a, st, exact, c->stream);
  if (e == hipSuccess && a.abc_on && a.n_abc > 0 && a.theta_method == LPS_THETA_IDW)
    e = launch_idw(a, st, exact, c->stream);
  if (e != hipSuccess) return e;
  if (record_events && (e = hipEventRecord(ev[1], c->stream)) != hipSuccess) return e;
  e = launch_force(a, st, exact, c->stream);
  if (e != hipSuccess) return e;
  if (record_events) {
    if ((e = hipEventRecord(ev[2], c->stream)) != hipSuccess) return e;
    c->ev_used++;
  }
  c->step++;
  return hipSuccess;
}

static int32_t flush_events(lps_ctx *c) {
  if (c->ev_used == 0) return LPS_OK;
  CTX_TRY(c, hipEventSynchronize(c->ev[3 * (size_t)(c->ev_used - 1) + 2]));
  for (int32_t k = 0; k < c->ev_used; k++) {
    float t01 = 0, t12 = 0, t02 = 0;
    hipEvent_t *e = &c->ev[3 * (size_t)k];
    CTX_TRY(c, hipEventElapsedTime(&t01, e[0], e[1]));
    CTX_TRY(c, hipEventElapsedTime(&t12, e[1], e[2]));
    CTX_TRY(c, hipEventElapsedTime(&t02, e[0], e[2]));
    c->t_theta += t01;
    c->t_force += t12;
    c->t_step += t02;
    c->t_count++;
  }
  c->ev_used = 0;
  return LPS_OK;
}

int32_t lps_set_timing(lps_ctx *c, int32_t on) {
  int32_t rc = flush_events(c);
  if (rc != LPS_OK) return rc;
  c->timing = on != 0;
  c->t_theta = c->t_force = c->t_step = 0;
  c->t_count = 0;
  if (c->timing && c->ev.empty()) {
    c->ev.resize(3 * 30);
    for (auto &e : c->ev) CTX_TRY(c, hipEventCreate(&e));
  }
  return LPS_OK;
}

int32_t lps_pass_times(lps_ctx *c, double *out3) {
  int32_t rc = flush_events(c);
  if (rc != LPS_OK) return rc;
  const double n = c->t_count > 0 ? (double)c->t_count : 1.0;
  out3[0] = c->t_theta / n;
  out3[1] = c->t_force / n;
  out3[2] = c->t_step / n;
  return LPS_OK;
}

static int32_t build_graph30(lps_ctx *c) {
  hipGraph_t g = nullptr;
  CTX_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int64_t s0 = c->step;
  hipError_t err = hipSuccess;
  for (int k = 0; k < 30 && err == hipSuccess; k++) err = step_once(c, false);
  c->step = s0;
  hipError_t e2 = hipStreamEndCapture(c->stream, &g);
  if (err != hipSuccess) return c->hip_fail(err, "capture of the step loop");
  if (e2 != hipSuccess) return c->hip_fail(e2, "hipStreamEndCapture");
  hipError_t e3 = hipGraphInstantiate(&c->graph30, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e3 != hipSuccess) {
    c->graph30 = nullptr;
    return c->hip_fail(e3, "hipGraphInstantiate");
  }
  return LPS_OK;
}

int32_t lps_step(lps_ctx *c, int32_t nsteps) {
  if (nsteps < 0) return c->fail(LPS_ERR_ARG, "nsteps must be >= 0");
  int32_t done = 0;
  while (done < nsteps) {
    const int32_t left = nsteps - done;
    if (c->opt.use_graph && !c->timing && c->step % 30 == 0 && left >= 30) {
      if (!c->graph30) {
        int32_t rc = build_graph30(c);
        if (rc != LPS_OK) return rc;
      }
      CTX_TRY(c, hipGraphLaunch(c->graph30, c->stream));
      c->step += 30;
      done += 30;
      continue;
    }
    if (c->timing && c->ev_used == 30) {
      int32_t rc = flush_events(c);
      if (rc != LPS_OK) return rc;
    }
    CTX_TRY(c, step_once(c, c->timing));
    done++;
  }
  return LPS_OK;
}

int32_t lps_diagnostics(lps_ctx *c, lps_diag *out) {
  const int64_t s = c->step;
  const StepArgs a = c->args();
  const Dev2 &U = c->U[s % 3], &V = c->V[s % 3], &A = c->A[s % 2];
  const double *th = c->TH[s % 3];
  CTX_TRY(c, launch_diag(a, U.x, U.y, V.x, V.y, A.x, A.y, th, c->hb.perm, c->exclude_r, !c->fast,
                         c->partials, c->n_partials, c->dout, c->stream));
  CTX_TRY(c, hipMemcpyAsync(c->hout, c->dout, sizeof(DiagPartial), hipMemcpyDeviceToHost, c->stream));
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  const DiagPartial &d = *c->hout;
  out->step = s;
  out->t = (double)s * c->dt;
  out->energy = d.energy;
  out->energy_k = d.ek;
  out->energy_p = d.ep;
  out->px = d.px;
  out->py = d.py;
  out->max_u = d.max_u;
  out->max_v = d.max_v;
  out->max_a = d.max_a;
  out->theta_max = d.th_max;
  out->theta_min = d.th_min;
  out->ub_rms = c->n_abc > 0 ? std::sqrt(d.ub / c->n_abc) : 0.0;
  out->vb_rms = c->n_abc > 0 ? std::sqrt(d.vb / c->n_abc) : 0.0;
  out->max_u_node = d.max_u_orig == INT64_MAX ? -1 : (int32_t)d.max_u_orig;
  out->max_v_node = d.max_v_orig == INT64_MAX ? -1 : (int32_t)d.max_v_orig;
  out->de_rel = (c->E0 > 0.0) ? (d.energy - c->E0) / c->E0 : 0.0;
  return LPS_OK;
}

int32_t lps_run(lps_ctx *c, int32_t nsteps, lps_diag *records, int32_t max_records) {
  if (nsteps < 0) return c->fail(LPS_ERR_ARG, "nsteps must be >= 0");
  int32_t done = 0, nrec = 0;
  while (done < nsteps) {
    const int32_t to_out = 30 - (int32_t)(c->step % 30);
    const int32_t chunk = std::min(to_out, nsteps - done);
    int32_t rc = lps_step(c, chunk);
    if (rc != LPS_OK) return rc;
    done += chunk;
    if (c->step % 30 == 0) {
      lps_diag d;
      rc = lps_diagnostics(c, &d);
      if (rc != LPS_OK) return rc;
      if (c->E0 < 0.0) c->E0 = d.energy;  // src/main.cpp:1046
      d.de_rel = (c->E0 > 0.0) ? (d.energy - c->E0) / c->E0 : 0.0;
      if (records && nrec < max_records) records[nrec] = d;
      nrec++;
      if (std::isnan(d.energy) || std::isinf(d.energy) || std::isnan(d.max_u) || std::isinf(d.max_u))
        return c->fail(LPS_ERR_NONFINITE, "Energy or displacement is NaN/Inf at step " +
                                              std::to_string(c->step));
    }
  }
  if (c->timing) {
    int32_t rc = flush_events(c);
    if (rc != LPS_OK) return rc;
  }
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  return nrec;
}

int32_t lps_fetch(lps_ctx *c, double *u, double *v, double *a, double *theta, double *m) {
  const int64_t s = c->step;
  const int32_t N = c->n;
  std::vector<double> bx(N), by(N);
  auto pull2 = [&](const Dev2 &d, double *out) -> int32_t {
    if (!out) return LPS_OK;
    CTX_TRY(c, hipMemcpyAsync(bx.data(), d.x, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    CTX_TRY(c, hipMemcpyAsync(by.data(), d.y, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    CTX_TRY(c, hipStreamSynchronize(c->stream));
    for (int32_t k = 0; k < N; k++) {
      const int64_t o = c->perm_h[k];
      out[2 * o] = bx[k];
      out[2 * o + 1] = by[k];
    }
    return LPS_OK;
  };
  auto pull1 = [&](const double *d, double *out) -> int32_t {
    if (!out) return LPS_OK;
    CTX_TRY(c, hipMemcpyAsync(bx.data(), d, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    CTX_TRY(c, hipStreamSynchronize(c->stream));
    for (int32_t k = 0; k < N; k++) out[c->perm_h[k]] = bx[k];
    return LPS_OK;
  };
  int32_t rc;
  if ((rc = pull2(c->U[s % 3], u)) != LPS_OK) return rc;
  if ((rc = pull2(c->V[s % 3], v)) != LPS_OK) return rc;
  if ((rc = pull2(c->A[s % 2], a)) != LPS_OK) return rc;
  if ((rc = pull1(c->TH[s % 3], theta)) != LPS_OK) return rc;
  if ((rc = pull1(c->m, m)) != LPS_OK) return rc;
  return LPS_OK;
}

int64_t lps_horizons(lps_ctx *c, int64_t *row_ptr, int32_t *cols) {
  const int32_t N = c->n;
  if (!cols && !row_ptr) return c->hb.nnz;
  std::vector<int64_t> rp(N + 1);
  std::vector<int32_t> cl(std::max<int64_t>(c->hb.nnz, 1));
  CTX_TRY(c, hipMemcpyAsync(rp.data(), c->hb.row_ptr, sizeof(int64_t) * (N + 1), hipMemcpyDeviceToHost, c->stream));
  CTX_TRY(c, hipMemcpyAsync(cl.data(), c->hb.col, sizeof(int32_t) * c->hb.nnz, hipMemcpyDeviceToHost, c->stream));
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  std::vector<int32_t> iperm(N);
  for (int32_t k = 0; k < N; k++) iperm[c->perm_h[k]] = k;
  int64_t pos = 0;
  if (row_ptr) row_ptr[0] = 0;
  for (int32_t o = 0; o < N; o++) {
    const int32_t k = iperm[o];
    const int64_t len = rp[k + 1] - rp[k];
    if (cols)
      for (int64_t t = 0; t < len; t++) cols[pos + t] = c->perm_h[cl[rp[k] + t]];
    pos += len;
    if (row_ptr) row_ptr[o + 1] = pos;
  }
  return c->hb.nnz;
}

}  // extern "C"
