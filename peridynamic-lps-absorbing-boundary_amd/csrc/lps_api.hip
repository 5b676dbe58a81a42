// C ABI of the LPS-ABC time loop: context setup, step loop, output block, fetch.
//
// Setup follows src/main.cpp:178-684, the time loop :718-1164 (lps_kernels.hip has the passes).
// Buffer rotation (src/main.cpp:1145-1163) is index arithmetic on fixed buffers: at step s
// (s completed steps) U^n lives in U[s % 3], U^{n-1} in U[(s + 2) % 3] and U^{n+1} is written to
// U[(s + 1) % 3]; V likewise; A and the prediction alternate between two buffers; the dilatation
// history uses three (theta of step s is written to TH[(s + 1) % 3], ThetaN is TH[s % 3],
// ThetaNm1 is TH[(s + 2) % 3]). The pattern repeats every 6 steps, so a captured 30-step graph
// replays from any step that is a multiple of 30.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <unordered_set>
#include <vector>

#include "../../include/lps_amd.h"
#include "lps_ctx.h"
#include "lps_internal.h"

using namespace lps;

namespace lps {
thread_local std::string g_create_error;  // message of the last failed call without a context
}  // namespace lps

StepArgs lps_ctx::args() const {
  StepArgs a{};
  a.n = n_slots; a.n_abc = n_abc; a.M = M;
  a.dt = dt; a.rho = rho; a.lam = lam; a.mu = mu; a.kappa = kappa;
  a.delta = delta; a.dx = dx; a.dy = dy;
  a.x = x; a.y = y; a.kind = kind; a.abc_slot = abc_slot;
  a.row_ptr = hb.row_ptr; a.col = hb.col;
  a.sell_off = hb.sell_off; a.sell_w = hb.sell_w; a.sell_col = hb.sell_col;
  a.m = m; a.c2m = c2m; a.is_bc = is_bc;
  a.abc_off = abc_off; a.cloud = cloud; a.gu = gu; a.gv = gv; a.gth = gth;
  a.abc_gofs = abc_gofs; a.abc_xf = abc_xf; a.abc_pmap = abc_pmap;
  a.theta_method = opt.theta_method;
  a.inst = inst_d;
  a.tile = tile;
  return a;
}

StepState lps_ctx::state(int64_t s) const {
  StepState st{};
  const Dev2 &Pc = P[s % 6], &Pn = P[(s + 1) % 6];
  const Dev2 &Uc = U[s % 3], &Um = U[(s + 2) % 3], &Un = U[(s + 1) % 3];
  const Dev2 &Vc = V[s % 3], &Vm = V[(s + 2) % 3], &Vn = V[(s + 1) % 3];
  const Dev2 &Ac = A[s % 2], &An = A[(s + 1) % 2];
  st.px = Pc.x; st.py = Pc.y;
  st.ux = Uc.x; st.uy = Uc.y; st.uxm = Um.x; st.uym = Um.y;
  st.vx = Vc.x; st.vy = Vc.y; st.vxm = Vm.x; st.vym = Vm.y;
  st.ax = Ac.x; st.ay = Ac.y;
  st.theta = TH[(s + 1) % 3];
  st.thn = TH[s % 3];
  st.thnm1 = TH[(s + 2) % 3];
  st.ux1 = Un.x; st.uy1 = Un.y; st.vx1 = Vn.x; st.vy1 = Vn.y; st.ax1 = An.x; st.ay1 = An.y;
  st.pnx = Pn.x; st.pny = Pn.y;
  return st;
}

static double beta_host(double xi, double R, double dx) {  // src/main.cpp:100-107
  if (xi <= R - dx / 2.0) return 1.0;
  else if (xi <= R + dx / 2.0) return (R + dx / 2.0 - xi) / dx;
  else return 0.0;
}

// Lattice coordinates of every node, when all nodes sit exactly on x0 + ix dx, y0 + iy dy.
static bool lattice_coords(const lps_problem *p, std::vector<int32_t> &ix, std::vector<int32_t> &iy) {
  const int32_t N = p->node_no;
  double xmin = p->nodes[0], ymin = p->nodes[1];
  for (int32_t i = 1; i < N; i++) {
    xmin = std::min(xmin, p->nodes[2 * (int64_t)i]);
    ymin = std::min(ymin, p->nodes[2 * (int64_t)i + 1]);
  }
  ix.resize(N);
  iy.resize(N);
  for (int32_t i = 0; i < N; i++) {
    const double x = p->nodes[2 * (int64_t)i], y = p->nodes[2 * (int64_t)i + 1];
    const long long a = std::llround((x - xmin) / p->dx), b = std::llround((y - ymin) / p->dy);
    if (a < 0 || b < 0 || a > 32767 * kTile || b > 32767 * kTile) return false;
    if (std::abs(xmin + (double)a * p->dx - x) > 1e-9 * p->dx ||
        std::abs(ymin + (double)b * p->dy - y) > 1e-9 * p->dy)
      return false;
    ix[i] = (int32_t)a;
    iy[i] = (int32_t)b;
  }
  return true;
}

extern "C" {

static int32_t reset_state(lps_ctx *c);
static int32_t flush_events(lps_ctx *c);
static int32_t choose_variant(lps_ctx *c);
static int32_t choose_abc_predict(lps_ctx *c, bool tune);
static void drop_graphs(lps_ctx *c);

void lps_default_options(lps_options *opt) {
  opt->theta_method = LPS_THETA_GTHETA;
  opt->abc_enabled = 1;
  opt->exact = 0;
  opt->use_graph = 1;
  opt->device = 0;
  opt->horizon_radius = 0.0;
  opt->variant = 0;
  opt->concurrent = -1;
}

// concurrent = -1 (the default): schedule 2 (the partial tiles on a third stream, in whole-tile
// blocks) when the partial tiles are at least a tenth of the tiles, as in batches of small
// domains (default workload: LSQ step 3.70 -> 3.58 ms; 64 domains 1.55 -> 1.50 ms), else 1 (one
// large domain: even). Every schedule gives the same bits.
static void resolve_schedule(lps_ctx *c) {
  if (c->opt.concurrent >= 0) return;
  const bool many_part = c->fast && (int64_t)c->tile.n_part * 10 >= (int64_t)c->tile.n_tiles;
  c->opt.concurrent = many_part ? 2 : 1;
  c->path += many_part ? "; schedule 2 (partial tiles on a third stream)" : "; schedule 1";
}

const char *lps_global_error(void) { return g_create_error.c_str(); }
const char *lps_last_error(const lps_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }
const char *lps_describe(const lps_ctx *ctx) { return ctx->path.c_str(); }

int32_t lps_tile_info(const lps_ctx *ctx, int64_t *out5) {
  for (int q = 0; q < 5; q++) out5[q] = 0;
  if (!ctx->fast) return LPS_OK;
  out5[0] = ctx->tile.n_tiles;
  out5[1] = ctx->tile.n_full;
  out5[2] = ctx->tile.n_part;
  out5[3] = ctx->n_full_nodes;
  out5[4] = ctx->n;
  return LPS_OK;
}
int64_t lps_step_index(const lps_ctx *ctx) { return ctx->step; }
double lps_m_ref(const lps_ctx *ctx) { return ctx->m_ref; }

void lps_destroy(lps_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  closure_free(ctx);
  if (ctx->graph30) (void)hipGraphExecDestroy(ctx->graph30);
  for (hipGraphExec_t g : ctx->graph30t)
    if (g) (void)hipGraphExecDestroy(g);
  for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
  for (void *p : ctx->allocs) (void)hipFree(p);
  free_horizons(&ctx->hb);
  if (ctx->hout) (void)hipHostFree(ctx->hout);
  if (ctx->hout_inst) (void)hipHostFree(ctx->hout_inst);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  for (int q = 0; q < 2; q++) {
    if (ctx->ev_fork[q]) (void)hipEventDestroy(ctx->ev_fork[q]);
    if (ctx->ev_join[q]) (void)hipEventDestroy(ctx->ev_join[q]);
  }
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  if (ctx->stream3) (void)hipStreamSynchronize(ctx->stream3);
  for (hipEvent_t ev : ctx->ev_part)
    if (ev) (void)hipEventDestroy(ev);
  if (ctx->stream3) (void)hipStreamDestroy(ctx->stream3);
  if (ctx->stream4) (void)hipStreamSynchronize(ctx->stream4);
  if (ctx->ev_clo) (void)hipEventDestroy(ctx->ev_clo);
  if (ctx->stream4) (void)hipStreamDestroy(ctx->stream4);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

// FNV-1a (on 8-byte words) over the ABC inputs of a problem: the ABC node list, the clouds and the operators the
// time loop reads (G_theta only for the G_theta method).
static uint64_t abc_inputs_hash(const lps_problem *p, bool abc_on, int32_t theta_method) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](const void *ptr, size_t n) {  // 8-byte words, then the tail bytes
    const unsigned char *b = (const unsigned char *)ptr;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
      uint64_t w;
      memcpy(&w, b + i, 8);
      h = (h ^ w) * 1099511628211ull;
    }
    for (; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
  };
  if (!abc_on) return h;
  const int64_t C = p->constrains_no, M = p->max_cloud_size;
  mix(p->inf_nodes, sizeof(int32_t) * C);
  mix(p->iclouds, sizeof(int32_t) * C * M);
  mix(p->ci_nodes_u, sizeof(double) * 2 * C * 4 * M);
  mix(p->ci_nodes_v, sizeof(double) * 2 * C * 4 * M);
  if (theta_method == LPS_THETA_GTHETA) mix(p->ci_nodes_theta, sizeof(double) * C * 2 * M);
  return h;
}

// The fast path's packed member state (AbcPack) for the clouds cl (slots, -1 pads; row r's entries
// from off[r] on): the distinct member slots in ascending order, each entry's index among them in
// the order of its row's operator block at a fixed stride per row (entry jj of row r at
// r * stride + pmap[off[r] + jj], so the row kernels read a member's coefficients at its own
// position, without the map), and the record buffers.
static int32_t build_member_pack(lps_ctx *c, const std::vector<int32_t> &cl,
                                 const std::vector<int32_t> &off, const std::vector<uint8_t> &pmap,
                                 int64_t n_slots, bool with_theta) {
  std::vector<int32_t> of((size_t)std::max<int64_t>(n_slots, 1), -1), slots;
  for (int32_t v : cl)
    if (v >= 0) of[v] = 0;
  for (int64_t i = 0; i < n_slots; i++)
    if (of[i] == 0) {
      of[i] = (int32_t)slots.size();
      slots.push_back((int32_t)i);
    }
  const size_t R = off.size() - 1;
  int32_t W = 1;
  for (size_t r = 0; r < R; r++) W = std::max(W, off[r + 1] - off[r]);
  std::vector<int32_t> idx(std::max<size_t>(R * W, 1), -1);
  for (size_t r = 0; r < R; r++)
    for (int32_t e = off[r]; e < off[r + 1]; e++)
      idx[r * W + pmap[e]] = cl[e] < 0 ? -1 : of[cl[e]];
  c->n_pk = (int32_t)slots.size();
  c->pk_stride = W;
  CTX_TRY(c, c->upload(&c->pk_slot, slots));
  CTX_TRY(c, c->upload(&c->pk_idx, idx));
  // at least one record: the row kernels read record 0 for pads
  for (int b = 0; b < 2; b++) CTX_TRY(c, c->alloc(&c->pk_lvl[b], (size_t)std::max(c->n_pk, 1)));
  c->pk_last_step = -1;
  if (with_theta) CTX_TRY(c, c->alloc(&c->pk_th, (size_t)std::max(c->n_pk, 1)));
  CTX_TRY(c, hipStreamSynchronize(c->stream));  // the host vectors end here
  return LPS_OK;
}

// Lattice images among the ABC rows. On a square lattice about the origin each row (lattice
// position (i, j)) has a canonical row, the one at (max(|i|, |j|), min(|i|, |j|)) in the octant
// 0 <= y <= x, and a signed permutation T with (i, j) = T (c1, c2). The row is an image when its
// cloud is T applied to the canonical cloud (same member count) and every 2 x 2 block of its G_u
// and G_v is T B T' of the canonical member's block B, and its G_theta weights are the canonical
// ones: exactly, bit for bit. Then it needs no operators of its own (the kernels map the canonical
// block, lattice_map). True when every row is an image; canon[r] the canonical row, xf[r] the
// transform (bit 0 swap, bit 1 cross terms negated), pmap[off[r] + jj] the canonical member of
// member jj. Operators solved row by row (the reference's preprocessor) are images only up to
// rounding, so they keep a block per row; operators built once per canonical row and mapped by
// the same transforms are exact images.
static bool abc_lattice_images(const lps_problem *p, const std::vector<int32_t> &rows, int32_t M,
                               const std::vector<int32_t> &off, bool gth_on,
                               std::vector<int32_t> &canon, std::vector<int8_t> &xf,
                               std::vector<uint8_t> &pmap) {
  const int32_t R = (int32_t)rows.size();
  if (R == 0 || M > 256) return false;
  auto lat = [p](int32_t k, long long &i, long long &j) {
    const double x = p->nodes[2 * (int64_t)k], y = p->nodes[2 * (int64_t)k + 1];
    i = std::llround(x / p->dx);
    j = std::llround(y / p->dy);
    return (double)i * p->dx == x && (double)j * p->dy == y && std::llabs(i) < (1 << 30) &&
           std::llabs(j) < (1 << 30);
  };
  auto key = [](long long i, long long j) { return (int64_t)((uint64_t)(uint32_t)i << 32 | (uint32_t)j); };
  std::vector<std::pair<int64_t, int32_t>> at(R);  // lattice key -> row
  std::vector<long long> ri(R), rj(R);
  for (int32_t r = 0; r < R; r++) {
    if (!lat(p->inf_nodes[rows[r]], ri[r], rj[r])) return false;
    at[r] = {key(ri[r], rj[r]), r};
  }
  std::sort(at.begin(), at.end());
  auto row_at = [&at](int64_t k) {
    auto it = std::lower_bound(at.begin(), at.end(), std::make_pair(k, (int32_t)INT32_MIN));
    return it != at.end() && it->first == k ? it->second : -1;
  };
  auto map_block = [](int t, const double *b0, const double *b1, double *o0, double *o1) {
    if (t & 1) {
      const double x0[4] = {b1[1], b1[0], b1[3], b1[2]}, x1[4] = {b0[1], b0[0], b0[3], b0[2]};
      for (int q = 0; q < 4; q++) { o0[q] = x0[q]; o1[q] = x1[q]; }
    } else {
      for (int q = 0; q < 4; q++) { o0[q] = b0[q]; o1[q] = b1[q]; }
    }
    if (t & 2) { o0[1] = -o0[1]; o0[3] = -o0[3]; o1[0] = -o1[0]; o1[2] = -o1[2]; }
  };
  canon.assign(R, -1);
  std::vector<std::pair<int64_t, int32_t>> mk;  // member keys of a row -> member
  std::vector<char> hit;
  for (int32_t r = 0; r < R; r++) {
    const long long ia = ri[r], ja = rj[r];
    const long long c1 = std::max(std::llabs(ia), std::llabs(ja)), c2 = std::min(std::llabs(ia), std::llabs(ja));
    const int32_t cr = row_at(key(c1, c2));
    if (cr < 0) return false;
    const int s1 = ia < 0 ? -1 : 1, s2 = ja < 0 ? -1 : 1;
    const bool sw = std::llabs(ia) < std::llabs(ja);
    const int t = (sw ? 1 : 0) | (s1 * s2 < 0 ? 2 : 0);
    const int32_t L = off[r + 1] - off[r];
    if (off[cr + 1] - off[cr] != L) return false;
    const int64_t a = rows[r], ac = rows[cr];
    mk.clear();
    for (int32_t jj = 0; jj < L; jj++) {
      const int32_t k = p->iclouds[a * M + jj];
      long long x, y;
      if (k < 0 || k >= p->node_no || !lat(k, x, y)) return false;
      mk.push_back({key(x, y), jj});
    }
    std::sort(mk.begin(), mk.end());
    hit.assign(L, 0);
    for (int32_t m = 0; m < L; m++) {
      const int32_t kc = p->iclouds[ac * M + m];
      long long x, y;
      if (kc < 0 || kc >= p->node_no || !lat(kc, x, y)) return false;
      const long long ox = x - c1, oy = y - c2;
      const long long qx = ia + (sw ? s1 * oy : s1 * ox), qy = ja + (sw ? s2 * ox : s2 * oy);
      auto it = std::lower_bound(mk.begin(), mk.end(), std::make_pair(key(qx, qy), (int32_t)INT32_MIN));
      if (it == mk.end() || it->first != key(qx, qy) || hit[it->second]) return false;
      const int32_t jj = it->second;
      hit[jj] = 1;
      pmap[(size_t)off[r] + jj] = (uint8_t)m;
      // the values, exactly
      for (const double *G : {p->ci_nodes_u, p->ci_nodes_v}) {
        double o0[4], o1[4];
        map_block(t, G + ac * 8 * M + 4 * m, G + ac * 8 * M + 4 * M + 4 * m, o0, o1);
        const double *g0 = G + a * 8 * M + 4 * jj, *g1 = G + a * 8 * M + 4 * M + 4 * jj;
        for (int q = 0; q < 4; q++)
          if (!(o0[q] == g0[q]) || !(o1[q] == g1[q])) return false;
      }
      if (gth_on && (!(p->ci_nodes_theta[ac * 2 * M + 2 * m] == p->ci_nodes_theta[a * 2 * M + 2 * jj]) ||
                     !(p->ci_nodes_theta[ac * 2 * M + 2 * m + 1] == p->ci_nodes_theta[a * 2 * M + 2 * jj + 1])))
        return false;
    }
    canon[r] = cr;
    xf[r] = (int8_t)t;
  }
  return true;
}

static int32_t create_impl(lps_ctx *c, const lps_problem *p) {
  const int32_t N = p->node_no, C = p->constrains_no, M = p->max_cloud_size;
  if (N <= 0 || !p->nodes || !p->bctypes || !p->icvalues || !p->icdvalues)
    return c->fail(LPS_ERR_ARG, "node_no must be positive and nodes/bctypes/icvalues/icdvalues set");
  if (C < 0 || (C > 0 && M <= 0)) return c->fail(LPS_ERR_ARG, "inconsistent constrains_no / max_cloud_size");
  if (C > 0 && M > 256) return c->fail(LPS_ERR_ARG, "max_cloud_size above 256");
  const bool abc_on = c->opt.abc_enabled && C > 0;
  if (abc_on && (!p->inf_nodes || !p->iclouds || !p->ci_nodes_u || !p->ci_nodes_v))
    return c->fail(LPS_ERR_ARG, "ABC rows need inf_nodes, iclouds, ci_nodes_u and ci_nodes_v");
  if (abc_on && c->opt.theta_method == LPS_THETA_GTHETA && !p->ci_nodes_theta)
    return c->fail(LPS_ERR_ARG, "the G_theta method needs ci_nodes_theta");
  if (c->opt.theta_method != LPS_THETA_GTHETA && c->opt.theta_method != LPS_THETA_IDW)
    return c->fail(LPS_ERR_ARG, "theta_method must be LPS_THETA_GTHETA or LPS_THETA_IDW");
  if (!(p->dx > 0.0) || !(p->dy > 0.0)) return c->fail(LPS_ERR_ARG, "dx and dy must be positive");
  if (p->prenotch_no < 0 || (p->prenotch_no > 0 && !p->prenotches))
    return c->fail(LPS_ERR_ARG, "prenotch_no must be >= 0, with prenotches set when positive");
  c->n = N;
  c->M = abc_on ? M : 0;
  c->abc_hash.assign(1, abc_inputs_hash(p, abc_on, c->opt.theta_method));
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
  const hipStream_t s = c->stream;
  const double search_r = c->opt.horizon_radius > 0.0 ? c->opt.horizon_radius
                                                      : c->delta + c->dx / 2.0 + 1e-5;  // src/main.cpp:352

  // ---- slot layout ----
  std::vector<int32_t> lix, liy;
  const int32_t reach = (int32_t)std::floor(search_r / std::min(c->dx, c->dy));
  bool tile_layout = !c->opt.exact && reach <= kMaxHalo && lattice_coords(p, lix, liy);
  std::vector<int32_t> slot_of(N);
  std::vector<int32_t> tile_xy, tile_map;
  if (tile_layout) {
    int32_t nx = 0, ny = 0;
    for (int32_t i = 0; i < N; i++) { nx = std::max(nx, lix[i] + 1); ny = std::max(ny, liy[i] + 1); }
    const int32_t ntx = (nx + kTile - 1) / kTile, nty = (ny + kTile - 1) / kTile;
    tile_map.assign((size_t)ntx * nty, -1);
    for (int32_t i = 0; i < N; i++) tile_map[(size_t)(lix[i] / kTile) * nty + liy[i] / kTile] = 0;
    for (int32_t tx = 0; tx < ntx; tx++)
      for (int32_t ty = 0; ty < nty; ty++)
        if (tile_map[(size_t)tx * nty + ty] == 0) {
          tile_map[(size_t)tx * nty + ty] = (int32_t)tile_xy.size();
          tile_xy.push_back((tx << 16) | ty);
        }
    for (int32_t i = 0; i < N; i++) {
      const int32_t t = tile_map[(size_t)(lix[i] / kTile) * nty + liy[i] / kTile];
      slot_of[i] = t * kTile * kTile + (lix[i] % kTile) * kTile + (liy[i] % kTile);
    }
    c->n_slots = (int32_t)tile_xy.size() * kTile * kTile;
    c->tile.n_tiles = (int32_t)tile_xy.size();
    c->tile.ntx = ntx;
    c->tile.nty = nty;
    c->tile.halo = reach;
    c->tile.win = kTile + 2 * reach;
  } else {
    std::vector<double> hx(N), hy(N);
    for (int32_t i = 0; i < N; i++) {
      hx[i] = p->nodes[2 * (int64_t)i];
      hy[i] = p->nodes[2 * (int64_t)i + 1];
    }
    double *xo, *yo;
    int8_t *dseg;
    int32_t *drank, *dperm;
    CTX_TRY(c, c->upload(&xo, hx));
    CTX_TRY(c, c->upload(&yo, hy));
    CTX_TRY(c, c->upload(&dseg, seg));
    CTX_TRY(c, c->upload(&drank, seg_rank));
    CTX_TRY(c, c->alloc(&dperm, N));
    CTX_TRY(c, morton_order(N, xo, yo, dseg, drank, search_r, dperm, s));
    std::vector<int32_t> perm(N);
    CTX_TRY(c, hipMemcpyAsync(perm.data(), dperm, sizeof(int32_t) * N, hipMemcpyDeviceToHost, s));
    CTX_TRY(c, hipStreamSynchronize(s));
    for (int32_t k = 0; k < N; k++) slot_of[perm[k]] = k;
    c->n_slots = N;
  }
  const int32_t NS = c->n_slots;
  c->perm_h.assign(NS, -1);
  std::vector<int8_t> kind(NS, 0), isbc(NS, 0);
  std::vector<double> xs(NS, 0.0), ys(NS, 0.0);
  for (int32_t i = 0; i < N; i++) {
    const int32_t k = slot_of[i];
    c->perm_h[k] = i;
    kind[k] = seg[i] ? 2 : 1;
    isbc[k] = p->bctypes[2 * (int64_t)i] == 1;
    xs[k] = p->nodes[2 * (int64_t)i];
    ys[k] = p->nodes[2 * (int64_t)i + 1];
  }
  std::vector<int32_t> abc_slot(c->n_abc);
  for (int32_t r = 0; r < c->n_abc; r++) abc_slot[r] = slot_of[p->inf_nodes[active_rows[r]]];
  CTX_TRY(c, c->upload(&c->x, xs));
  CTX_TRY(c, c->upload(&c->y, ys));
  CTX_TRY(c, c->upload(&c->kind, kind));
  CTX_TRY(c, c->upload(&c->is_bc, isbc));
  CTX_TRY(c, c->upload(&c->perm_d, c->perm_h));
  CTX_TRY(c, c->upload(&c->abc_slot, abc_slot));
  CTX_TRY(c, hipStreamSynchronize(s));

  // ---- horizons ----
  bool overflow = false;
  double *d_notch = nullptr;
  if (p->prenotch_no > 0)
    CTX_TRY(c, c->upload(&d_notch, std::vector<double>(p->prenotches,
                                                       p->prenotches + 4 * (size_t)p->prenotch_no)));
  CTX_TRY(c, build_horizons(NS, c->x, c->y, c->kind, c->perm_d, search_r, kMaxHorizon, d_notch,
                            p->prenotch_no, &c->hb, &overflow, s));
  if (overflow) return c->fail(LPS_ERR_HORIZON, "found_node_no > max_node_no; increase max_node_no.");

  // ---- m_i on the device, m_ref and the ABC override on the host (src/main.cpp:408-480) ----
  CTX_TRY(c, c->alloc(&c->m, NS));
  {
    StepArgs a = c->args();
    CTX_TRY(c, launch_weighted_volume(a, c->m, s));
  }
  std::vector<double> hm(NS);
  CTX_TRY(c, hipMemcpyAsync(hm.data(), c->m, sizeof(double) * NS, hipMemcpyDeviceToHost, s));
  CTX_TRY(c, hipStreamSynchronize(s));
  double m_max = 0.0;
  for (int32_t k = 0; k < NS; k++)
    if (kind[k] && hm[k] > 0.0) m_max = std::max(m_max, hm[k]);
  c->m_ref = m_max;
  if (abc_on)
    for (int32_t a = 0; a < C; a++) hm[slot_of[p->inf_nodes[a]]] = c->m_ref;
  CTX_TRY(c, hipMemcpyAsync(c->m, hm.data(), sizeof(double) * NS, hipMemcpyHostToDevice, s));
  std::vector<double> hc(NS);
  for (int32_t k = 0; k < NS; k++) hc[k] = hm[k] > 0.0 ? 2.0 / hm[k] : 0.0;
  CTX_TRY(c, c->upload(&c->c2m, hc));
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
  // the problem's instance entry (the fast path's kernels read the material through it)
  InstParams ip{};
  ip.dt = c->dt; ip.rho = c->rho; ip.lam = c->lam; ip.mu = c->mu; ip.kappa = c->kappa;
  ip.exclude_r = c->exclude_r;
  ip.th_scale = (c->lam - c->mu) / c->mu;
  ip.f_scale = c->mu;

  // ---- tile path: the stencil of a full horizon and the check against the CSR ----
  c->fast = false;
  c->path = tile_layout ? "exact kernels on the tile layout" : "exact kernels, Morton layout";
  if (tile_layout) {
    struct Off { double d; int32_t ox, oy; };
    std::vector<Off> offs;
    const int32_t K = c->tile.halo;
    for (int32_t ox = -K; ox <= K; ox++)
      for (int32_t oy = -K; oy <= K; oy++) {
        if (ox == 0 && oy == 0) continue;
        const double xx = ox * c->dx, yy = oy * c->dy;
        const double d = std::sqrt(xx * xx + yy * yy);
        if (d <= search_r) offs.push_back({d, ox, oy});
      }
    std::sort(offs.begin(), offs.end(), [](const Off &a, const Off &b) {
      if (a.d != b.d) return a.d < b.d;
      if (a.ox != b.ox) return a.ox < b.ox;
      return a.oy < b.oy;
    });
    const int32_t nt = (int32_t)offs.size();
    std::vector<int32_t> dl(nt);
    std::vector<double> thx(nt), thy(nt), tg1(nt), tg2(nt), txi(nt);
    for (int32_t k = 0; k < nt; k++) {
      const double xx = offs[k].ox * c->dx, yy = offs[k].oy * c->dy, d = offs[k].d;
      const double bt = beta_host(d, c->delta, c->dx);
      dl[k] = offs[k].ox * c->tile.win + offs[k].oy;
      thx[k] = xx / d;
      thy[k] = yy / d;
      tg1[k] = d * bt * c->dx * c->dy;
      tg2[k] = bt * c->dx * c->dy;
      txi[k] = d;
    }
    // 3 x 3 neighbour tiles and "full" windows (every cell a node with m == m_ref)
    const int32_t ntl = c->tile.n_tiles;
    std::vector<int32_t> nbr((size_t)ntl * 9, -1);
    std::vector<int8_t> full(ntl, 0);
    for (int32_t t = 0; t < ntl; t++) {
      const int32_t tx = tile_xy[t] >> 16, ty = tile_xy[t] & 0xffff;
      for (int32_t ax = -1; ax <= 1; ax++)
        for (int32_t ay = -1; ay <= 1; ay++) {
          const int32_t qx = tx + ax, qy = ty + ay;
          if (qx < 0 || qy < 0 || qx >= c->tile.ntx || qy >= c->tile.nty) continue;
          nbr[(size_t)t * 9 + (ax + 1) * 3 + (ay + 1)] = tile_map[(size_t)qx * c->tile.nty + qy];
        }
      bool f = true;
      const int32_t W = c->tile.win;
      for (int32_t cc = 0; cc < W * W && f; cc++) {
        const int32_t rx = cc / W - K, ry = cc % W - K;
        const int32_t ax = rx < 0 ? 0 : (rx < kTile ? 1 : 2), ay = ry < 0 ? 0 : (ry < kTile ? 1 : 2);
        const int32_t tt = nbr[(size_t)t * 9 + ax * 3 + ay];
        if (tt < 0) { f = false; break; }
        const int32_t sl = tt * kTile * kTile + (rx - (ax - 1) * kTile) * kTile + (ry - (ay - 1) * kTile);
        f = kind[sl] != 0 && hm[sl] == c->m_ref;
      }
      full[t] = f;
    }
    // the reference stencil (dx == dy, the 68 offsets 0 < ox^2 + oy^2 <= 20) runs unrolled
    bool s68 = (K == 4) && (c->dx == c->dy) && (nt == 68);
    for (int32_t k = 0; k < nt && s68; k++) {
      const int32_t r2 = offs[k].ox * offs[k].ox + offs[k].oy * offs[k].oy;
      s68 = r2 > 0 && r2 <= 20;
    }
    c->stencil68_ok = s68;
    c->tile.theta_kernel = c->tile.force_kernel = kGeneric;
    // nodes per lane of the register-blocked kernels: the force pass two wavefronts per tile (4%
    // faster on R = 500 mm, same-box A/B), the dilatation one (no faster with two);
    // LPS_RB_NODES=<dilatation>,<force> (4 or 2 each) overrides
    c->tile.theta_nr = 4;
    c->tile.force_nr = 2;
    if (const char *e = std::getenv("LPS_RB_NODES")) {
      int th = 4, fo = 2;
      if (std::sscanf(e, "%d,%d", &th, &fo) == 2) {
        c->tile.theta_nr = th == 2 ? 2 : 4;
        c->tile.force_nr = fo == 2 ? 2 : 4;
      }
    }
    if (s68) {
      const double C = 2.0 / c->m_ref;
      for (int32_t q = 0; q < 12; q++) {
        const double xi = c->dx * std::sqrt((double)kStencilR2[q]);
        const double bt = beta_host(xi, c->delta, c->dx);
        const double cc = c->dx / xi;
        c->tile.c1[q] = xi * bt * c->dx * c->dy * cc;
        c->tile.c2[q] = bt * c->dx * c->dy * cc * cc;
        int32_t ra = 0, rb = 0;  // class representative a >= b >= 0
        for (int32_t aa = 0; aa <= 4; aa++)
          for (int32_t bb = 0; bb <= aa; bb++)
            if (aa * aa + bb * bb == kStencilR2[q]) { ra = aa; rb = bb; }
        const double ms[2] = {(double)ra, (double)rb};
        const double ps[3] = {(double)(ra * ra), (double)(ra * rb), (double)(rb * rb)};
        for (int32_t z = 0; z < 2; z++) c->tile.wt[q][z] = C * c->tile.c1[q] * ms[z];
        for (int32_t z = 0; z < 3; z++) c->tile.wb[q][z] = 8.0 * C * c->tile.c2[q] * ps[z];
      }
      // sums of the bond-extension weights of o_x^2 and of o_y^2 over the 68 offsets
      c->tile.wbsum[0] = c->tile.wbsum[1] = 0.0;
      for (int32_t ox = -4; ox <= 4; ox++)
        for (int32_t oy = -4; oy <= 4; oy++) {
          const int32_t r2 = ox * ox + oy * oy;
          if (r2 == 0 || r2 > 20) continue;
          int32_t q = 0;
          while (kStencilR2[q] != r2) q++;
          int32_t ra = 0, rb = 0;
          for (int32_t aa = 0; aa <= 4; aa++)
            for (int32_t bb = 0; bb <= aa; bb++)
              if (aa * aa + bb * bb == r2) { ra = aa; rb = bb; }
          auto slot = [&](int32_t p2) { return p2 == ra * ra ? 0 : (p2 == ra * rb ? 1 : 2); };
          c->tile.wbsum[0] += c->tile.wb[q][slot(ox * ox)];
          c->tile.wbsum[1] += c->tile.wb[q][slot(oy * oy)];
        }
    }
    std::vector<int32_t> tinst(c->tile.n_tiles, 0);
    int32_t *d_tinst;
    CTX_TRY(c, c->upload(&d_tinst, tinst));
    c->tile.tile_inst = d_tinst;
    int32_t *d_txy, *d_tmap, *d_delta, *d_nbr;
    int8_t *d_full;
    double *d_hx, *d_hy, *d_g1, *d_g2, *d_xi;
    CTX_TRY(c, c->upload(&d_txy, tile_xy));
    CTX_TRY(c, c->upload(&d_tmap, tile_map));
    CTX_TRY(c, c->upload(&d_nbr, nbr));
    CTX_TRY(c, c->upload(&d_full, full));
    c->tile.tile_nbr = d_nbr;
    c->tile.tile_full = d_full;
    std::vector<int32_t> full_list, part_list;
    for (int32_t t = 0; t < ntl; t++) (full[t] ? full_list : part_list).push_back(t);
    int32_t *d_fl, *d_pl;
    CTX_TRY(c, c->upload(&d_fl, full_list));
    CTX_TRY(c, c->upload(&d_pl, part_list));
    c->tile.full_list = d_fl;
    c->tile.part_list = d_pl;
    c->tile.n_full = (int32_t)full_list.size();
    c->tile.n_part = (int32_t)part_list.size();
    // deep full tiles: the 3 x 3 tiles around them full and every cell of the 24 x 24 window an
    // interior node (kind 1); only with the 68-offset stencil (halo 4). The lists stay alive until
    // the stream synchronisation below (asynchronous uploads).
    std::vector<int32_t> deep_list, shallow_list;
    {
      for (int32_t t : full_list) {
        bool deep = K == 4;
        for (int32_t k = 0; k < 9 && deep; k++) {
          const int32_t tt = nbr[(size_t)t * 9 + k];
          deep = tt >= 0 && full[tt];
        }
        for (int32_t cc = 0; cc < (kTile + 8) * (kTile + 8) && deep; cc++) {
          const int32_t rx = cc / (kTile + 8) - 4, ry = cc % (kTile + 8) - 4;
          const int32_t ax = rx < 0 ? 0 : (rx < kTile ? 1 : 2), ay = ry < 0 ? 0 : (ry < kTile ? 1 : 2);
          const int32_t tt = nbr[(size_t)t * 9 + ax * 3 + ay];
          deep = kind[(size_t)tt * kTile * kTile + (rx - (ax - 1) * kTile) * kTile + (ry - (ay - 1) * kTile)] == 1;
        }
        (deep ? deep_list : shallow_list).push_back(t);
      }
      CTX_TRY(c, c->upload(&c->deep_list, deep_list));
      CTX_TRY(c, c->upload(&c->shallow_list, shallow_list));
      c->n_deep = (int32_t)deep_list.size();
      c->n_shallow = (int32_t)shallow_list.size();
    }
    c->abc_in_part = true;  // every ABC node in a partial tile (its horizon reaches past the rim)
    for (int32_t r = 0; r < c->n_abc; r++)
      if (full[abc_slot[r] / (kTile * kTile)]) c->abc_in_part = false;
    CTX_TRY(c, c->upload(&d_delta, dl));
    CTX_TRY(c, c->upload(&d_hx, thx));
    CTX_TRY(c, c->upload(&d_hy, thy));
    CTX_TRY(c, c->upload(&d_g1, tg1));
    CTX_TRY(c, c->upload(&d_g2, tg2));
    CTX_TRY(c, c->upload(&d_xi, txi));
    c->tile.xi = d_xi;
    c->tile.tile_xy = d_txy;
    c->tile.tile_map = d_tmap;
    c->tile.nt = nt;
    c->tile.delta = d_delta;
    c->tile.hx = d_hx;
    c->tile.hy = d_hy;
    c->tile.g1 = d_g1;
    c->tile.g2 = d_g2;
    int32_t *d_mis;
    CTX_TRY(c, c->alloc(&d_mis, 1));
    CTX_TRY(c, hipMemsetAsync(d_mis, 0, sizeof(int32_t), s));
    const int32_t E = 2 * K + 1;
    std::vector<int16_t> bond_of((size_t)E * E, -1);
    for (int32_t k = 0; k < nt; k++) bond_of[(size_t)(offs[k].ox + K) * E + offs[k].oy + K] = (int16_t)k;
    int16_t *d_bond_of;
    CTX_TRY(c, c->upload(&d_bond_of, bond_of));
    CTX_TRY(c, launch_tile_check(c->args(), d_bond_of, d_mis, s));
    int32_t mis = 0;
    CTX_TRY(c, hipMemcpyAsync(&mis, d_mis, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CTX_TRY(c, hipStreamSynchronize(s));
    if (mis == 0) {
      c->fast = true;
      int32_t nfull = 0;
      for (int8_t f : full) nfull += f;
      for (int32_t t = 0; t < ntl; t++)
        if (full[t])
          for (int32_t k = 0; k < kTile * kTile; k++) c->n_full_nodes += kind[(size_t)t * kTile * kTile + k] != 0;
      c->path = " (16x16 tiles, " + std::to_string(ntl) + " tiles, " + std::to_string(nfull) +
                " full, " + std::to_string(c->n_deep) + " deep, " + std::to_string(nt) +
                " bonds, halo " + std::to_string(K) + ")";
    } else {
      c->path += " (" + std::to_string(mis) + " horizons differ from the stencil)";
    }
  }

  // ---- ABC rows on the device, active rows only, cloud ids as slots ----
  // Ragged rows. On the tile path a row ends after its last member that adds anything: trailing
  // members that are pads (-1) or have all-zero G_u, G_v (and G_theta) coefficients are dropped.
  // Their terms are zeros at the end of the row, so the wavefront sums keep their bits. The exact
  // path keeps every member, for the reference's order of additions.
  if (c->n_abc > 0) {
    const int32_t R = c->n_abc;
    const bool gth_on = c->opt.theta_method == LPS_THETA_GTHETA;
    std::vector<int32_t> off(R + 1, 0);
    for (int32_t r = 0; r < R; r++) {
      const int64_t a = active_rows[r];
      int32_t L = (int32_t)M;
      if (c->fast) {
        auto adds = [&](int32_t j) {
          if (p->iclouds[a * M + j] < 0) return false;
          for (int b = 0; b < 2; b++)
            for (int q = 0; q < 4; q++)
              if (p->ci_nodes_u[a * 8 * M + b * 4 * M + 4 * j + q] != 0.0 ||
                  p->ci_nodes_v[a * 8 * M + b * 4 * M + 4 * j + q] != 0.0)
                return true;
          return gth_on && (p->ci_nodes_theta[a * 2 * M + 2 * j] != 0.0 ||
                            p->ci_nodes_theta[a * 2 * M + 2 * j + 1] != 0.0);
        };
        while (L > 0 && !adds(L - 1)) L--;
      }
      if ((int64_t)off[r] + L > INT32_MAX) return c->fail(LPS_ERR_ARG, "too many ABC cloud members");
      off[r + 1] = off[r] + L;
    }
    const size_t NM = (size_t)off[R];
    c->abc_members = (int64_t)NM;
    // operator blocks: shared by lattice images (tile path), else one per row
    std::vector<int32_t> canon;
    std::vector<int8_t> xf(R, 0);
    std::vector<uint8_t> pmap(std::max<size_t>(NM, 1), 0);
    const char *sym_env = std::getenv("LPS_ABC_SYMMETRY");
    const bool try_sym = c->fast && !(sym_env && !strcmp(sym_env, "off"));
    bool shared = try_sym && abc_lattice_images(p, active_rows, M, off, gth_on, canon, xf, pmap);
    if (!shared) {
      canon.resize(R);
      for (int32_t r = 0; r < R; r++) {
        canon[r] = r;
        xf[r] = 0;
        for (int32_t j = 0; j < off[r + 1] - off[r]; j++) pmap[(size_t)off[r] + j] = (uint8_t)j;
      }
    }
    std::vector<int32_t> gofs(R), bofs(R, -1);
    int64_t GM = 0;  // members of the operator blocks
    for (int32_t r = 0; r < R; r++)
      if (canon[r] == r) {
        bofs[r] = (int32_t)GM;
        GM += off[r + 1] - off[r];
      }
    for (int32_t r = 0; r < R; r++) gofs[r] = bofs[canon[r]];
    c->abc_block_members = GM;
    c->abc_shared_blocks = shared;
    std::vector<int32_t> hcl(std::max<size_t>(NM, 1), -1);
    // one member of zeros past the last block: a row without members (L = 0) still reads its
    // first member's coefficients (discarded) without a branch on L in the row kernels
    std::vector<double> hgu(8 * GM + 8), hgv(8 * GM + 8), hgt;
    if (gth_on) hgt.resize(2 * GM + 2);
    for (int32_t r = 0; r < R; r++) {
      const int64_t a = active_rows[r];
      const size_t o = (size_t)off[r];
      const int32_t L = off[r + 1] - off[r];
      for (int32_t j = 0; j < L; j++) {
        const int32_t k = p->iclouds[a * M + j];
        hcl[o + j] = (k < 0) ? -1 : slot_of[k];
      }
      if (canon[r] != r) continue;
      const size_t g = (size_t)bofs[r];
      for (int b = 0; b < 2; b++) {  // the x block, then the y block, 4 L doubles each
        std::memcpy(&hgu[8 * g + 4 * b * L], p->ci_nodes_u + a * 8 * M + 4 * b * M, sizeof(double) * 4 * L);
        std::memcpy(&hgv[8 * g + 4 * b * L], p->ci_nodes_v + a * 8 * M + 4 * b * M, sizeof(double) * 4 * L);
      }
      if (gth_on) std::memcpy(&hgt[2 * g], p->ci_nodes_theta + a * 2 * M, sizeof(double) * 2 * L);
    }
    CTX_TRY(c, c->upload(&c->abc_gofs, gofs));
    CTX_TRY(c, c->upload(&c->abc_xf, xf));
    CTX_TRY(c, c->upload(&c->abc_pmap, pmap));
    if (shared)
      c->path += "; ABC operators of " + std::to_string(GM) + " of " + std::to_string(NM) +
                 " members stored, the rest lattice images";
    CTX_TRY(c, c->upload(&c->abc_off, off));
    CTX_TRY(c, c->upload(&c->cloud, hcl));
    if (c->fast) {
      const int32_t rc = build_member_pack(c, hcl, off, pmap, c->n_slots, gth_on);
      if (rc != LPS_OK) return rc;
    }
    CTX_TRY(c, c->upload(&c->gu, hgu));
    CTX_TRY(c, c->upload(&c->gv, hgv));
    if (!hgt.empty()) CTX_TRY(c, c->upload(&c->gth, hgt));
    CTX_TRY(c, hipStreamSynchronize(s));
  }

  // ---- state buffers; the initial conditions stay on the device for lps_reset ----
  for (int b = 0; b < 3; b++) {
    CTX_TRY(c, c->alloc(&c->U[b].x, NS)); CTX_TRY(c, c->alloc(&c->U[b].y, NS));
    CTX_TRY(c, c->alloc(&c->V[b].x, NS)); CTX_TRY(c, c->alloc(&c->V[b].y, NS));
    CTX_TRY(c, c->alloc(&c->TH[b], NS));
  }
  for (int b = 0; b < 2; b++) {
    CTX_TRY(c, c->alloc(&c->A[b].x, NS)); CTX_TRY(c, c->alloc(&c->A[b].y, NS));
  }
  for (int b = 0; b < 6; b++) {
    CTX_TRY(c, c->alloc(&c->P[b].x, NS)); CTX_TRY(c, c->alloc(&c->P[b].y, NS));
  }
  {
    std::vector<double> ux(NS, 0.0), uy(NS, 0.0), vx(NS, 0.0), vy(NS, 0.0);
    for (int32_t i = 0; i < N; i++) {
      const int32_t k = slot_of[i];
      ux[k] = p->icvalues[2 * (int64_t)i]; uy[k] = p->icvalues[2 * (int64_t)i + 1];
      vx[k] = p->icdvalues[2 * (int64_t)i]; vy[k] = p->icdvalues[2 * (int64_t)i + 1];
    }
    CTX_TRY(c, c->upload(&c->IC.x, ux));
    CTX_TRY(c, c->upload(&c->IC.y, uy));
    CTX_TRY(c, c->upload(&c->ICD.x, vx));
    CTX_TRY(c, c->upload(&c->ICD.y, vy));
    CTX_TRY(c, hipStreamSynchronize(s));
  }
  c->inst_h.assign(1, ip);
  CTX_TRY(c, c->upload(&c->inst_d, c->inst_h));
  // output block buffers
  c->n_partials = diag_partials(NS);
  CTX_TRY(c, c->alloc(&c->partials, std::max(c->n_partials, c->tile.n_tiles)));
  CTX_TRY(c, c->alloc(&c->dout, 1));
  CTX_TRY(c, hipHostMalloc((void **)&c->hout, sizeof(DiagPartial)));
  return reset_state(c);
}

// The prediction ring at step 0 (P[0], the prediction of step 1, written by the caller): U^0 = IC
// in P[5], U^{-1} = U^{-2} = 0 in P[4], P[3].
static hipError_t reset_ring(lps_ctx *c) {
  const size_t bytes = sizeof(double) * c->n_slots;
  hipError_t e = hipSuccess;
  for (int b = 1; b < 5 && e == hipSuccess; b++) {
    e = hipMemsetAsync(c->P[b].x, 0, bytes, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->P[b].y, 0, bytes, c->stream);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(c->P[5].x, c->IC.x, bytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(c->P[5].y, c->IC.y, bytes, hipMemcpyDeviceToDevice, c->stream);
  c->u_synced = 3;
  return e;
}

// Initial state (src/main.cpp:624-634), theta history (:664-682) and the prediction of step 1.
static int32_t reset_state(lps_ctx *c) {
  const int32_t NS = c->n_slots;
  const hipStream_t s = c->stream;
  const size_t bytes = sizeof(double) * NS;
  if (c->batch) {  // the instances' setup state, kept from lps_create_batch
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
    CTX_TRY(c, hipMemcpyAsync(c->TH[0], c->TH0, bytes, hipMemcpyDeviceToDevice, s));
    CTX_TRY(c, hipMemcpyAsync(c->P[0].x, c->P0.x, bytes, hipMemcpyDeviceToDevice, s));
    CTX_TRY(c, hipMemcpyAsync(c->P[0].y, c->P0.y, bytes, hipMemcpyDeviceToDevice, s));
    CTX_TRY(c, reset_ring(c));
    CTX_TRY(c, hipStreamSynchronize(s));
    c->step = 0;
    c->pk_last_step = -1;
    c->E0 = -1.0;
    c->E0_inst.assign(c->n_inst, -1.0);
    return LPS_OK;
  }
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
  CTX_TRY(c, reset_ring(c));
  CTX_TRY(c, hipStreamSynchronize(s));
  c->step = 0;
  c->pk_last_step = -1;
  c->E0 = -1.0;
  return LPS_OK;
}

int32_t lps_reset(lps_ctx *c) {
  int32_t rc = flush_events(c);
  if (rc != LPS_OK) return rc;
  rc = reset_state(c);
  if (rc == LPS_OK && c->clo) {
    CTX_TRY(c, closure_reset_hook(c));
    CTX_TRY(c, hipStreamSynchronize(c->stream));
  }
  return rc;
}

int32_t lps_set_initial(lps_ctx *c, const double *ic, const double *icd) {
  if (c->batch) return c->fail(LPS_ERR_ARG, "lps_set_initial takes a single-problem context");
  if (!ic || !icd) return c->fail(LPS_ERR_ARG, "ic and icd are required");
  const int32_t NS = c->n_slots;
  std::vector<double> ux(NS, 0.0), uy(NS, 0.0), vx(NS, 0.0), vy(NS, 0.0);
  for (int32_t k = 0; k < NS; k++) {
    const int32_t i = c->perm_h[k];
    if (i < 0) continue;
    ux[k] = ic[2 * (int64_t)i]; uy[k] = ic[2 * (int64_t)i + 1];
    vx[k] = icd[2 * (int64_t)i]; vy[k] = icd[2 * (int64_t)i + 1];
  }
  const size_t bytes = sizeof(double) * NS;
  CTX_TRY(c, hipMemcpyAsync(c->IC.x, ux.data(), bytes, hipMemcpyHostToDevice, c->stream));
  CTX_TRY(c, hipMemcpyAsync(c->IC.y, uy.data(), bytes, hipMemcpyHostToDevice, c->stream));
  CTX_TRY(c, hipMemcpyAsync(c->ICD.x, vx.data(), bytes, hipMemcpyHostToDevice, c->stream));
  CTX_TRY(c, hipMemcpyAsync(c->ICD.y, vy.data(), bytes, hipMemcpyHostToDevice, c->stream));
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  return lps_reset(c);
}

// A context with its device, streams and events, nothing else.
static int32_t new_context(const lps_options *options, lps_ctx **out) {
  lps_ctx *c = new lps_ctx();
  if (options) c->opt = *options;
  else lps_default_options(&c->opt);
  if (c->opt.concurrent < -1 || c->opt.concurrent > 4) {
    g_create_error = "concurrent must be -1 (automatic), 0, 1, 2, 3 or 4";
    delete c;
    return LPS_ERR_ARG;
  }
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
  // both streams at the default priority
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream4, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_clo, hipEventDisableTiming);
  if (const char *cf = std::getenv("LPS_CLOSURE_FIRST")) c->clo_first = std::atoi(cf) != 0;
  for (int q = 0; q < 3; q++)
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_part[q], hipEventDisableTiming);
  for (int q = 0; q < 2; q++) {
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork[q], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join[q], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    g_create_error = std::string("HIP setup failed: ") + hipGetErrorString(e);
    lps_destroy(c);
    return LPS_ERR_HIP;
  }
  *out = c;
  return LPS_OK;
}

int32_t lps_create(const lps_problem *problem, const lps_options *options, lps_ctx **out) {
  *out = nullptr;
  if (!problem) {
    g_create_error = "problem is NULL";
    return LPS_ERR_ARG;
  }
  lps_ctx *c = nullptr;
  int32_t rc = new_context(options, &c);
  if (rc != LPS_OK) return rc;
  rc = create_impl(c, problem);
  if (rc == LPS_OK) resolve_schedule(c);
  if (rc == LPS_OK && c->fast) rc = choose_variant(c);
  if (rc != LPS_OK) {
    g_create_error = c->err;
    lps_destroy(c);
    return rc;
  }
  *out = c;
  return LPS_OK;
}

// ---- batches: instances set up one by one (lps_create), then packed into one context ----

extern "C++" {
// dst[total] = the parts one after another (device to device).
template <class T>
static int32_t cat_dev(lps_ctx *c, T **dst, size_t total,
                       const std::vector<std::pair<const T *, size_t>> &parts) {
  CTX_TRY(c, c->alloc(dst, total));
  size_t off = 0;
  for (const auto &pt : parts) {
    if (pt.second)
      CTX_TRY(c, hipMemcpyAsync(*dst + off, pt.first, sizeof(T) * pt.second,
                                hipMemcpyDeviceToDevice, c->stream));
    off += pt.second;
  }
  return LPS_OK;
}

template <class T>
static int32_t to_host(lps_ctx *c, std::vector<T> &h, const T *d, size_t n) {
  h.resize(n);
  if (n) CTX_TRY(c, hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost));
  return LPS_OK;
}
}  // extern "C++"

static int32_t merge_batch(lps_ctx *c, const std::vector<lps_ctx *> &sub) {
  const int32_t B = (int32_t)sub.size();
  const lps_ctx &s0 = *sub[0];
  for (int32_t i = 0; i < B; i++) {
    const lps_ctx &q = *sub[i];
    if (!q.fast || !q.stencil68_ok)
      return c->fail(LPS_ERR_ARG, "batch instance " + std::to_string(i) +
                                      ": a batch needs lattice inputs on the 68-bond tile path");
    if (q.dx != s0.dx || q.dy != s0.dy || q.delta != s0.delta || q.m_ref != s0.m_ref ||
        q.tile.halo != s0.tile.halo)
      return c->fail(LPS_ERR_ARG, "batch instance " + std::to_string(i) +
                                      ": every instance needs the same dx, dy and delta");
  }
  c->batch = true;
  c->fast = true;
  c->stencil68_ok = true;
  c->abc_in_part = true;
  for (const lps_ctx *q : sub) c->abc_in_part = c->abc_in_part && q->abc_in_part;
  c->n_inst = B;
  // scalars shared by the instances (the fast kernels take the material from InstParams)
  c->E = s0.E; c->nu = s0.nu; c->rho = s0.rho; c->dt = s0.dt;
  c->delta = s0.delta; c->dx = s0.dx; c->dy = s0.dy;
  c->lam = s0.lam; c->mu = s0.mu; c->kappa = s0.kappa; c->m_ref = s0.m_ref;
  c->exclude_r = s0.exclude_r;
  // offsets
  c->slot_off.assign(B + 1, 0);
  c->node_off.assign(B + 1, 0);
  c->tile_off.assign(B + 1, 0);
  c->abc_n.resize(B);
  c->node_n.resize(B);
  int32_t M = 0, n_full = 0, n_part = 0;
  for (int32_t i = 0; i < B; i++) {
    c->slot_off[i + 1] = c->slot_off[i] + sub[i]->n_slots;
    c->node_off[i + 1] = c->node_off[i] + sub[i]->n;
    c->tile_off[i + 1] = c->tile_off[i] + sub[i]->tile.n_tiles;
    c->abc_n[i] = sub[i]->n_abc;
    c->node_n[i] = sub[i]->n;
    M = std::max(M, sub[i]->M);
    n_full += sub[i]->tile.n_full;
    n_part += sub[i]->tile.n_part;
    c->n_full_nodes += sub[i]->n_full_nodes;
  }
  if (c->slot_off[B] > INT32_MAX || c->node_off[B] > INT32_MAX)
    return c->fail(LPS_ERR_ARG, "a batch holds at most 2^31 - 1 slots");
  const int32_t NS = (int32_t)c->slot_off[B], NT = c->tile_off[B];
  c->n_slots = NS;
  c->n = (int32_t)c->node_off[B];
  c->M = M;
  // per-slot arrays
  auto parts_d = [&](auto get) {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(get(*sub[0]))>>;
    std::vector<std::pair<const T *, size_t>> v;
    for (int32_t i = 0; i < B; i++) v.push_back({get(*sub[i]), (size_t)sub[i]->n_slots});
    return v;
  };
  int32_t rc;
  if ((rc = cat_dev(c, &c->x, NS, parts_d([](const lps_ctx &q) { return (const double *)q.x; }))) ||
      (rc = cat_dev(c, &c->y, NS, parts_d([](const lps_ctx &q) { return (const double *)q.y; }))) ||
      (rc = cat_dev(c, &c->kind, NS, parts_d([](const lps_ctx &q) { return (const int8_t *)q.kind; }))) ||
      (rc = cat_dev(c, &c->is_bc, NS, parts_d([](const lps_ctx &q) { return (const int8_t *)q.is_bc; }))) ||
      (rc = cat_dev(c, &c->m, NS, parts_d([](const lps_ctx &q) { return (const double *)q.m; }))) ||
      (rc = cat_dev(c, &c->c2m, NS, parts_d([](const lps_ctx &q) { return (const double *)q.c2m; }))) ||
      (rc = cat_dev(c, &c->perm_d, NS, parts_d([](const lps_ctx &q) { return (const int32_t *)q.perm_d; }))) ||
      (rc = cat_dev(c, &c->IC.x, NS, parts_d([](const lps_ctx &q) { return (const double *)q.IC.x; }))) ||
      (rc = cat_dev(c, &c->IC.y, NS, parts_d([](const lps_ctx &q) { return (const double *)q.IC.y; }))) ||
      (rc = cat_dev(c, &c->ICD.x, NS, parts_d([](const lps_ctx &q) { return (const double *)q.ICD.x; }))) ||
      (rc = cat_dev(c, &c->ICD.y, NS, parts_d([](const lps_ctx &q) { return (const double *)q.ICD.y; }))) ||
      (rc = cat_dev(c, &c->TH0, NS, parts_d([](const lps_ctx &q) { return (const double *)q.TH[0]; }))) ||
      (rc = cat_dev(c, &c->P0.x, NS, parts_d([](const lps_ctx &q) { return (const double *)q.P[0].x; }))) ||
      (rc = cat_dev(c, &c->P0.y, NS, parts_d([](const lps_ctx &q) { return (const double *)q.P[0].y; }))))
    return rc;
  // slot -> node of the batch (instances' nodes one after another); perm_d stays per instance
  c->perm_h.assign(NS, -1);
  for (int32_t i = 0; i < B; i++)
    for (int32_t k = 0; k < sub[i]->n_slots; k++) {
      const int32_t o = sub[i]->perm_h[k];
      c->perm_h[c->slot_off[i] + k] = o < 0 ? -1 : (int32_t)(o + c->node_off[i]);
    }
  // tiles: neighbour tables and lists shifted by the instances' tile offsets
  std::vector<int32_t> nbr, full_list, part_list, tinst(NT), txy, deep_list, shallow_list;
  std::vector<int8_t> tfull;
  for (int32_t i = 0; i < B; i++) {
    const TileArgs &T = sub[i]->tile;
    std::vector<int32_t> h;
    if ((rc = to_host(c, h, sub[i]->deep_list, (size_t)sub[i]->n_deep))) return rc;
    for (int32_t v : h) deep_list.push_back(v + c->tile_off[i]);
    if ((rc = to_host(c, h, sub[i]->shallow_list, (size_t)sub[i]->n_shallow))) return rc;
    for (int32_t v : h) shallow_list.push_back(v + c->tile_off[i]);
    if ((rc = to_host(c, h, T.tile_nbr, (size_t)9 * T.n_tiles))) return rc;
    for (int32_t v : h) nbr.push_back(v < 0 ? -1 : v + c->tile_off[i]);
    if ((rc = to_host(c, h, T.full_list, (size_t)T.n_full))) return rc;
    for (int32_t v : h) full_list.push_back(v + c->tile_off[i]);
    if ((rc = to_host(c, h, T.part_list, (size_t)T.n_part))) return rc;
    for (int32_t v : h) part_list.push_back(v + c->tile_off[i]);
    if ((rc = to_host(c, h, T.tile_xy, (size_t)T.n_tiles))) return rc;
    txy.insert(txy.end(), h.begin(), h.end());
    std::vector<int8_t> f;
    if ((rc = to_host(c, f, T.tile_full, (size_t)T.n_tiles))) return rc;
    tfull.insert(tfull.end(), f.begin(), f.end());
    for (int32_t t = 0; t < T.n_tiles; t++) tinst[c->tile_off[i] + t] = i;
  }
  TileArgs &T = c->tile;
  T = s0.tile;  // stencil constants, halo, window, class weights
  T.n_tiles = NT;
  T.ntx = T.nty = 0;
  T.tile_map = nullptr;  // one lattice per instance: the batch kernels use the neighbour tables
  T.n_full = n_full;
  T.n_part = n_part;
  T.theta_kernel = s0.tile.theta_kernel;
  T.force_kernel = s0.tile.force_kernel;
  int32_t *d_nbr, *d_fl, *d_pl, *d_ti, *d_txy;
  int8_t *d_full;
  CTX_TRY(c, c->upload(&d_nbr, nbr));
  CTX_TRY(c, c->upload(&d_fl, full_list));
  CTX_TRY(c, c->upload(&d_pl, part_list));
  CTX_TRY(c, c->upload(&c->deep_list, deep_list));
  CTX_TRY(c, c->upload(&c->shallow_list, shallow_list));
  c->n_deep = (int32_t)deep_list.size();
  c->n_shallow = (int32_t)shallow_list.size();
  CTX_TRY(c, c->upload(&d_ti, tinst));
  CTX_TRY(c, c->upload(&d_txy, txy));
  CTX_TRY(c, c->upload(&d_full, tfull));
  T.tile_nbr = d_nbr; T.full_list = d_fl; T.part_list = d_pl; T.tile_inst = d_ti;
  T.tile_xy = d_txy; T.tile_full = d_full;
  {  // the stencil tables (the same for every instance)
    const size_t nt = (size_t)s0.tile.nt;
    int32_t *d_delta;
    double *d_hx, *d_hy, *d_g1, *d_g2, *d_xi;
    if ((rc = cat_dev(c, &d_delta, nt, {{s0.tile.delta, nt}})) ||
        (rc = cat_dev(c, &d_hx, nt, {{s0.tile.hx, nt}})) ||
        (rc = cat_dev(c, &d_hy, nt, {{s0.tile.hy, nt}})) ||
        (rc = cat_dev(c, &d_g1, nt, {{s0.tile.g1, nt}})) ||
        (rc = cat_dev(c, &d_g2, nt, {{s0.tile.g2, nt}})) ||
        (rc = cat_dev(c, &d_xi, nt, {{s0.tile.xi, nt}})))
      return rc;
    T.delta = d_delta; T.hx = d_hx; T.hy = d_hy; T.g1 = d_g1; T.g2 = d_g2; T.xi = d_xi;
  }
  // ABC rows: slots shifted; the instances' ragged rows and operator blocks one after the other
  int32_t R = 0;
  int64_t NM = 0, GM = 0;
  bool shared = true;
  const bool gth = c->opt.theta_method == LPS_THETA_GTHETA;
  // Instances with the same ABC inputs (equal hashes of node list, clouds and operators, and
  // byte-equal operator blocks) share one copy of the blocks: a batch of runs of one geometry
  // and material then streams one set of operators, which stays in the caches, instead of one
  // per run. gbase: the first block member of each instance's blocks.
  std::vector<int64_t> gbase(B, -1);
  int32_t n_dedup = 0;
  for (int32_t i = 0; i < B; i++) {
    R += sub[i]->n_abc;
    NM += sub[i]->abc_members;
    shared = shared && (sub[i]->n_abc == 0 || sub[i]->abc_shared_blocks);
    const lps_ctx &q = *sub[i];
    const int64_t Gi = q.abc_block_members;
    if (q.n_abc == 0) continue;
    for (int32_t k = 0; k < i && gbase[i] < 0; k++) {
      const lps_ctx &p0 = *sub[k];
      if (gbase[k] < 0 || p0.abc_hash[0] != q.abc_hash[0] || p0.abc_block_members != Gi) continue;
      bool same = true;
      std::vector<double> ha, hb;
      const std::pair<const double *, const double *> blk[3] = {
          {p0.gu, q.gu}, {p0.gv, q.gv}, {gth ? p0.gth : nullptr, gth ? q.gth : nullptr}};
      for (int b = 0; b < 3 && same; b++) {
        if (!blk[b].first) continue;
        const size_t len = (size_t)(b < 2 ? 8 : 2) * Gi;
        if ((rc = to_host(c, ha, blk[b].first, len)) || (rc = to_host(c, hb, blk[b].second, len)))
          return rc;
        same = std::memcmp(ha.data(), hb.data(), sizeof(double) * len) == 0;
      }
      if (same) {
        gbase[i] = gbase[k];
        n_dedup++;
      }
    }
    if (gbase[i] < 0) {
      gbase[i] = GM;
      GM += Gi;
    }
  }
  c->n_abc = R;
  c->abc_members = NM;
  c->abc_block_members = GM;
  c->abc_shared_blocks = shared && R > 0;
  if (NM > INT32_MAX) return c->fail(LPS_ERR_ARG, "too many ABC cloud members in the batch");
  if (R > 0) {
    const size_t nm = (size_t)std::max<int64_t>(NM, 1), gm = (size_t)std::max<int64_t>(GM, 1);
    std::vector<int32_t> aslot, off, gofs, cl;
    std::vector<int8_t> xf;
    std::vector<uint8_t> pmap;
    off.reserve(R + 1);
    gofs.reserve(R);
    xf.reserve(R);
    cl.reserve(nm);
    pmap.reserve(nm);
    off.push_back(0);
    // one member past the last block, as in the single setup (read, never used, by rows with L = 0)
    CTX_TRY(c, c->alloc(&c->gu, 8 * gm + 8));
    CTX_TRY(c, c->alloc(&c->gv, 8 * gm + 8));
    if (gth) CTX_TRY(c, c->alloc(&c->gth, 2 * gm + 2));
    int64_t m0 = 0;
    for (int32_t i = 0; i < B; i++) {
      const lps_ctx &q = *sub[i];
      const int32_t Ri = q.n_abc;
      const int64_t Mi = q.abc_members, Gi = q.abc_block_members, g0 = gbase[i];
      if (Ri == 0) continue;
      // the blocks are copied by the first instance that holds them
      bool first = true;
      for (int32_t k = 0; k < i && first; k++) first = !(sub[k]->n_abc > 0 && gbase[k] == g0);
      std::vector<int32_t> h;
      if ((rc = to_host(c, h, q.abc_slot, (size_t)Ri))) return rc;
      for (int32_t v : h) aslot.push_back((int32_t)(v + c->slot_off[i]));
      if ((rc = to_host(c, h, q.abc_off, (size_t)Ri + 1))) return rc;
      for (int32_t r = 1; r <= Ri; r++) off.push_back((int32_t)(m0 + h[r]));
      if ((rc = to_host(c, h, q.abc_gofs, (size_t)Ri))) return rc;
      for (int32_t v : h) gofs.push_back((int32_t)(g0 + v));
      if ((rc = to_host(c, h, q.cloud, (size_t)Mi))) return rc;
      for (int32_t v : h) cl.push_back(v < 0 ? -1 : (int32_t)(v + c->slot_off[i]));
      std::vector<int8_t> hx;
      if ((rc = to_host(c, hx, q.abc_xf, (size_t)Ri))) return rc;
      xf.insert(xf.end(), hx.begin(), hx.end());
      std::vector<uint8_t> hp;
      if ((rc = to_host(c, hp, q.abc_pmap, (size_t)Mi))) return rc;
      pmap.insert(pmap.end(), hp.begin(), hp.end());
      if (first) {
        CTX_TRY(c, hipMemcpyAsync(c->gu + 8 * g0, q.gu, sizeof(double) * 8 * Gi, hipMemcpyDeviceToDevice, c->stream));
        CTX_TRY(c, hipMemcpyAsync(c->gv + 8 * g0, q.gv, sizeof(double) * 8 * Gi, hipMemcpyDeviceToDevice, c->stream));
        if (gth)
          CTX_TRY(c, hipMemcpyAsync(c->gth + 2 * g0, q.gth, sizeof(double) * 2 * Gi, hipMemcpyDeviceToDevice, c->stream));
      }
      m0 += Mi;
    }
    if (cl.empty()) {
      cl.push_back(-1);
      pmap.push_back(0);
    }
    CTX_TRY(c, c->upload(&c->abc_slot, aslot));
    CTX_TRY(c, c->upload(&c->abc_off, off));
    CTX_TRY(c, c->upload(&c->abc_gofs, gofs));
    CTX_TRY(c, c->upload(&c->abc_xf, xf));
    CTX_TRY(c, c->upload(&c->abc_pmap, pmap));
    CTX_TRY(c, c->upload(&c->cloud, cl));
    if ((rc = build_member_pack(c, cl, off, pmap, NS, gth)) != LPS_OK) return rc;
  }
  // instances
  c->inst_h.clear();
  c->abc_hash.clear();
  for (int32_t i = 0; i < B; i++) {
    c->inst_h.push_back(sub[i]->inst_h[0]);
    c->abc_hash.push_back(sub[i]->abc_hash[0]);
  }
  CTX_TRY(c, c->upload(&c->inst_d, c->inst_h));
  CTX_TRY(c, c->upload(&c->tile_off_d, c->tile_off));
  // state and output buffers
  for (int b = 0; b < 3; b++) {
    CTX_TRY(c, c->alloc(&c->U[b].x, NS)); CTX_TRY(c, c->alloc(&c->U[b].y, NS));
    CTX_TRY(c, c->alloc(&c->V[b].x, NS)); CTX_TRY(c, c->alloc(&c->V[b].y, NS));
    CTX_TRY(c, c->alloc(&c->TH[b], NS));
  }
  for (int b = 0; b < 2; b++) {
    CTX_TRY(c, c->alloc(&c->A[b].x, NS)); CTX_TRY(c, c->alloc(&c->A[b].y, NS));
  }
  for (int b = 0; b < 6; b++) {
    CTX_TRY(c, c->alloc(&c->P[b].x, NS)); CTX_TRY(c, c->alloc(&c->P[b].y, NS));
  }
  CTX_TRY(c, c->alloc(&c->partials, NT));
  CTX_TRY(c, c->alloc(&c->dout, 1));
  CTX_TRY(c, c->alloc(&c->dout_inst, B));
  CTX_TRY(c, hipHostMalloc((void **)&c->hout, sizeof(DiagPartial)));
  CTX_TRY(c, hipHostMalloc((void **)&c->hout_inst, sizeof(DiagPartial) * B));
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  const char *names[3] = {"generic", "unrolled", "blocked"};
  c->path = "batch of " + std::to_string(B) + " instances: tile kernels: dilatation " +
            names[T.theta_kernel] + ", force " + names[T.force_kernel] + " (16x16 tiles, " +
            std::to_string(NT) + " tiles, " + std::to_string(n_full) + " full, " +
            std::to_string(c->n_deep) + " deep, " + std::to_string(c->n) + " nodes, " + std::to_string(R) + " ABC rows)";
  if (c->abc_shared_blocks)
    c->path += "; ABC operators of " + std::to_string(GM) + " of " + std::to_string(NM) +
               " members stored, the rest lattice images";
  if (n_dedup > 0)
    c->path += "; " + std::to_string(n_dedup) + " instances share an earlier one's ABC operators";
  return reset_state(c);
}

// An instance context once set up holds more than merge_batch reads: its horizons (CSR and
// SELL, the largest arrays), the rotating state and the output partials. They are freed right
// after its creation, so that a batch's peak device memory is the merged context plus the
// instances' setup state (about 100 B per slot) instead of every instance in full (about 900 B
// per node): 3e8 nodes fit on one GPU.
static int32_t release_for_merge(lps_ctx *q) {
  CTX_TRY(q, hipStreamSynchronize(q->stream));
  free_horizons(&q->hb);
  auto release = [q](auto *&p) {
    if (!p) return;
    auto it = std::find(q->allocs.begin(), q->allocs.end(), (void *)p);
    if (it != q->allocs.end()) {
      (void)hipFree(p);
      q->allocs.erase(it);
    }
    p = nullptr;
  };
  for (int b = 0; b < 3; b++) {
    release(q->U[b].x); release(q->U[b].y);
    release(q->V[b].x); release(q->V[b].y);
  }
  for (int b = 1; b < 3; b++) release(q->TH[b]);  // TH[0]: the setup's theta (merged)
  for (int b = 0; b < 2; b++) {
    release(q->A[b].x); release(q->A[b].y);
  }
  for (int b = 1; b < 6; b++) {  // P[0]: the first prediction (merged)
    release(q->P[b].x); release(q->P[b].y);
  }
  release(q->partials);
  return LPS_OK;
}

int32_t lps_create_batch(const lps_problem *problems, int32_t count, const lps_options *options,
                         lps_ctx **out) {
  *out = nullptr;
  if (!problems || count < 1) {
    g_create_error = "a batch needs at least one problem";
    return LPS_ERR_ARG;
  }
  lps_options opt;
  if (options) opt = *options;
  else lps_default_options(&opt);
  if (opt.exact) {
    g_create_error = "batches run the tile path only (exact = 0)";
    return LPS_ERR_ARG;
  }
  // the instances run the register-blocked kernels unless the unrolled ones are asked for
  lps_options sub_opt = opt;
  sub_opt.variant = opt.variant == 2 ? 2 : 3;
  sub_opt.use_graph = 0;
  lps_ctx *c = nullptr;
  int32_t rc = new_context(&opt, &c);
  if (rc != LPS_OK) return rc;
  std::vector<lps_ctx *> sub;
  for (int32_t i = 0; i < count && rc == LPS_OK; i++) {
    lps_ctx *q = nullptr;
    rc = lps_create(&problems[i], &sub_opt, &q);
    if (rc != LPS_OK) {
      c->err = "batch instance " + std::to_string(i) + ": " + g_create_error;
    } else {
      sub.push_back(q);
      if ((rc = release_for_merge(q)) != LPS_OK) c->err = "batch instance " + std::to_string(i) + ": " + q->err;
    }
  }
  if (rc == LPS_OK) rc = merge_batch(c, sub);
  if (rc == LPS_OK) resolve_schedule(c);
  for (lps_ctx *q : sub) lps_destroy(q);
  if (rc == LPS_OK) rc = choose_abc_predict(c, opt.variant == 0);
  if (rc != LPS_OK) {
    g_create_error = c->err;
    lps_destroy(c);
    return rc;
  }
  *out = c;
  return LPS_OK;
}

int32_t lps_instance_count(const lps_ctx *ctx) { return ctx->n_inst; }

// Fast path with ABC rows on their own stream: the ABC nodes' prediction folded into the partial
// tiles' force pass (which then waits for G_u / G_v) or a kernel of its own after G_u / G_v (the
// partial tiles then run beside them). The same bits either way. With tune, two rounds of 3
// warm-up and 10 timed steps each and the faster by the whole step; LPS_ABC_PREDICT=fold or
// =apart sets the choice (tests of both schedules).
static int32_t choose_abc_predict(lps_ctx *c, bool tune) {
  const char *force_pred = std::getenv("LPS_ABC_PREDICT");
  if (force_pred && (!strcmp(force_pred, "fold") || !strcmp(force_pred, "apart"))) {
    c->fold_pred = !strcmp(force_pred, "fold");
  } else if (tune && c->n_abc > 0 && c->opt.concurrent == 1 && c->abc_in_part &&
             c->tile.force_kernel != kGeneric) {
    double tf[2] = {1e300, 1e300};
    for (int round = 0; round < 2; round++) {
      for (int q = 0; q < 2; q++) {
        double t[4] = {0, 0, 0, 0};
        c->fold_pred = q == 0;
        int32_t rc = lps_set_timing(c, 1);
        if (rc == LPS_OK) rc = lps_step(c, 3);
        if (rc == LPS_OK) rc = lps_set_timing(c, 1);
        if (rc == LPS_OK) rc = lps_step(c, 10);
        if (rc == LPS_OK) rc = lps_pass_times(c, t);
        if (rc == LPS_OK) rc = lps_set_timing(c, 0);
        if (rc == LPS_OK) rc = reset_state(c);
        if (rc != LPS_OK) return rc;
        tf[q] = std::min(tf[q], t[3]);
      }
    }
    c->fold_pred = tf[0] <= tf[1];
  } else {
    return LPS_OK;
  }
  drop_graphs(c);
  c->path += c->fold_pred ? "; ABC prediction in the partial tiles' force pass" : "; ABC prediction apart";
  return LPS_OK;
}

// Fast path: pick the node kernels of each pass. With variant 0 every candidate (generic tile
// kernels, unrolled 68-bond kernels, register-blocked 68-bond kernels) runs a few timed steps on
// the problem itself and each pass keeps its fastest; the state is reset afterwards.
static int32_t choose_variant(lps_ctx *c) {
  const bool tune = c->stencil68_ok && c->opt.variant == 0;
  int32_t th = kGeneric, fo = kGeneric;
  if (!c->stencil68_ok || c->opt.variant == 1) {
    th = fo = kGeneric;
  } else if (c->opt.variant == 2) {
    th = fo = kUnrolled;
  } else if (c->opt.variant == 3) {
    th = fo = kBlocked;
  } else {
    // two rounds of 3 warm-up and 10 timed steps per family, the minimum of the rounds; the
    // dilatation family by its pass, the force family by the rest of the step (with the partial
    // tiles on a third stream the force's share moves between the timed intervals)
    double best_th = 1e300, best_fo = 1e300, best_step = 1e300;
    double tth[3] = {1e300, 1e300, 1e300}, tfo[3] = {1e300, 1e300, 1e300};
    double tst[3] = {1e300, 1e300, 1e300};
    for (int round = 0; round < 2; round++) {
      for (int32_t m : {kGeneric, kUnrolled, kBlocked}) {
        double t[4] = {0, 0, 0, 0};
        c->tile.theta_kernel = c->tile.force_kernel = m;
        int32_t rc = lps_set_timing(c, 1);
        if (rc == LPS_OK) rc = lps_step(c, 3);
        if (rc == LPS_OK) rc = lps_set_timing(c, 1);
        if (rc == LPS_OK) rc = lps_step(c, 10);
        if (rc == LPS_OK) rc = lps_pass_times(c, t);
        if (rc == LPS_OK) rc = lps_set_timing(c, 0);
        if (rc == LPS_OK) rc = reset_state(c);
        if (rc != LPS_OK) return rc;
        tth[m] = std::min(tth[m], t[0]);
        tfo[m] = std::min(tfo[m], t[3] - t[0]);
        tst[m] = std::min(tst[m], t[3]);
      }
    }
    for (int32_t m : {kGeneric, kUnrolled, kBlocked}) {
      if (tth[m] < best_th) { best_th = tth[m]; th = m; }
      if (tfo[m] < best_fo) { best_fo = tfo[m]; fo = m; }
    }
    // concurrent == 4 fuses the deep tiles' two passes when both are register-blocked, so its
    // first interval holds force work: one family for both passes, by the whole step
    if (c->opt.concurrent == 4)
      for (int32_t m : {kGeneric, kUnrolled, kBlocked})
        if (tst[m] < best_step) { best_step = tst[m]; th = fo = m; }
  }
  c->tile.theta_kernel = th;
  c->tile.force_kernel = fo;
  drop_graphs(c);  // graphs hold the kernels they were captured with
  const char *names[3] = {"generic", "unrolled", "blocked"};
  c->path = std::string("tile kernels: dilatation ") + names[th] + ", force " + names[fo] +
            (tune ? " [autotuned]" : "") + c->path;
  return choose_abc_predict(c, tune);
}

constexpr int kEv = 4;  // events per timed step

// A timing event on the context's stream; under capture it becomes an event-record node of the
// graph (hipEventRecordExternal), so a replay records it like a direct launch would.
static hipError_t record_timing(lps_ctx *c, hipEvent_t e) {
  return hipEventRecordWithFlags(e, c->stream, c->capturing ? hipEventRecordExternal : 0);
}

// The ABC members' records of step c->step: its U, V level, and the previous step's unless the
// records already hold it (written as the previous step's level by the step just before). The
// first step of a captured 30-step graph always writes both, so a replay is right after a reset
// or a restore too.
static hipError_t pack_members(lps_ctx *c, const StepState &st, hipStream_t s) {
  const int64_t n = c->step;
  const bool prev = n % 30 == 0 || c->pk_last_step != n - 1;
  c->pk_last_step = n;
  return launch_abc_pack(c->n_pk, c->pk_slot, c->kind, st, c->ring(n), c->pk_lvl[n & 1],
                         prev ? c->pk_lvl[(n + 1) & 1] : nullptr, c->pk_th, s);
}

// Fast path, concurrent == 3: one stream, three launches. The ABC rows' G_theta and G_u / G_v run as
// extra workgroups of the dilatation launch (they read only the previous steps' state), then the
// full tiles' force, then the partial tiles' force with the ABC nodes' prediction. A learned
// closure runs on the fourth stream from the step's start and its corrections are added before
// the ABC nodes' prediction.
static hipError_t step_fused(lps_ctx *c, const StepArgs &a, const StepState &st, hipEvent_t *ev,
                             bool record_events) {
  hipError_t e;
  if (c->clo) {
    if ((e = hipEventRecord(c->ev_fork[0], c->stream)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(c->stream4, c->ev_fork[0], 0)) != hipSuccess) return e;
    if ((e = closure_step_hook(c, st, c->stream4, true)) != hipSuccess) return e;
    if ((e = hipEventRecord(c->ev_clo, c->stream4)) != hipSuccess) return e;
  }
  if ((e = pack_members(c, st, c->stream)) != hipSuccess) return e;
  if ((e = launch_theta_rows(a, st, c->pack(c->step), c->stream)) != hipSuccess) return e;
  if (a.theta_method == LPS_THETA_IDW && (e = launch_idw(a, st, c->stream)) != hipSuccess) return e;
  const bool ev_force = record_events && !c->timing_closure;
  if (ev_force && (e = record_timing(c, ev[1])) != hipSuccess) return e;
  if ((e = launch_force(a, st, true, c->stream)) != hipSuccess) return e;
  if (ev_force && (e = record_timing(c, ev[2])) != hipSuccess) return e;
  if (!c->clo && c->abc_in_part && a.tile.force_kernel != kGeneric)
    return launch_force_part(a, st, c->stream, true);
  if ((e = launch_force_part(a, st, c->stream)) != hipSuccess) return e;
  if (c->clo) {
    if ((e = hipStreamWaitEvent(c->stream, c->ev_clo, 0)) != hipSuccess) return e;
    if ((e = closure_apply_hook(c, st, c->stream)) != hipSuccess) return e;
  }
  return launch_abc_predict(a, st, c->stream);
}

static hipError_t step_once(lps_ctx *c, bool record_events) {
  const StepArgs a = c->args();
  const StepState st = c->state(c->step);
  hipEvent_t *ev = nullptr;
  hipError_t e;
  if (record_events) {
    ev = &c->ev[kEv * (30 * (size_t)c->ev_set + c->ev_used)];
    if (!c->timing_force_only && (e = record_timing(c, ev[0])) != hipSuccess) return e;
  }
  c->clo_ev = record_events && c->timing_closure ? ev : nullptr;  // recorded by the closure hook
  if (c->fast && a.n_abc > 0 && c->opt.concurrent == 3) {
    if ((e = step_fused(c, a, st, ev, record_events)) != hipSuccess) return e;
    if (record_events) {
      if (!c->timing_force_only && (e = record_timing(c, ev[3])) != hipSuccess) return e;
      c->ev_used++;
    }
    c->step++;
    c->u_synced = 0;
    return hipSuccess;
  }
  // fast path: the ABC rows' G_theta and G_u / G_v read only the previous steps' state and run on
  // the second stream beside the tile passes; the force pass waits for G_theta (ev_join[0]), the
  // ABC prediction for G_u / G_v (ev_join[1])
  const bool abc = c->fast && a.n_abc > 0;
  const bool fork = abc && c->opt.concurrent;
  // concurrent == 2: the partial tiles' kernels on a third stream, beside the full tiles' kernels
  const bool split = c->fast && c->opt.concurrent == 2 && a.tile.n_part > 0 &&
                     a.tile.theta_kernel == kBlocked && a.tile.force_kernel != kGeneric;
  // concurrent == 4: the deep full tiles' dilatation + force in one launch on the main stream, the
  // other tiles' dilatation on the third stream beside it; the shallow full tiles' force then
  // waits for that dilatation and for G_theta
  const bool deep = c->fast && c->opt.concurrent == 4 && c->n_deep > 0 &&
                    a.tile.theta_kernel == kBlocked && a.tile.force_kernel == kBlocked;
  // the closure's history shift and network first, alone on the GPU, on the main stream (its
  // corrections stored, added after G_u / G_v): beside the tile kernels its 152 KB workgroups wait
  // for CUs with that much LDS free, so the step gains nothing from the overlap
  const bool clo_first = fork && c->clo && c->clo_first;
  if (clo_first && (e = closure_step_hook(c, st, c->stream, true)) != hipSuccess) return e;
  if (fork || split || deep) {
    if ((e = hipEventRecord(c->ev_fork[0], c->stream)) != hipSuccess) return e;
  }
  if (deep) {
    if ((e = hipStreamWaitEvent(c->stream3, c->ev_fork[0], 0)) != hipSuccess) return e;
    if ((e = launch_theta_list(a, st, c->shallow_list, c->n_shallow, c->stream3)) != hipSuccess)
      return e;
    if ((e = hipEventRecord(c->ev_part[0], c->stream3)) != hipSuccess) return e;
  }
  if (split) {
    if ((e = hipStreamWaitEvent(c->stream3, c->ev_fork[0], 0)) != hipSuccess) return e;
    if ((e = launch_theta_part(a, st, c->stream3)) != hipSuccess) return e;
    if ((e = hipEventRecord(c->ev_part[0], c->stream3)) != hipSuccess) return e;
  }
  if (fork) {
    if ((e = hipStreamWaitEvent(c->stream2, c->ev_fork[0], 0)) != hipSuccess) return e;
    // the closure's history shift and network read only step n's state: on the fourth stream
    // from the step's start, their corrections stored and added after G_u / G_v
    if (c->clo && !clo_first) {
      if ((e = hipStreamWaitEvent(c->stream4, c->ev_fork[0], 0)) != hipSuccess) return e;
      if ((e = closure_step_hook(c, st, c->stream4, true)) != hipSuccess) return e;
      if ((e = hipEventRecord(c->ev_clo, c->stream4)) != hipSuccess) return e;
    }
    if ((e = pack_members(c, st, c->stream2)) != hipSuccess) return e;
    if ((e = launch_abc_theta(a, st, c->pack(c->step), c->stream2)) != hipSuccess) return e;
    if ((e = hipEventRecord(c->ev_join[0], c->stream2)) != hipSuccess) return e;
    if ((e = launch_abc_uv(a, st, c->pack(c->step), c->stream2)) != hipSuccess) return e;
    if (c->clo) {
      if (!clo_first && (e = hipStreamWaitEvent(c->stream2, c->ev_clo, 0)) != hipSuccess) return e;
      if ((e = closure_apply_hook(c, st, c->stream2)) != hipSuccess) return e;
    }
    if ((e = hipEventRecord(c->ev_join[1], c->stream2)) != hipSuccess) return e;
  } else if (abc) {
    if ((e = pack_members(c, st, c->stream)) != hipSuccess) return e;
    if ((e = launch_abc_theta(a, st, c->pack(c->step), c->stream)) != hipSuccess) return e;
  }
  if (deep) {
    if ((e = launch_force_deep(a, st, c->deep_list, c->n_deep, c->stream)) != hipSuccess) return e;
  } else if ((e = launch_theta(a, st, c->fast, c->stream, !split)) != hipSuccess) {
    return e;
  }
  if ((split || deep) && (e = hipStreamWaitEvent(c->stream, c->ev_part[0], 0)) != hipSuccess)
    return e;
  if (fork && (e = hipStreamWaitEvent(c->stream, c->ev_join[0], 0)) != hipSuccess) return e;
  if (a.n_abc > 0 && a.theta_method == LPS_THETA_IDW && (e = launch_idw(a, st, c->stream)) != hipSuccess)
    return e;
  if (split) {  // every theta is written: the partial tiles' force may start
    if ((e = hipEventRecord(c->ev_part[1], c->stream)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(c->stream3, c->ev_part[1], 0)) != hipSuccess) return e;
    if ((e = launch_force_part(a, st, c->stream3)) != hipSuccess) return e;
    if ((e = hipEventRecord(c->ev_part[2], c->stream3)) != hipSuccess) return e;
  }
  const bool ev_force = record_events && !c->timing_closure;
  if (ev_force && (e = record_timing(c, ev[1])) != hipSuccess) return e;
  if (deep) {
    if ((e = launch_force_list(a, st, c->shallow_list, c->n_shallow, c->stream)) != hipSuccess)
      return e;
  } else if ((e = launch_force(a, st, c->fast, c->stream)) != hipSuccess) {
    return e;
  }
  if (ev_force && (e = record_timing(c, ev[2])) != hipSuccess) return e;
  // the ABC nodes' prediction folded into the partial tiles' force pass (every ABC node lies in a
  // partial tile), which then waits for G_u / G_v (and a closure's stored corrections, when its
  // network ran first in the step; not when the network runs beside the tile kernels, a long ABC
  // branch the partial tiles would wait for)
  const bool fold = abc && !split && (!c->clo || clo_first) && c->abc_in_part && c->fold_pred &&
                    a.tile.force_kernel != kGeneric;
  if (c->fast && fold) {
    if (fork && (e = hipStreamWaitEvent(c->stream, c->ev_join[1], 0)) != hipSuccess) return e;
    if (!fork && (e = launch_abc_uv(a, st, c->pack(c->step), c->stream)) != hipSuccess) return e;
    if ((e = launch_force_part(a, st, c->stream, true)) != hipSuccess) return e;
  } else if (c->fast) {
    if (!split && (e = launch_force_part(a, st, c->stream)) != hipSuccess) return e;
    if (split && (e = hipStreamWaitEvent(c->stream, c->ev_part[2], 0)) != hipSuccess) return e;
    if (fork && (e = hipStreamWaitEvent(c->stream, c->ev_join[1], 0)) != hipSuccess) return e;
    if (abc && !fork && (e = launch_abc_uv(a, st, c->pack(c->step), c->stream)) != hipSuccess) return e;
    if (abc && !fork && c->clo && (e = closure_step_hook(c, st, c->stream, false)) != hipSuccess) return e;
    if (abc && (e = launch_abc_predict(a, st, c->stream)) != hipSuccess) return e;
  }
  if (c->clo && !abc && (e = closure_step_hook(c, st, c->stream, false)) != hipSuccess) return e;
  if (record_events) {
    if (!c->timing_force_only && (e = record_timing(c, ev[3])) != hipSuccess) return e;
    c->ev_used++;
  }
  c->step++;
  c->u_synced = 0;
  return hipSuccess;
}

// Reads the times of `used` steps of event set `set`. Events recorded by a graph the runtime cannot
// time switch timing mode to separate launches (the samples are dropped).
static int32_t flush_set(lps_ctx *c, int32_t set, int32_t used, bool from_graph) {
  if (used == 0) return LPS_OK;
  hipEvent_t *e0 = &c->ev[kEv * 30 * (size_t)set];
  const bool fo = c->timing_force_only;  // events 1 and 2 only
  CTX_TRY(c, hipEventSynchronize(e0[kEv * (used - 1) + (fo ? 2 : kEv - 1)]));
  std::vector<float> t((size_t)used * kEv, 0.0f);
  for (int32_t k = 0; k < used; k++) {
    hipEvent_t *e = e0 + kEv * k;
    for (int q = 0; q < kEv; q++) {
      if (fo && q != 1) continue;
      hipError_t r = q < kEv - 1 ? hipEventElapsedTime(&t[kEv * k + q], e[q], e[q + 1])
                                 : hipEventElapsedTime(&t[kEv * k + q], e[0], e[kEv - 1]);
      if (r != hipSuccess && from_graph) {
        (void)hipGetLastError();
        c->graph_timing = false;
        return LPS_OK;
      }
      if (r != hipSuccess) return c->hip_fail(r, "hipEventElapsedTime");
    }
  }
  for (int32_t k = 0; k < used; k++) {
    for (int q = 0; q < kEv; q++) c->t_pass[q] += t[kEv * k + q];
    c->t_count++;
  }
  return LPS_OK;
}

// Reads every recorded step's times: the pending replay's set, then the steps launched one by one.
static int32_t flush_events(lps_ctx *c) {
  const int32_t prev = c->ev_prev_used, used = c->ev_used;
  c->ev_prev_used = 0;
  c->ev_used = 0;
  int32_t rc = flush_set(c, 1 - c->ev_set, prev, true);
  if (rc != LPS_OK) return rc;
  return flush_set(c, c->ev_set, used, false);
}

int32_t lps_set_timing(lps_ctx *c, int32_t on) {
  if (on < 0 || on > 3) return c->fail(LPS_ERR_ARG, "timing mode must be 0, 1, 2 or 3");
  if (on == 3 && !c->clo) return c->fail(LPS_ERR_ARG, "timing mode 3 needs a closure attached");
  int32_t rc = flush_events(c);
  if (rc != LPS_OK) return rc;
  const bool fo = on == 2 || on == 3, tc = on == 3;
  if (fo != c->timing_force_only || tc != c->timing_closure) {
    // the timed graphs hold the other mode's event nodes
    for (hipGraphExec_t &g : c->graph30t) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  }
  c->timing_force_only = fo;
  c->timing_closure = tc;
  c->timing = on != 0;
  for (double &t : c->t_pass) t = 0.0;
  c->t_count = 0;
  if (c->timing && c->ev.empty()) {
    c->ev.resize(2 * kEv * 30);
    for (auto &e : c->ev) CTX_TRY(c, hipEventCreate(&e));
  }
  return LPS_OK;
}

int32_t lps_pass_times(lps_ctx *c, double *out4) {
  int32_t rc = flush_events(c);
  if (rc != LPS_OK) return rc;
  const double n = c->t_count > 0 ? (double)c->t_count : 1.0;
  for (int q = 0; q < kEv; q++) out4[q] = c->t_pass[q] / n;
  return LPS_OK;
}

// 30 steps from a step that is a multiple of 30 (the buffer rotation repeats every 6 steps),
// captured once and replayed.
// timed: with the timing events of event set c->ev_set inside, into graph30t[c->ev_set].
static int32_t build_graph30(lps_ctx *c, bool timed) {
  hipGraphExec_t *out = timed ? &c->graph30t[c->ev_set] : &c->graph30;
  hipGraph_t g = nullptr;
  CTX_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int64_t s0 = c->step;
  hipError_t err = hipSuccess;
  c->capturing = true;
  c->ev_used = 0;
  for (int k = 0; k < 30 && err == hipSuccess; k++) err = step_once(c, timed);
  c->capturing = false;
  c->ev_used = 0;
  c->step = s0;
  hipError_t e2 = hipStreamEndCapture(c->stream, &g);
  if (err != hipSuccess) return c->hip_fail(err, "capture of the step loop");
  if (e2 != hipSuccess) return c->hip_fail(e2, "hipStreamEndCapture");
  hipError_t e3 = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e3 != hipSuccess) {
    *out = nullptr;
    return c->hip_fail(e3, "hipGraphInstantiate");
  }
  return LPS_OK;
}

static void drop_graphs(lps_ctx *c) {
  if (c->graph30) (void)hipGraphExecDestroy(c->graph30);
  for (hipGraphExec_t &g : c->graph30t) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  c->graph30 = nullptr;
}

int32_t lps_step(lps_ctx *c, int32_t nsteps) {
  if (nsteps < 0) return c->fail(LPS_ERR_ARG, "nsteps must be >= 0");
  int32_t done = 0;
  while (done < nsteps) {
    const int32_t left = nsteps - done;
    if (c->opt.use_graph && c->step % 30 == 0 && left >= 30 && (!c->timing || c->graph_timing)) {
      if (c->timing) {
        // steps launched one by one are read first; then this replay records its 30 steps into
        // event set k while the previous replay's set is read
        int32_t rc = flush_set(c, c->ev_set, c->ev_used, false);
        c->ev_used = 0;
        if (rc != LPS_OK) return rc;
        const int32_t k = c->ev_set;
        if (!c->graph30t[k] && (rc = build_graph30(c, true)) != LPS_OK) return rc;
        CTX_TRY(c, hipGraphLaunch(c->graph30t[k], c->stream));
        rc = flush_set(c, 1 - k, c->ev_prev_used, true);
        if (rc != LPS_OK) return rc;
        c->ev_prev_used = 30;
        c->ev_set = 1 - k;
        if (!c->graph_timing) {  // this runtime cannot time graph events: separate launches
          c->ev_prev_used = 0;
          c->step += 30;
          c->u_synced = 0;
          done += 30;
          continue;
        }
      } else {
        if (!c->graph30) {
          int32_t rc = build_graph30(c, false);
          if (rc != LPS_OK) return rc;
        }
        CTX_TRY(c, hipGraphLaunch(c->graph30, c->stream));
      }
      c->step += 30;
      c->u_synced = 0;
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

static int32_t diag_fill(lps_ctx *c, const DiagPartial &d, double dt, double E0, int32_t n_abc,
                         lps_diag *out) {
  const int64_t s = c->step;
  out->step = s;
  out->t = (double)s * dt;
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
  out->ub_rms = n_abc > 0 ? std::sqrt(d.ub / n_abc) : 0.0;
  out->vb_rms = n_abc > 0 ? std::sqrt(d.vb / n_abc) : 0.0;
  out->max_u_node = d.max_u_orig == INT64_MAX ? -1 : (int32_t)d.max_u_orig;
  out->max_v_node = d.max_v_orig == INT64_MAX ? -1 : (int32_t)d.max_v_orig;
  out->de_rel = (E0 > 0.0) ? (d.energy - E0) / E0 : 0.0;
  return LPS_OK;
}

int32_t lps_diagnostics_batch(lps_ctx *c, lps_diag *out) {
  const int64_t s = c->step;
  const StepArgs a = c->args();
  if (!c->fast) return c->fail(LPS_ERR_ARG, "per-instance diagnostics need the tile path");
  int32_t rc = ctx_sync_u(c, 1);
  if (rc != LPS_OK) return rc;
  const Dev2 &U = c->U[s % 3], &V = c->V[s % 3], &A = c->A[s % 2];
  const double *th = c->TH[s % 3];
  if (c->batch) {
    CTX_TRY(c, launch_diag_tile(a, U.x, U.y, V.x, V.y, A.x, A.y, th, c->perm_d, c->partials,
                                nullptr, c->stream));
    CTX_TRY(c, launch_diag_final_inst(c->partials, c->tile_off_d, c->n_inst, c->dout_inst,
                                      c->stream));
    CTX_TRY(c, hipMemcpyAsync(c->hout_inst, c->dout_inst, sizeof(DiagPartial) * c->n_inst,
                              hipMemcpyDeviceToHost, c->stream));
    CTX_TRY(c, hipStreamSynchronize(c->stream));
    if ((int32_t)c->E0_inst.size() != c->n_inst) c->E0_inst.assign(c->n_inst, -1.0);
    for (int32_t i = 0; i < c->n_inst; i++)
      diag_fill(c, c->hout_inst[i], c->inst_h[i].dt, c->E0_inst[i], c->abc_n[i], &out[i]);
    return LPS_OK;
  }
  return lps_diagnostics(c, out);
}

int32_t lps_run_batch(lps_ctx *c, int32_t nsteps, lps_diag *records, int32_t max_records) {
  if (nsteps < 0) return c->fail(LPS_ERR_ARG, "nsteps must be >= 0");
  const int32_t B = c->n_inst;
  if ((int32_t)c->E0_inst.size() != B) c->E0_inst.assign(B, -1.0);
  std::vector<lps_diag> d(B);
  int32_t done = 0, nrec = 0;
  while (done < nsteps) {
    const int32_t to_out = 30 - (int32_t)(c->step % 30);
    const int32_t chunk = std::min(to_out, nsteps - done);
    int32_t rc = lps_step(c, chunk);
    if (rc != LPS_OK) return rc;
    done += chunk;
    if (c->step % 30 == 0) {
      rc = c->batch ? lps_diagnostics_batch(c, d.data()) : lps_diagnostics(c, d.data());
      if (rc != LPS_OK) return rc;
      for (int32_t i = 0; i < B; i++) {
        double &E0 = c->batch ? c->E0_inst[i] : c->E0;
        if (E0 < 0.0) E0 = d[i].energy;  // src/main.cpp:1046, per instance
        d[i].de_rel = (E0 > 0.0) ? (d[i].energy - E0) / E0 : 0.0;
        if (records && nrec < max_records) records[(size_t)nrec * B + i] = d[i];
      }
      nrec++;
      if (c->clo && (rc = closure_record_hook(c, d.data(), B)) != LPS_OK) return rc;
      for (int32_t i = 0; i < B; i++)
        if (std::isnan(d[i].energy) || std::isinf(d[i].energy) || std::isnan(d[i].max_u) ||
            std::isinf(d[i].max_u))
          return c->fail(LPS_ERR_NONFINITE, "instance " + std::to_string(i) +
                                                ": energy or displacement is NaN/Inf at step " +
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

int32_t lps_diagnostics(lps_ctx *c, lps_diag *out) {
  if (c->batch) return c->fail(LPS_ERR_ARG, "a batched context reports per instance: use lps_diagnostics_batch");
  const int64_t s = c->step;
  const StepArgs a = c->args();
  int32_t rc = ctx_sync_u(c, 1);
  if (rc != LPS_OK) return rc;
  const Dev2 &U = c->U[s % 3], &V = c->V[s % 3], &A = c->A[s % 2];
  const double *th = c->TH[s % 3];
  if (c->fast)
    CTX_TRY(c, launch_diag_tile(a, U.x, U.y, V.x, V.y, A.x, A.y, th, c->perm_d, c->partials,
                                c->dout, c->stream));
  else
    CTX_TRY(c, launch_diag(a, U.x, U.y, V.x, V.y, A.x, A.y, th, c->perm_d, c->exclude_r,
                           c->partials, c->n_partials, c->dout, c->stream));
  CTX_TRY(c, hipMemcpyAsync(c->hout, c->dout, sizeof(DiagPartial), hipMemcpyDeviceToHost, c->stream));
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  return diag_fill(c, *c->hout, c->dt, c->E0, c->n_abc, out);
}

int32_t lps_run(lps_ctx *c, int32_t nsteps, lps_diag *records, int32_t max_records) {
  if (c->batch) return c->fail(LPS_ERR_ARG, "a batched context runs with lps_run_batch");
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
      if (c->clo && (rc = closure_record_hook(c, &d, 1)) != LPS_OK) return rc;
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

// Fields of the slots [s0, s1) scattered to node order; node ids are perm_h minus node0.
static int32_t fetch_range(lps_ctx *c, int64_t s0, int64_t s1, int64_t node0, double *u, double *v,
                           double *a, double *theta, double *m) {
  const int64_t s = c->step;
  const size_t NS = (size_t)(s1 - s0);
  std::vector<double> bx(NS), by(NS);
  auto pull2 = [&](const Dev2 &d, double *out) -> int32_t {
    if (!out) return LPS_OK;
    CTX_TRY(c, hipMemcpyAsync(bx.data(), d.x + s0, sizeof(double) * NS, hipMemcpyDeviceToHost, c->stream));
    CTX_TRY(c, hipMemcpyAsync(by.data(), d.y + s0, sizeof(double) * NS, hipMemcpyDeviceToHost, c->stream));
    CTX_TRY(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < NS; k++) {
      const int64_t o = c->perm_h[s0 + k];
      if (o < 0) continue;
      out[2 * (o - node0)] = bx[k];
      out[2 * (o - node0) + 1] = by[k];
    }
    return LPS_OK;
  };
  auto pull1 = [&](const double *d, double *out) -> int32_t {
    if (!out) return LPS_OK;
    CTX_TRY(c, hipMemcpyAsync(bx.data(), d + s0, sizeof(double) * NS, hipMemcpyDeviceToHost, c->stream));
    CTX_TRY(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < NS; k++)
      if (c->perm_h[s0 + k] >= 0) out[c->perm_h[s0 + k] - node0] = bx[k];
    return LPS_OK;
  };
  int32_t rc;
  if (u && (rc = ctx_sync_u(c, 1)) != LPS_OK) return rc;
  if ((rc = pull2(c->U[s % 3], u)) != LPS_OK) return rc;
  if ((rc = pull2(c->V[s % 3], v)) != LPS_OK) return rc;
  if ((rc = pull2(c->A[s % 2], a)) != LPS_OK) return rc;
  if ((rc = pull1(c->TH[s % 3], theta)) != LPS_OK) return rc;
  if ((rc = pull1(c->m, m)) != LPS_OK) return rc;
  return LPS_OK;
}

int32_t lps_fetch(lps_ctx *c, double *u, double *v, double *a, double *theta, double *m) {
  if (c->batch) return c->fail(LPS_ERR_ARG, "a batched context fetches per instance: use lps_fetch_instance");
  return fetch_range(c, 0, c->n_slots, 0, u, v, a, theta, m);
}

int32_t lps_fetch_instance(lps_ctx *c, int32_t inst, double *u, double *v, double *a,
                           double *theta, double *m) {
  if (inst < 0 || inst >= c->n_inst) return c->fail(LPS_ERR_ARG, "instance out of range");
  if (!c->batch) return fetch_range(c, 0, c->n_slots, 0, u, v, a, theta, m);
  return fetch_range(c, c->slot_off[inst], c->slot_off[inst + 1], c->node_off[inst], u, v, a,
                     theta, m);
}

int64_t lps_horizons(lps_ctx *c, int64_t *row_ptr, int32_t *cols) {
  if (c->batch) return c->fail(LPS_ERR_ARG, "a batched context keeps no horizons");
  const int32_t N = c->n, NS = c->n_slots;
  if (!cols && !row_ptr) return c->hb.nnz;
  std::vector<int64_t> rp(NS + 1);
  std::vector<int32_t> cl(std::max<int64_t>(c->hb.nnz, 1));
  CTX_TRY(c, hipMemcpyAsync(rp.data(), c->hb.row_ptr, sizeof(int64_t) * (NS + 1), hipMemcpyDeviceToHost, c->stream));
  CTX_TRY(c, hipMemcpyAsync(cl.data(), c->hb.col, sizeof(int32_t) * c->hb.nnz, hipMemcpyDeviceToHost, c->stream));
  CTX_TRY(c, hipStreamSynchronize(c->stream));
  std::vector<int32_t> slot_of(N);
  for (int32_t k = 0; k < NS; k++)
    if (c->perm_h[k] >= 0) slot_of[c->perm_h[k]] = k;
  int64_t pos = 0;
  if (row_ptr) row_ptr[0] = 0;
  for (int32_t o = 0; o < N; o++) {
    const int32_t k = slot_of[o];
    const int64_t len = rp[k + 1] - rp[k];
    if (cols)
      for (int64_t t = 0; t < len; t++) cols[pos + t] = c->perm_h[cl[rp[k] + t]];
    pos += len;
    if (row_ptr) row_ptr[o + 1] = pos;
  }
  return c->hb.nnz;
}

}  // extern "C"

void lps::ctx_drop_graphs(lps_ctx *c) { drop_graphs(c); }

int32_t lps::ctx_sync_u(lps_ctx *c, int32_t depth) {
  if (!c->fast || c->u_synced >= depth) return LPS_OK;
  const StepArgs a = c->args();
  for (int32_t j = c->u_synced; j < depth; j++) {
    const Dev2 &R = c->ring_u(c->step, j), &U = c->U[(c->step + 3 - j) % 3];
    CTX_TRY(c, launch_u_from_ring(a, R.x, R.y, U.x, U.y, c->stream));
  }
  c->u_synced = depth;
  return LPS_OK;
}
