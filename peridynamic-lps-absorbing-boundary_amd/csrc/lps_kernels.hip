// HIP kernels of the LPS-ABC time loop for gfx950 (MI355X).
//
// One time step is two launches:
//   dilatation pass  theta_i of every non-ABC node and, in the same grid, G_theta at the ABC rows
//                    (one wavefront per row);
//   force pass       force density of every node plus the fused epilogue: velocity-Verlet
//                    corrector, G_u / G_v at the ABC rows (one wavefront per row) and the next
//                    step's predictor, written to ping-pong buffers so that no thread overwrites a
//                    value another thread of the same launch still gathers.
// The output block (energy, momentum, maxima) is a third, every-30-steps pass with a deterministic
// two-level reduction.
//
// Two variants of the node passes:
//   exact  thread per slot, SELL-64 horizons, per-bond arithmetic in the reference's operation
//          order (src/main.cpp:736-874) with floating-point contraction off: the results equal a
//          CPU build without FMA bit for bit, given the same neighbour order;
//   tile   (fast path, lattice inputs) one block per 16 x 16 tile: the block stages its 24 x 24
//          halo window of the fields in LDS with coalesced loads, then every bond is an LDS read
//          at a fixed window offset; the bond geometry (unit vector, |xi| beta dV, beta dV) is a
//          wave-uniform table entry (scalar loads). No index stream, no square root or division
//          per bond, FMA on.
#include "lps_internal.h"

namespace lps {

// ---------------------------------------------------------------------------------------------
// Scalar pieces of src/main.cpp, evaluated in the reference's order.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double beta_exact(double xi, double R, double dx) {
#pragma clang fp contract(off)
  // src/main.cpp:100-107
  if (xi <= R - dx / 2.0) return 1.0;
  else if (xi <= R + dx / 2.0) return (R + dx / 2.0 - xi) / dx;
  else return 0.0;
}

// One bond of the dilatation sum: w |xi| e beta dx dy (src/main.cpp:743-757).
__device__ __forceinline__ double theta_term_exact(double xi_x, double xi_y, double eta_x,
                                                   double eta_y, double R, double dx, double dy) {
#pragma clang fp contract(off)
  const double xi_norm = sqrt(xi_x * xi_x + xi_y * xi_y);
  const double e = (eta_x * xi_x + eta_y * xi_y) / xi_norm;
  const double w = 1.0;
  const double b = beta_exact(xi_norm, R, dx);
  return w * xi_norm * e * b * dx * dy;
}

// One bond of the force sum (src/main.cpp:837-869); returns false when the bond is skipped.
__device__ __forceinline__ bool force_term_exact(double xi_x, double xi_y, double eta_x,
                                                 double eta_y, double m_i, double m_j,
                                                 double th_i, double th_j, double lam, double mu,
                                                 double R, double dx, double dy, double *fx,
                                                 double *fy) {
#pragma clang fp contract(off)
  const double xi_norm = sqrt(xi_x * xi_x + xi_y * xi_y);
  const double e = (eta_x * xi_x + eta_y * xi_y) / xi_norm;
  const double w = 1.0;
  const double b = beta_exact(xi_norm, R, dx);
  if (m_j <= 0.0) return false;
  const double t_i = (2.0 * (lam - mu) / m_i) * th_i * xi_norm * w + (8.0 * mu / m_i) * w * e;
  const double t_j = (2.0 * (lam - mu) / m_j) * th_j * xi_norm * w + (8.0 * mu / m_j) * w * e;
  const double scale = (t_i + t_j) * b * dx * dy;
  *fx = scale * xi_x / xi_norm;
  *fy = scale * xi_y / xi_norm;
  return true;
}

__device__ __forceinline__ double add_exact(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}

// Velocity-Verlet corrector and next predictor for one component (src/main.cpp:924, :727).
__device__ __forceinline__ void verlet_exact(double u1, double v, double a, double a1, double dt,
                                             double *v1, double *p1) {
#pragma clang fp contract(off)
  *v1 = v + (dt / 2.0) * a + (dt / 2.0) * a1;
  *p1 = u1 + dt * (*v1 + (dt / 2.0) * a1);
}

__device__ __forceinline__ double predict_exact(double u, double v, double a, double dt) {
#pragma clang fp contract(off)
  return u + dt * (v + (dt / 2.0) * a);  // src/main.cpp:727
}

// G_u / G_v term of one cloud slot: g0 U_x + g1 U_y + g2 Unm1_x + g3 Unm1_y (src/main.cpp:903-912).
__device__ __forceinline__ double cloud_term_exact(const double *g, double a, double b, double c,
                                                   double d) {
#pragma clang fp contract(off)
  return g[0] * a + g[1] * b + g[2] * c + g[3] * d;
}

__device__ __forceinline__ double gtheta_term_exact(double g0, double g1, double a, double b) {
#pragma clang fp contract(off)
  return g0 * a + g1 * b;  // src/main.cpp:805-806
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------------------------
// Horizon rows
// ---------------------------------------------------------------------------------------------
struct SellRow {
  const int32_t *p;  // first entry of the row
  int32_t w;         // slice width
  __device__ __forceinline__ int32_t operator[](int k) const { return p[(int64_t)k * kSlice]; }
};

__device__ __forceinline__ SellRow sell_row(const StepArgs &a, int32_t i) {
  const int32_t s = i / kSlice;
  SellRow r;
  r.p = a.sell_col + a.sell_off[s] + (i % kSlice);
  r.w = a.sell_w[s];
  return r;
}

// ---------------------------------------------------------------------------------------------
// Predictor (first step only; later steps get the prediction from the force pass's epilogue)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_predict(int32_t n, double dt, const double *ux,
                                                    const double *uy, const double *vx,
                                                    const double *vy, const double *ax,
                                                    const double *ay, double *px, double *py) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  px[i] = predict_exact(ux[i], vx[i], ax[i], dt);
  py[i] = predict_exact(uy[i], vy[i], ay[i], dt);
}

// Dilatation of one node from a field, sequential over the sorted horizon (src/main.cpp:736-761).
__device__ __forceinline__ double theta_node_exact(const StepArgs &a, const double *fx,
                                                   const double *fy, int32_t i) {
  const double xi0 = a.x[i], yi0 = a.y[i], ui = fx[i], vi = fy[i];
  double sum = 0.0;
  const SellRow r = sell_row(a, i);
  for (int k = 0; k < r.w; k++) {
    const int32_t j = r[k];
    if (j < 0) break;
    sum = add_exact(sum, theta_term_exact(a.x[j] - xi0, a.y[j] - yi0, fx[j] - ui, fy[j] - vi,
                                          a.delta, a.dx, a.dy));
  }
  const double m = a.m[i];
  double th;
  {
#pragma clang fp contract(off)
    th = (m > 0.0) ? (2.0 / m) * sum : 0.0;
  }
  return th;
}

// theta of a whole field (exact order): the theta-history initialisation (src/main.cpp:664-682).
__global__ void __launch_bounds__(kBlock) k_theta_field(StepArgs a, const double *fx,
                                                        const double *fy, double *theta) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  theta[i] = a.kind[i] ? theta_node_exact(a, fx, fy, i) : 0.0;
}

// Weighted volume m_i = sum w |xi|^2 beta dx dy over the CSR row (src/main.cpp:412-426).
__global__ void __launch_bounds__(kBlock) k_weighted_volume(StepArgs a, double *m) {
#pragma clang fp contract(off)
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  double m_i = 0.0;
  const double xi0 = a.x[i], yi0 = a.y[i];
  for (int64_t q = a.row_ptr[i] + 1; q < a.row_ptr[i + 1]; q++) {
    const int32_t j = a.col[q];
    const double xi_x = a.x[j] - xi0;
    const double xi_y = a.y[j] - yi0;
    const double xi_norm = sqrt(xi_x * xi_x + xi_y * xi_y);
    const double w = 1.0;
    const double b = beta_exact(xi_norm, a.delta, a.dx);
    m_i += w * xi_norm * xi_norm * b * a.dx * a.dy;
  }
  m[i] = m_i;
}

// ---------------------------------------------------------------------------------------------
// ABC rows, one wavefront per row. EXACT sums in the reference's order on lane 0 (per-term values
// staged in LDS); otherwise wavefront tree sums.
// ---------------------------------------------------------------------------------------------
struct AbcScratch {
  double f[2][kMaxHorizon];
  double c[4][64];
  int8_t flag[kMaxHorizon];
};

// G_theta at one ABC row: theta_a = sum_k G[2k] ThN[c_k] + G[2k+1] ThNm1[c_k] (src/main.cpp:795-812).
template <bool EXACT>
__device__ __forceinline__ double abc_gtheta(const StepArgs &a, const StepState &st, int32_t row,
                                             int32_t lane, AbcScratch *sw) {
  const double *g = a.gth + (int64_t)row * 2 * a.M;
  const int32_t *cl = a.cloud + (int64_t)row * a.M;
  double acc = 0.0;
  for (int base = 0; base < a.M; base += 64) {
    const int jj = base + lane;
    double t = 0.0;
    int32_t k = -1;
    if (jj < a.M) {
      k = cl[jj];
      if (k >= 0) t = gtheta_term_exact(g[2 * jj], g[2 * jj + 1], st.thn[k], st.thnm1[k]);
    }
    if (EXACT) {
      sw->c[0][lane] = t;
      sw->flag[lane] = (k >= 0);
      __threadfence_block();
      if (lane == 0) {
        const int lim = min(64, a.M - base);
        for (int q = 0; q < lim; q++)
          if (sw->flag[q]) acc = add_exact(acc, sw->c[0][q]);
      }
      __threadfence_block();
    } else {
      acc += wave_sum(t);
    }
  }
  return acc;  // complete on lane 0 (EXACT) or on every lane
}

// Force pass of one ABC row: force on the node, G_u / G_v, next prediction.
template <bool EXACT>
__device__ __forceinline__ void abc_force_row(const StepArgs &a, const StepState &st, int32_t row,
                                              int32_t lane, AbcScratch *sw) {
  const int32_t node = a.abc_slot[row];
  const double dt = a.dt;
  // force on the ABC node (src/main.cpp:821-874): bonds in parallel
  double F0 = 0.0, F1 = 0.0;
  const double m_i = a.m[node];
  {
    const int64_t q0 = a.row_ptr[node] + 1;
    const int32_t nb = (int32_t)(a.row_ptr[node + 1] - q0);
    const double pxi = st.px[node], pyi = st.py[node];
    const double xi0 = a.x[node], yi0 = a.y[node], th_i = st.theta[node];
    for (int k = lane; k < nb; k += 64) {
      const int32_t j = a.col[q0 + k];
      double fx = 0.0, fy = 0.0;
      bool ok = false;
      if (m_i > 0.0)
        ok = force_term_exact(a.x[j] - xi0, a.y[j] - yi0, st.px[j] - pxi, st.py[j] - pyi, m_i,
                              a.m[j], th_i, st.theta[j], a.lam, a.mu, a.delta, a.dx, a.dy, &fx,
                              &fy);
      if (EXACT) {
        sw->f[0][k] = fx;
        sw->f[1][k] = fy;
        sw->flag[k] = ok;
      } else if (ok) {
        F0 += fx;
        F1 += fy;
      }
    }
    if (EXACT) {
      __threadfence_block();
      if (lane == 0 && m_i > 0.0) {
        for (int k = 0; k < nb; k++) {
          if (!sw->flag[k]) continue;
          F0 = add_exact(F0, sw->f[0][k]);
          F1 = add_exact(F1, sw->f[1][k]);
        }
      }
      __threadfence_block();
    } else {
      F0 = wave_sum(F0);
      F1 = wave_sum(F1);
    }
  }
  // G_u and G_v over the cloud (src/main.cpp:883-916, 933-974)
  const double *gu0 = a.gu + (int64_t)row * 8 * a.M, *gu1 = gu0 + 4 * a.M;
  const double *gv0 = a.gv + (int64_t)row * 8 * a.M, *gv1 = gv0 + 4 * a.M;
  const int32_t *cl = a.cloud + (int64_t)row * a.M;
  double Ux = 0.0, Uy = 0.0, Vx = 0.0, Vy = 0.0;
  for (int base = 0; base < a.M; base += 64) {
    const int jj = base + lane;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
    int32_t k = -1;
    if (jj < a.M) {
      k = cl[jj];
      if (k >= 0) {
        const double ux = st.ux[k], uy = st.uy[k], uxm = st.uxm[k], uym = st.uym[k];
        const double vx = st.vx[k], vy = st.vy[k], vxm = st.vxm[k], vym = st.vym[k];
        t0 = cloud_term_exact(gu0 + 4 * jj, ux, uy, uxm, uym);
        t1 = cloud_term_exact(gu1 + 4 * jj, ux, uy, uxm, uym);
        t2 = cloud_term_exact(gv0 + 4 * jj, vx, vy, vxm, vym);
        t3 = cloud_term_exact(gv1 + 4 * jj, vx, vy, vxm, vym);
      }
    }
    if (EXACT) {
      sw->c[0][lane] = t0;
      sw->c[1][lane] = t1;
      sw->c[2][lane] = t2;
      sw->c[3][lane] = t3;
      sw->flag[lane] = (k >= 0);
      __threadfence_block();
      if (lane == 0) {
        const int lim = min(64, a.M - base);
        for (int q = 0; q < lim; q++) {
          if (!sw->flag[q]) continue;
          Ux = add_exact(Ux, sw->c[0][q]);
          Uy = add_exact(Uy, sw->c[1][q]);
          Vx = add_exact(Vx, sw->c[2][q]);
          Vy = add_exact(Vy, sw->c[3][q]);
        }
      }
      __threadfence_block();
    } else {
      Ux += wave_sum(t0);
      Uy += wave_sum(t1);
      Vx += wave_sum(t2);
      Vy += wave_sum(t3);
    }
  }
  if (lane == 0) {
    double a1x, a1y;
    if (m_i <= 0.0) {
      a1x = 0.0;
      a1y = 0.0;
    } else {
#pragma clang fp contract(off)
      a1x = F0 / a.rho;
      a1y = F1 / a.rho;
    }
    st.ax1[node] = a1x; st.ay1[node] = a1y;
    st.ux1[node] = Ux; st.uy1[node] = Uy;
    st.vx1[node] = Vx; st.vy1[node] = Vy;
    st.pnx[node] = predict_exact(Ux, Vx, a1x, dt);
    st.pny[node] = predict_exact(Uy, Vy, a1y, dt);
  }
}

// Dilatation pass of one ABC row: G_theta, or for IDW the direct value (overridden by k_idw where a
// non-ABC neighbour exists, src/main.cpp:787-789).
template <bool EXACT>
__device__ __forceinline__ void abc_theta_row(const StepArgs &a, const StepState &st, int32_t row,
                                              int32_t lane, AbcScratch *sw) {
  const int32_t node = a.abc_slot[row];
  double th;
  if (a.theta_method == 1) th = abc_gtheta<EXACT>(a, st, row, lane, sw);
  else th = (lane == 0) ? theta_node_exact(a, st.px, st.py, node) : 0.0;
  if (lane == 0) st.theta[node] = th;
}

// ---------------------------------------------------------------------------------------------
// Exact node passes: thread per slot (kind 1), SELL rows
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void force_node_exact(const StepArgs &a, const StepState &st, int32_t i,
                                                 double *fx_out, double *fy_out) {
  const double m_i = a.m[i];
  if (m_i <= 0.0) { *fx_out = 0.0; *fy_out = 0.0; return; }
  const double pxi = st.px[i], pyi = st.py[i];
  const SellRow r = sell_row(a, i);
  const double xi0 = a.x[i], yi0 = a.y[i], th_i = st.theta[i];
  double F0 = 0.0, F1 = 0.0;
  for (int k = 0; k < r.w; k++) {
    const int32_t j = r[k];
    if (j < 0) break;
    double fx, fy;
    if (force_term_exact(a.x[j] - xi0, a.y[j] - yi0, st.px[j] - pxi, st.py[j] - pyi, m_i, a.m[j],
                         th_i, st.theta[j], a.lam, a.mu, a.delta, a.dx, a.dy, &fx, &fy)) {
      F0 = add_exact(F0, fx);
      F1 = add_exact(F1, fy);
    }
  }
  {
#pragma clang fp contract(off)
    *fx_out = F0 / a.rho;
    *fy_out = F1 / a.rho;
  }
}

__global__ void __launch_bounds__(kBlock) k_theta_exact(StepArgs a, StepState st,
                                                        int32_t n_node_blocks) {
  __shared__ AbcScratch sabc[kWavesPerBlock];
  const int32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if ((int32_t)blockIdx.x < n_node_blocks) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n || a.kind[i] != 1) return;
    st.theta[i] = theta_node_exact(a, st.px, st.py, i);
    return;
  }
  const int32_t row = (blockIdx.x - n_node_blocks) * kWavesPerBlock + wid;
  if (row < a.n_abc) abc_theta_row<true>(a, st, row, lane, &sabc[wid]);
}

__global__ void __launch_bounds__(kBlock) k_force_exact(StepArgs a, StepState st,
                                                        int32_t n_node_blocks) {
  __shared__ AbcScratch sabc[kWavesPerBlock];
  const int32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if ((int32_t)blockIdx.x < n_node_blocks) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n || a.kind[i] != 1) return;
    double a1x, a1y, v1x, v1y, p1x, p1y;
    force_node_exact(a, st, i, &a1x, &a1y);
    const double u1x = st.px[i], u1y = st.py[i];
    verlet_exact(u1x, st.vx[i], st.ax[i], a1x, a.dt, &v1x, &p1x);
    verlet_exact(u1y, st.vy[i], st.ay[i], a1y, a.dt, &v1y, &p1y);
    st.ax1[i] = a1x; st.ay1[i] = a1y;
    st.ux1[i] = u1x; st.uy1[i] = u1y;
    st.vx1[i] = v1x; st.vy1[i] = v1y;
    st.pnx[i] = p1x; st.pny[i] = p1y;
    return;
  }
  const int32_t row = (blockIdx.x - n_node_blocks) * kWavesPerBlock + wid;
  if (row < a.n_abc) abc_force_row<true>(a, st, row, lane, &sabc[wid]);
}

// IDW dilatation at ABC rows (src/main.cpp:769-790), one wavefront per row: per-bond weights in
// parallel, sums in horizon order by lane 0.
__global__ void __launch_bounds__(kBlock) k_idw(StepArgs a, StepState st) {
  __shared__ double swt[kWavesPerBlock][kMaxHorizon];
  __shared__ double sval[kWavesPerBlock][kMaxHorizon];
  const int32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int32_t row = blockIdx.x * kWavesPerBlock + wid;
  if (row >= a.n_abc) return;
  const int32_t node = a.abc_slot[row];
  const int64_t q0 = a.row_ptr[node] + 1, q1 = a.row_ptr[node + 1];
  const int nb = (int)(q1 - q0);
  for (int k = lane; k < nb; k += 64) {
    const int32_t j = a.col[q0 + k];
    double wt = -1.0, v = 0.0;
    if (!a.is_bc[j]) {
#pragma clang fp contract(off)
      const double rx = a.x[j] - a.x[node];
      const double ry = a.y[j] - a.y[node];
      const double dist = sqrt(rx * rx + ry * ry);
      wt = 1.0 / (dist + 1e-30);
      v = wt * st.theta[j];
    }
    swt[wid][k] = wt;
    sval[wid][k] = v;
  }
  __threadfence_block();
  if (lane == 0) {
    double theta_avg = 0.0, w_sum = 0.0;
    for (int k = 0; k < nb; k++) {
      if (swt[wid][k] < 0.0) continue;
      theta_avg = add_exact(theta_avg, sval[wid][k]);
      w_sum = add_exact(w_sum, swt[wid][k]);
    }
    if (w_sum > 0.0) {
#pragma clang fp contract(off)
      st.theta[node] = theta_avg / w_sum;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Tile passes (fast path)
// ---------------------------------------------------------------------------------------------
// Blocks are spread over the 8 XCDs round-robin (block b runs on XCD b % 8). Tile t is given to
// block b with t = (b % 8) * per_xcd + b / 8, so each XCD sweeps a contiguous band of tile rows and
// a tile's halo neighbours sit in the same XCD's L2.
__device__ __forceinline__ int32_t tile_of_block(int32_t b, int32_t n_tiles) {
  const int32_t per_xcd = (n_tiles + 7) / 8;
  return (b % 8) * per_xcd + b / 8;
}

// Slot of lattice point (gx, gy), or -1 outside the tiled area / in an absent tile.
__device__ __forceinline__ int32_t lattice_slot(const TileArgs &t, int32_t gx, int32_t gy) {
  if (gx < 0 || gy < 0) return -1;
  const int32_t tx = gx / kTile, ty = gy / kTile;
  if (tx >= t.ntx || ty >= t.nty) return -1;
  const int32_t tt = t.tile_map[tx * t.nty + ty];
  if (tt < 0) return -1;
  return tt * (kTile * kTile) + (gx % kTile) * kTile + (gy % kTile);
}

union TileShared {
  struct {
    double2 u[kMaxWin * kMaxWin];
    int8_t ok[kMaxWin * kMaxWin];
  } th;
  struct {
    double2 u[kMaxWin * kMaxWin];
    double2 ab[kMaxWin * kMaxWin];  // (2 (lambda-mu) theta / m, 8 mu / m); ab.y == 0: no bond
  } fo;
  AbcScratch abc[kWavesPerBlock];
};

__global__ void __launch_bounds__(kBlock) k_theta_tile(StepArgs a, StepState st,
                                                       int32_t n_tile_blocks) {
  __shared__ TileShared sh;
  const TileArgs &T = a.tile;
  if ((int32_t)blockIdx.x >= n_tile_blocks) {  // ABC rows
    const int32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int32_t row = (blockIdx.x - n_tile_blocks) * kWavesPerBlock + wid;
    if (row < a.n_abc) abc_theta_row<false>(a, st, row, lane, &sh.abc[wid]);
    return;
  }
  const int32_t t = tile_of_block(blockIdx.x, T.n_tiles);
  if (t >= T.n_tiles) return;
  const int32_t txy = T.tile_xy[t];
  const int32_t gx0 = (txy >> 16) * kTile - T.halo, gy0 = (txy & 0xffff) * kTile - T.halo;
  const int32_t W = T.win;
  for (int32_t c = threadIdx.x; c < W * W; c += kBlock) {
    const int32_t s = lattice_slot(T, gx0 + c / W, gy0 + c % W);
    const bool ok = s >= 0 && a.kind[s] != 0;
    sh.th.u[c] = ok ? make_double2(st.px[s], st.py[s]) : make_double2(0.0, 0.0);
    sh.th.ok[c] = ok;
  }
  __syncthreads();
  const int32_t i = t * (kTile * kTile) + threadIdx.x;
  if (a.kind[i] != 1) return;
  const int32_t c = (threadIdx.x / kTile + T.halo) * W + (threadIdx.x % kTile + T.halo);
  const double2 ui = sh.th.u[c];
  double sum = 0.0;
  for (int32_t k = 0; k < T.nt; k++) {
    const int32_t cj = c + T.delta[k];
    const double hx = T.hx[k], hy = T.hy[k], g1 = T.g1[k];
    const double2 uj = sh.th.u[cj];
    const double term = g1 * ((uj.x - ui.x) * hx + (uj.y - ui.y) * hy);
    sum += sh.th.ok[cj] ? term : 0.0;
  }
  st.theta[i] = a.c2m[i] * sum;
}

// ---------------------------------------------------------------------------------------------
// Unrolled tile passes for the reference stencil (68 bonds, 0 < ox^2 + oy^2 <= 20, dx == dy):
// every LDS address is the thread's cell plus a compile-time offset, the unit vector is
// (ox, oy) c with integer ox, oy, and the per-class constants sit in scalar registers.
// Tiles whose whole window holds nodes with m > 0 ("full") run without masks.
// ---------------------------------------------------------------------------------------------
struct Off68 {
  int ox, oy, cls;
};
struct Stencil68 {
  Off68 o[68];
};

constexpr int stencil_class(int r2) {
  for (int i = 0; i < 12; i++)
    if (kStencilR2[i] == r2) return i;
  return -1;
}

constexpr Stencil68 make_stencil68() {
  Stencil68 s{};
  int n = 0;
  for (int ox = -4; ox <= 4; ox++)
    for (int oy = -4; oy <= 4; oy++) {
      const int r2 = ox * ox + oy * oy;
      if (r2 > 0 && r2 <= 20) s.o[n++] = Off68{ox, oy, stencil_class(r2)};
    }
  return s;
}

constexpr int kW68 = kTile + 8;          // window edge (halo 4)
constexpr int kCells68 = kW68 * kW68;    // 576

// (dx, dy) . (ox, oy) with the zero products dropped at compile time.
__device__ __forceinline__ double dot_int(double dx, double dy, int ox, int oy) {
  if (ox == 0) return dy * oy;
  if (oy == 0) return dx * ox;
  return dx * ox + dy * oy;
}

// Slot of window cell c of tile t (halo 4) through the 3 x 3 neighbour table, -1 if absent.
__device__ __forceinline__ int32_t window_slot68(const int32_t *nbr, int32_t c) {
  const int32_t rx = c / kW68 - 4, ry = c % kW68 - 4;
  const int32_t tx = rx < 0 ? 0 : (rx < kTile ? 1 : 2);
  const int32_t ty = ry < 0 ? 0 : (ry < kTile ? 1 : 2);
  const int32_t tt = nbr[tx * 3 + ty];
  if (tt < 0) return -1;
  return tt * (kTile * kTile) + (rx - (tx - 1) * kTile) * kTile + (ry - (ty - 1) * kTile);
}

union Tile68Shared {
  struct {
    double2 u[kCells68];
    int8_t ok[kCells68];
  } th;
  struct {
    double2 u[kCells68];
    double2 ab[kCells68];
  } fo;
  AbcScratch abc[kWavesPerBlock];
};

__global__ void __launch_bounds__(kBlock) k_theta_t68(StepArgs a, StepState st,
                                                      int32_t n_tile_blocks) {
  __shared__ Tile68Shared sh;
  const TileArgs &T = a.tile;
  if ((int32_t)blockIdx.x >= n_tile_blocks) {  // ABC rows
    const int32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int32_t row = (blockIdx.x - n_tile_blocks) * kWavesPerBlock + wid;
    if (row < a.n_abc) abc_theta_row<false>(a, st, row, lane, &sh.abc[wid]);
    return;
  }
  const int32_t t = tile_of_block(blockIdx.x, T.n_tiles);
  if (t >= T.n_tiles) return;
  const int32_t *nbr = T.tile_nbr + 9 * t;
  const bool full = T.tile_full[t] != 0;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int32_t c = threadIdx.x + q * kBlock;
    if (c < kCells68) {
      const int32_t s = window_slot68(nbr, c);
      const bool ok = s >= 0 && (full || a.kind[s] != 0);
      sh.th.u[c] = ok ? make_double2(st.px[s], st.py[s]) : make_double2(0.0, 0.0);
      sh.th.ok[c] = ok;
    }
  }
  __syncthreads();
  const int32_t i = t * (kTile * kTile) + threadIdx.x;
  if (a.kind[i] != 1) return;
  const int32_t c = (threadIdx.x / kTile + 4) * kW68 + (threadIdx.x % kTile + 4);
  const double2 ui = sh.th.u[c];
  constexpr Stencil68 S = make_stencil68();
  double acc[4] = {0.0, 0.0, 0.0, 0.0};  // bond k feeds chain k % 4
  if (full) {
#pragma unroll
    for (int k = 0; k < 68; k++) {
      const Off68 o = S.o[k];
      const double2 uj = sh.th.u[c + o.ox * kW68 + o.oy];
      acc[k % 4] += T.c1[o.cls] * dot_int(uj.x - ui.x, uj.y - ui.y, o.ox, o.oy);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 68; k++) {
      const Off68 o = S.o[k];
      const int32_t cj = c + o.ox * kW68 + o.oy;
      const double2 uj = sh.th.u[cj];
      const double term = T.c1[o.cls] * dot_int(uj.x - ui.x, uj.y - ui.y, o.ox, o.oy);
      acc[k % 4] += sh.th.ok[cj] ? term : 0.0;
    }
  }
  st.theta[i] = a.c2m[i] * ((acc[0] + acc[1]) + (acc[2] + acc[3]));
}

template <bool FULL>
__device__ __forceinline__ void force_sum68(const double2 *su, const double2 *sab, int32_t c,
                                            double2 ui, double2 abi, const TileArgs &T,
                                            double *F0, double *F1) {
  constexpr Stencil68 S = make_stencil68();
  // four independent accumulation chains (bond k feeds chain k % 4): short dependence chains keep
  // few bond terms live
  double fx[4] = {0.0, 0.0, 0.0, 0.0}, fy[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 68; k++) {
    const Off68 o = S.o[k];
    const int32_t cj = c + o.ox * kW68 + o.oy;
    const double2 uj = su[cj];
    const double2 abj = sab[cj];
    const double tt = dot_int(uj.x - ui.x, uj.y - ui.y, o.ox, o.oy);
    double sp = (abi.x + abj.x) * T.c1[o.cls] + (abi.y + abj.y) * tt * T.c2[o.cls];
    if (!FULL) sp = (abj.y != 0.0) ? sp : 0.0;
    if (o.ox != 0) fx[k % 4] += sp * o.ox;
    if (o.oy != 0) fy[k % 4] += sp * o.oy;
  }
  *F0 = (fx[0] + fx[1]) + (fx[2] + fx[3]);
  *F1 = (fy[0] + fy[1]) + (fy[2] + fy[3]);
}

// Unrolled force kernel for the full tiles (no masks: every window cell holds a node with m > 0).
// Five waves per SIMD fit without spills.
__global__ void __launch_bounds__(kBlock, 5) k_force_t68(StepArgs a, StepState st) {
  __shared__ Tile68Shared sh;
  const TileArgs &T = a.tile;
  const int32_t q = tile_of_block(blockIdx.x, T.n_full);
  if (q >= T.n_full) return;
  const int32_t t = T.full_list[q];
  const int32_t *nbr = T.tile_nbr + 9 * t;
  const double lm = a.lam - a.mu, mu4 = 4.0 * a.mu;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const int32_t c = threadIdx.x + r * kBlock;
    if (c < kCells68) {
      const int32_t s = window_slot68(nbr, c);
      const double c2 = a.c2m[s];
      sh.fo.u[c] = make_double2(st.px[s], st.py[s]);
      sh.fo.ab[c] = make_double2(lm * c2 * st.theta[s], mu4 * c2);
    }
  }
  __syncthreads();
  const int32_t i = t * (kTile * kTile) + threadIdx.x;
  if (a.kind[i] != 1) return;
  const int32_t c = (threadIdx.x / kTile + 4) * kW68 + (threadIdx.x % kTile + 4);
  const double2 ui = sh.fo.u[c];
  const double2 abi = sh.fo.ab[c];
  double F0 = 0.0, F1 = 0.0;
  force_sum68<true>(sh.fo.u, sh.fo.ab, c, ui, abi, T, &F0, &F1);
  const double inv_rho = 1.0 / a.rho;
  const double a1x = F0 * inv_rho, a1y = F1 * inv_rho;
  const double h = 0.5 * a.dt;
  const double v1x = st.vx[i] + h * (st.ax[i] + a1x);
  const double v1y = st.vy[i] + h * (st.ay[i] + a1y);
  st.ax1[i] = a1x; st.ay1[i] = a1y;
  st.ux1[i] = ui.x; st.uy1[i] = ui.y;
  st.vx1[i] = v1x; st.vy1[i] = v1y;
  st.pnx[i] = ui.x + a.dt * (v1x + h * a1x);
  st.pny[i] = ui.y + a.dt * (v1y + h * a1y);
}

// Force pass of the ABC rows alone (the tile kernels leave kind-2 slots to it).
template <bool EXACT>
__global__ void __launch_bounds__(kBlock) k_force_abc(StepArgs a, StepState st) {
  __shared__ AbcScratch sabc[kWavesPerBlock];
  const int32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int32_t row = blockIdx.x * kWavesPerBlock + wid;
  if (row < a.n_abc) abc_force_row<EXACT>(a, st, row, lane, &sabc[wid]);
}

// Generic force kernel over a list of tiles (all tiles when list is null).
__global__ void __launch_bounds__(kBlock) k_force_tile(StepArgs a, StepState st,
                                                       const int32_t *list, int32_t count) {
  __shared__ TileShared sh;
  const TileArgs &T = a.tile;
  const int32_t q = tile_of_block(blockIdx.x, count);
  if (q >= count) return;
  const int32_t t = list ? list[q] : q;
  const int32_t txy = T.tile_xy[t];
  const int32_t gx0 = (txy >> 16) * kTile - T.halo, gy0 = (txy & 0xffff) * kTile - T.halo;
  const int32_t W = T.win;
  const double lm = a.lam - a.mu, mu4 = 4.0 * a.mu;
  for (int32_t c = threadIdx.x; c < W * W; c += kBlock) {
    const int32_t s = lattice_slot(T, gx0 + c / W, gy0 + c % W);
    double2 u = make_double2(0.0, 0.0), ab = make_double2(0.0, 0.0);
    if (s >= 0 && a.kind[s] != 0) {
      const double c2 = a.c2m[s];  // 0 when m <= 0: the bond is skipped (src/main.cpp:852)
      u = make_double2(st.px[s], st.py[s]);
      ab = make_double2(lm * c2 * st.theta[s], mu4 * c2);
    }
    sh.fo.u[c] = u;
    sh.fo.ab[c] = ab;
  }
  __syncthreads();
  const int32_t i = t * (kTile * kTile) + threadIdx.x;
  if (a.kind[i] != 1) return;
  const int32_t c = (threadIdx.x / kTile + T.halo) * W + (threadIdx.x % kTile + T.halo);
  const double2 ui = sh.fo.u[c];
  const double2 abi = sh.fo.ab[c];
  double F0 = 0.0, F1 = 0.0;
  if (abi.y != 0.0) {
    // branch-free body: the per-bond geometry stays wave-uniform (scalar loads)
    for (int32_t k = 0; k < T.nt; k++) {
      const int32_t cj = c + T.delta[k];
      const double hx = T.hx[k], hy = T.hy[k], g1 = T.g1[k], g2 = T.g2[k];
      const double2 abj = sh.fo.ab[cj];
      const double2 uj = sh.fo.u[cj];
      const double e = (uj.x - ui.x) * hx + (uj.y - ui.y) * hy;
      double s = (abi.x + abj.x) * g1 + (abi.y + abj.y) * e * g2;
      s = (abj.y != 0.0) ? s : 0.0;
      F0 += s * hx;
      F1 += s * hy;
    }
  }
  const double inv_rho = 1.0 / a.rho;
  const double a1x = F0 * inv_rho, a1y = F1 * inv_rho;
  const double h = 0.5 * a.dt;
  const double u1x = ui.x, u1y = ui.y;
  const double v1x = st.vx[i] + h * (st.ax[i] + a1x);
  const double v1y = st.vy[i] + h * (st.ay[i] + a1y);
  st.ax1[i] = a1x; st.ay1[i] = a1y;
  st.ux1[i] = u1x; st.uy1[i] = u1y;
  st.vx1[i] = v1x; st.vy1[i] = v1y;
  st.pnx[i] = u1x + a.dt * (v1x + h * a1x);
  st.pny[i] = u1y + a.dt * (v1y + h * a1y);
}

// For every node slot, the stencil cells holding a node must be exactly the set of its CSR row
// (self excluded) for the tile path to reproduce the reference's horizons.
__global__ void __launch_bounds__(kBlock) k_tile_check(StepArgs a, int32_t *mismatch) {
  const TileArgs &T = a.tile;
  const int32_t t = blockIdx.x;
  const int32_t i = t * (kTile * kTile) + threadIdx.x;
  if (a.kind[i] == 0) return;
  const int32_t txy = T.tile_xy[t];
  const int32_t gx = (txy >> 16) * kTile + threadIdx.x / kTile;
  const int32_t gy = (txy & 0xffff) * kTile + threadIdx.x % kTile;
  const int64_t q0 = a.row_ptr[i] + 1, q1 = a.row_ptr[i + 1];
  int64_t found = 0;
  bool bad = false;
  for (int32_t k = 0; k < T.nt && !bad; k++) {
    // delta = ox * win + oy with |ox|, |oy| <= halo < win / 2
    const int32_t d = T.delta[k];
    const int32_t ox = (d + T.halo + T.halo * T.win) / T.win - T.halo;
    const int32_t oy = d - ox * T.win;
    const int32_t s = lattice_slot(T, gx + ox, gy + oy);
    if (s < 0 || a.kind[s] == 0) continue;
    bool in_row = false;
    for (int64_t q = q0; q < q1 && !in_row; q++) in_row = (a.col[q] == s);
    bad = !in_row;
    found++;
  }
  if (bad || found != q1 - q0) atomicAdd(mismatch, 1);
}

// ---------------------------------------------------------------------------------------------
// Output block (src/main.cpp:986-1062): per-node terms in the reference's order, block partials,
// final reduction in a fixed order.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void diag_merge(DiagPartial &d, const DiagPartial &o) {
  d.energy += o.energy; d.ek += o.ek; d.ep += o.ep; d.px += o.px; d.py += o.py;
  d.ub += o.ub; d.vb += o.vb;
  if (o.max_u > d.max_u || (o.max_u == d.max_u && o.max_u_orig < d.max_u_orig)) {
    d.max_u = o.max_u; d.max_u_orig = o.max_u_orig;
  }
  if (o.max_v > d.max_v || (o.max_v == d.max_v && o.max_v_orig < d.max_v_orig)) {
    d.max_v = o.max_v; d.max_v_orig = o.max_v_orig;
  }
  d.max_a = (d.max_a < o.max_a) ? o.max_a : d.max_a;
  d.th_max = (d.th_max < o.th_max) ? o.th_max : d.th_max;
  d.th_min = (o.th_min < d.th_min) ? o.th_min : d.th_min;
}

__device__ __forceinline__ void diag_identity(DiagPartial &d) {
  d.energy = d.ek = d.ep = d.px = d.py = d.ub = d.vb = 0.0;
  d.max_u = 0.0; d.max_v = 0.0; d.max_a = 0.0;
  d.th_max = -1e30; d.th_min = 1e30;
  d.max_u_orig = INT64_MAX; d.max_v_orig = INT64_MAX;
}

__device__ void block_reduce_diag(DiagPartial &d, DiagPartial *sh) {
  sh[threadIdx.x] = d;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      DiagPartial t = sh[threadIdx.x];
      diag_merge(t, sh[threadIdx.x + s]);
      sh[threadIdx.x] = t;
    }
    __syncthreads();
  }
  d = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_diag(StepArgs a, const double *ux, const double *uy,
                                                 const double *vx, const double *vy,
                                                 const double *ax, const double *ay,
                                                 const double *theta, const int32_t *perm,
                                                 double exclude_r, int32_t per_block,
                                                 DiagPartial *partials) {
  __shared__ DiagPartial sh[kBlock];
  DiagPartial d;
  diag_identity(d);
  const int32_t base = blockIdx.x * per_block;
  const int32_t end = min(a.n, base + per_block);
  for (int32_t i = base + threadIdx.x; i < end; i += kBlock) {
#pragma clang fp contract(off)
    const int8_t kd = a.kind[i];
    if (kd == 0) continue;
    const double x = a.x[i];
    const double y = a.y[i];
    const double r = sqrt(x * x + y * y);
    const double u_mag = sqrt(ux[i] * ux[i] + uy[i] * uy[i]);
    const double v_mag = sqrt(vx[i] * vx[i] + vy[i] * vy[i]);
    const double a_mag = sqrt(ax[i] * ax[i] + ay[i] * ay[i]);
    const int64_t orig = perm[i];
    if (u_mag > d.max_u || (u_mag == d.max_u && u_mag > 0.0 && orig < d.max_u_orig)) {
      d.max_u = u_mag; d.max_u_orig = orig;
    }
    if (v_mag > d.max_v || (v_mag == d.max_v && v_mag > 0.0 && orig < d.max_v_orig)) {
      d.max_v = v_mag; d.max_v_orig = orig;
    }
    d.max_a = (d.max_a < a_mag) ? a_mag : d.max_a;
    const double th = theta[i];
    d.th_min = (th < d.th_min) ? th : d.th_min;
    d.th_max = (d.th_max < th) ? th : d.th_max;
    if (kd == 2) {  // ABC rows: boundary norms (src/main.cpp:1050-1060)
      d.ub += ux[i] * ux[i] + uy[i] * uy[i];
      d.vb += vx[i] * vx[i] + vy[i] * vy[i];
    }
    if (exclude_r > 0.0 && r > exclude_r) continue;
    d.px += a.rho * vx[i] * a.dx * a.dy;
    d.py += a.rho * vy[i] * a.dx * a.dy;
    const double ek = 0.5 * a.rho * (vx[i] * vx[i] + vy[i] * vy[i]) * a.dx * a.dy;
    d.ek += ek;
    const double m_i = a.m[i];
    double W_i = 0.5 * a.kappa * th * th;
    const SellRow row = sell_row(a, i);
    for (int k = 0; k < row.w; k++) {
      const int32_t n = row[k];
      if (n < 0) break;
      const double xi_x = a.x[n] - x;
      const double xi_y = a.y[n] - y;
      const double xi_norm = sqrt(xi_x * xi_x + xi_y * xi_y);
      const double eta_x = ux[n] - ux[i];
      const double eta_y = uy[n] - uy[i];
      const double e = (eta_x * xi_x + eta_y * xi_y) / xi_norm;
      const double ed = e - th * xi_norm / 2.0;
      const double w = 1.0;
      const double b = beta_exact(xi_norm, a.delta, a.dx);
      W_i += (4.0 * a.mu / m_i) * w * ed * ed * b * a.dx * a.dy;
    }
    const double ep = W_i * a.dx * a.dy;
    d.ep += ep;
    d.energy += ek + ep;
  }
  block_reduce_diag(d, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = d;
}

// Output block on the tile layout: the per-node terms of src/main.cpp:992-1043 with the strain
// energy sum taken over the stencil on the LDS-staged window of U^{n+1}.
__global__ void __launch_bounds__(kBlock) k_diag_tile(StepArgs a, const double *ux,
                                                      const double *uy, const double *vx,
                                                      const double *vy, const double *ax,
                                                      const double *ay, const double *theta,
                                                      const int32_t *perm, double exclude_r,
                                                      DiagPartial *partials) {
  __shared__ DiagPartial sh[kBlock];
  __shared__ double2 su[kMaxWin * kMaxWin];
  __shared__ int8_t sok[kMaxWin * kMaxWin];
  const TileArgs &T = a.tile;
  const int32_t t = blockIdx.x;
  const int32_t txy = T.tile_xy[t];
  const int32_t gx0 = (txy >> 16) * kTile - T.halo, gy0 = (txy & 0xffff) * kTile - T.halo;
  const int32_t W = T.win;
  for (int32_t c = threadIdx.x; c < W * W; c += kBlock) {
    const int32_t s = lattice_slot(T, gx0 + c / W, gy0 + c % W);
    const bool ok = s >= 0 && a.kind[s] != 0;
    su[c] = ok ? make_double2(ux[s], uy[s]) : make_double2(0.0, 0.0);
    sok[c] = ok;
  }
  __syncthreads();
  DiagPartial d;
  diag_identity(d);
  const int32_t i = t * (kTile * kTile) + threadIdx.x;
  const int8_t kd = a.kind[i];
  if (kd != 0) {
    const double x = a.x[i], y = a.y[i];
    const double r = sqrt(x * x + y * y);
    const double uxi = ux[i], uyi = uy[i], vxi = vx[i], vyi = vy[i];
    const double u_mag = sqrt(uxi * uxi + uyi * uyi);
    const double v_mag = sqrt(vxi * vxi + vyi * vyi);
    const double a_mag = sqrt(ax[i] * ax[i] + ay[i] * ay[i]);
    const int64_t orig = perm[i];
    if (u_mag > 0.0) { d.max_u = u_mag; d.max_u_orig = orig; }
    if (v_mag > 0.0) { d.max_v = v_mag; d.max_v_orig = orig; }
    d.max_a = a_mag;
    const double th = theta[i];
    d.th_min = th;
    d.th_max = th;
    if (kd == 2) {
      d.ub = uxi * uxi + uyi * uyi;
      d.vb = vxi * vxi + vyi * vyi;
    }
    if (!(exclude_r > 0.0 && r > exclude_r)) {
      d.px = a.rho * vxi * a.dx * a.dy;
      d.py = a.rho * vyi * a.dx * a.dy;
      const double ek = 0.5 * a.rho * (vxi * vxi + vyi * vyi) * a.dx * a.dy;
      d.ek = ek;
      const int32_t c = (threadIdx.x / kTile + T.halo) * W + (threadIdx.x % kTile + T.halo);
      double acc = 0.0;
      for (int32_t k = 0; k < T.nt; k++) {
        const int32_t cj = c + T.delta[k];
        const double hx = T.hx[k], hy = T.hy[k], g2 = T.g2[k], xin = T.xi[k];
        const double2 uj = su[cj];
        const double ed = (uj.x - uxi) * hx + (uj.y - uyi) * hy - th * xin * 0.5;
        acc += sok[cj] ? ed * ed * g2 : 0.0;
      }
      const double W_i = 0.5 * a.kappa * th * th + (4.0 * a.mu / a.m[i]) * acc;
      const double ep = W_i * a.dx * a.dy;
      d.ep = ep;
      d.energy = ek + ep;
    }
  }
  block_reduce_diag(d, sh);
  if (threadIdx.x == 0) partials[t] = d;
}

__global__ void __launch_bounds__(kBlock) k_diag_final(const DiagPartial *partials, int32_t np,
                                                       DiagPartial *out) {
  __shared__ DiagPartial sh[kBlock];
  DiagPartial d;
  diag_identity(d);
  for (int32_t t = threadIdx.x; t < np; t += kBlock) diag_merge(d, partials[t]);
  block_reduce_diag(d, sh);
  if (threadIdx.x == 0) *out = d;
}

// ---------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------
static inline int32_t blocks_for(int64_t n) { return (int32_t)((n + kBlock - 1) / kBlock); }
static inline int32_t abc_blocks(const StepArgs &a) {
  return (a.n_abc + kWavesPerBlock - 1) / kWavesPerBlock;
}
static inline int32_t tile_blocks(const StepArgs &a) { return (a.tile.n_tiles + 7) / 8 * 8; }

hipError_t launch_predict(const StepArgs &a, const double *ux, const double *uy, const double *vx,
                          const double *vy, const double *ax, const double *ay, double *px,
                          double *py, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_predict, dim3(blocks_for(a.n)), dim3(kBlock), 0, s, a.n, a.dt, ux, uy, vx,
                     vy, ax, ay, px, py);
  return hipGetLastError();
}

hipError_t launch_theta(const StepArgs &a, const StepState &st, bool fast, hipStream_t s) {
  if (fast) {
    const int32_t nt = tile_blocks(a);
    if (a.tile.stencil68 & 1)
      hipLaunchKernelGGL(k_theta_t68, dim3(nt + abc_blocks(a)), dim3(kBlock), 0, s, a, st, nt);
    else
      hipLaunchKernelGGL(k_theta_tile, dim3(nt + abc_blocks(a)), dim3(kBlock), 0, s, a, st, nt);
  } else {
    const int32_t nb = blocks_for(a.n);
    hipLaunchKernelGGL(k_theta_exact, dim3(nb + abc_blocks(a)), dim3(kBlock), 0, s, a, st, nb);
  }
  return hipGetLastError();
}

hipError_t launch_idw(const StepArgs &a, const StepState &st, hipStream_t s) {
  if (abc_blocks(a) == 0) return hipSuccess;
  hipLaunchKernelGGL(k_idw, dim3(abc_blocks(a)), dim3(kBlock), 0, s, a, st);
  return hipGetLastError();
}

static inline int32_t list_blocks(int32_t count) { return (count + 7) / 8 * 8; }

hipError_t launch_force(const StepArgs &a, const StepState &st, bool fast, hipStream_t s) {
  if (fast) {
    if (a.tile.stencil68 & 2) {
      if (a.tile.n_full > 0)
        hipLaunchKernelGGL(k_force_t68, dim3(list_blocks(a.tile.n_full)), dim3(kBlock), 0, s, a,
                           st);
    } else {
      hipLaunchKernelGGL(k_force_tile, dim3(tile_blocks(a)), dim3(kBlock), 0, s, a, st,
                         (const int32_t *)nullptr, a.tile.n_tiles);
    }
  } else {
    const int32_t nb = blocks_for(a.n);
    hipLaunchKernelGGL(k_force_exact, dim3(nb + abc_blocks(a)), dim3(kBlock), 0, s, a, st, nb);
  }
  return hipGetLastError();
}

hipError_t launch_force_abc(const StepArgs &a, const StepState &st, hipStream_t s) {
  if ((a.tile.stencil68 & 2) && a.tile.n_part > 0) {
    hipLaunchKernelGGL(k_force_tile, dim3(list_blocks(a.tile.n_part)), dim3(kBlock), 0, s, a, st,
                       a.tile.part_list, a.tile.n_part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (abc_blocks(a) == 0) return hipSuccess;
  hipLaunchKernelGGL(k_force_abc<false>, dim3(abc_blocks(a)), dim3(kBlock), 0, s, a, st);
  return hipGetLastError();
}

hipError_t launch_tile_check(const StepArgs &a, int32_t *d_mismatch, hipStream_t s) {
  if (a.tile.n_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tile_check, dim3(a.tile.n_tiles), dim3(kBlock), 0, s, a, d_mismatch);
  return hipGetLastError();
}

hipError_t launch_theta_field(const StepArgs &a, const double *ux, const double *uy, double *theta,
                              hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_theta_field, dim3(blocks_for(a.n)), dim3(kBlock), 0, s, a, ux, uy, theta);
  return hipGetLastError();
}

hipError_t launch_weighted_volume(const StepArgs &a, double *m, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_weighted_volume, dim3(blocks_for(a.n)), dim3(kBlock), 0, s, a, m);
  return hipGetLastError();
}

// Partials: one per block of kDiagPerBlock slots.
constexpr int32_t kDiagPerBlock = 4096;

int32_t diag_partials(int32_t n) {
  return std::max<int32_t>(1, (n + kDiagPerBlock - 1) / kDiagPerBlock);
}

hipError_t launch_diag(const StepArgs &a, const double *ux, const double *uy, const double *vx,
                       const double *vy, const double *ax, const double *ay, const double *theta,
                       const int32_t *perm, double exclude_r, DiagPartial *partials,
                       int32_t n_partials, DiagPartial *out, hipStream_t s) {
  hipLaunchKernelGGL(k_diag, dim3(n_partials), dim3(kBlock), 0, s, a, ux, uy, vx, vy, ax, ay,
                     theta, perm, exclude_r, kDiagPerBlock, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kBlock), 0, s, partials, n_partials, out);
  return hipGetLastError();
}

hipError_t launch_diag_tile(const StepArgs &a, const double *ux, const double *uy,
                            const double *vx, const double *vy, const double *ax, const double *ay,
                            const double *theta, const int32_t *perm, double exclude_r,
                            DiagPartial *partials, DiagPartial *out, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_tile, dim3(a.tile.n_tiles), dim3(kBlock), 0, s, a, ux, uy, vx, vy, ax,
                     ay, theta, perm, exclude_r, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kBlock), 0, s, partials, a.tile.n_tiles, out);
  return hipGetLastError();
}

}  // namespace lps
