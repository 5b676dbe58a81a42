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
#include <utility>

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
__device__ __forceinline__ double cloud_term_exact(double4 g, double a, double b, double c, double d) {
#pragma clang fp contract(off)
  return g.x * a + g.y * b + g.z * c + g.w * d;
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
  int64_t o;
  int32_t L;
  abc_row_span(a, row, o, L);
  const double *g = a.gth + 2 * (int64_t)a.abc_gofs[row];  // G_theta needs no lattice transform
  const int32_t *cl = a.cloud + o;
  double acc = 0.0;
  for (int base = 0; base < L; base += 64) {
    const int jj = base + lane;
    double t = 0.0;
    int32_t k = -1;
    if (jj < L) {
      k = cl[jj];
      if (k >= 0) {
        const double2 gm = *(const double2 *)(g + 2 * a.abc_pmap[o + jj]);
        t = gtheta_term_exact(gm.x, gm.y, st.thn[k], st.thnm1[k]);
      }
    }
    if (EXACT) {
      sw->c[0][lane] = t;
      sw->flag[lane] = (k >= 0);
      __threadfence_block();
      if (lane == 0) {
        const int lim = min(64, L - base);
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
  int64_t o;
  int32_t L;
  abc_row_span(a, row, o, L);
  const int64_t gb = a.abc_gofs[row];
  const int xf = a.abc_xf[row];
  const int32_t *cl = a.cloud + o;
  double Ux = 0.0, Uy = 0.0, Vx = 0.0, Vy = 0.0;
  for (int base = 0; base < L; base += 64) {
    const int jj = base + lane;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
    int32_t k = -1;
    if (jj < L) {
      k = cl[jj];
      if (k >= 0) {
        double4 g0, g1, h0, h1;
        abc_uv_coeffs(a, o, L, gb, xf, jj, g0, g1, h0, h1);
        const double ux = st.ux[k], uy = st.uy[k], uxm = st.uxm[k], uym = st.uym[k];
        const double vx = st.vx[k], vy = st.vy[k], vxm = st.vxm[k], vym = st.vym[k];
        t0 = cloud_term_exact(g0, ux, uy, uxm, uym);
        t1 = cloud_term_exact(g1, ux, uy, uxm, uym);
        t2 = cloud_term_exact(h0, vx, vy, vxm, vym);
        t3 = cloud_term_exact(h1, vx, vy, vxm, vym);
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
        const int lim = min(64, L - base);
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
// Tile t of block b (xcd_block): each XCD sweeps a contiguous band of tile rows, so a tile's halo
// neighbours sit in the same XCD's L2.
__device__ __forceinline__ int32_t tile_of_block(int32_t b, int32_t n_tiles) {
  return xcd_block(b, n_tiles);
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

__global__ void __launch_bounds__(kBlock) k_theta_tile(StepArgs a, StepState st) {
  __shared__ TileShared sh;
  const TileArgs &T = a.tile;
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

// The masked 68-bond dilatation sum of the node at window cell c (row stride kW68), before its
// 2 / m_i: the per-bond terms of src/main.cpp:736-761 over the cells with ok set, four chains (bond
// k in chain k % 4). The roundings are fixed here (explicit fma, no contraction left to the
// compiler), so the partial tiles' whole-tile blocks (k_theta_t68) and quarter-tile blocks
// (theta_part_quarter) give the same bits.
__device__ __forceinline__ double theta_masked68(const double2 *su, const int8_t *ok, int32_t c,
                                                 double2 ui, const TileArgs &T) {
#pragma clang fp contract(off)
  constexpr Stencil68 S = make_stencil68();
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 68; k++) {
    const Off68 o = S.o[k];
    const int32_t cj = c + o.ox * kW68 + o.oy;
    const double2 uj = su[cj];
    const double dx = uj.x - ui.x, dy = uj.y - ui.y;
    const double d = o.ox == 0 ? dy * o.oy : (o.oy == 0 ? dx * o.ox : fma(dx, (double)o.ox, dy * o.oy));
    const double term = T.c1[o.cls] * d;
    acc[k % 4] = ok[cj] ? acc[k % 4] + term : acc[k % 4];
  }
  return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// Dilatation over the tiles of list (all tiles when list is null), one node per thread.
__global__ void __launch_bounds__(kBlock) k_theta_t68(StepArgs a, StepState st,
                                                      const int32_t *list, int32_t count) {
  __shared__ Tile68Shared sh;
  const TileArgs &T = a.tile;
  const int32_t q = tile_of_block(blockIdx.x, count);
  if (q >= count) return;
  const int32_t t = list ? list[q] : q;
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
    st.theta[i] = a.c2m[i] * theta_masked68(sh.th.u, sh.th.ok, c, ui, T);
    return;
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

// Force-pass epilogue of node slot i: acceleration, velocity-Verlet corrector and the next step's
// prediction (src/main.cpp:874, 922-925, 725-728). At an ABC node (kind 2) only the acceleration:
// G_u / G_v replace its displacement and velocity (k_abc_uv) and k_abc_predict then predicts it.
// U^{n+1} of an interior node is its prediction ui (src/main.cpp:922 keeps u): it is not stored
// again; it stays in the prediction ring (RingU), from which launch_u_from_ring fills U on demand.
// vi, ai: V^n and A^n of the node (read by the caller, possibly early); P: its instance.
__device__ __forceinline__ void node_update(const InstParams &P, const StepState &st, int32_t i,
                                            int8_t kd, double2 ui, double2 vi, double2 ai,
                                            double F0, double F1) {
  const double inv_rho = 1.0 / P.rho;
  const double a1x = F0 * inv_rho, a1y = F1 * inv_rho;
  st.ax1[i] = a1x; st.ay1[i] = a1y;
  if (kd != 1) return;
  const double dt = P.dt, h = 0.5 * dt;
  const double v1x = vi.x + h * (ai.x + a1x);
  const double v1y = vi.y + h * (ai.y + a1y);
  st.vx1[i] = v1x; st.vy1[i] = v1y;
  st.pnx[i] = ui.x + dt * (v1x + h * a1x);
  st.pny[i] = ui.y + dt * (v1y + h * a1y);
}

// Unrolled force kernel over the tiles of list. FULL: the full tiles, no masks (every window cell
// holds a node with m == m_ref); otherwise a bond to a cell without a node or with m <= 0 is
// skipped, and a node with m <= 0 gets no force (src/main.cpp:852).
// PRED: at ABC nodes also the next prediction, as k_abc_predict (G_u / G_v have written
// U^{n+1}, V^{n+1} there).
template <bool FULL, bool PRED = false>
__global__ void __launch_bounds__(kBlock, FULL ? 5 : 2) k_force_t68(StepArgs a, StepState st,
                                                                    const int32_t *list,
                                                                    int32_t count) {
  __shared__ Tile68Shared sh;
  const TileArgs &T = a.tile;
  const int32_t q = tile_of_block(blockIdx.x, count);
  if (q >= count) return;
  const int32_t t = list[q];
  const int32_t *nbr = T.tile_nbr + 9 * t;
  const InstParams &P = a.inst[T.tile_inst[t]];
  const double lm = P.lam - P.mu, mu4 = 4.0 * P.mu;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const int32_t c = threadIdx.x + r * kBlock;
    if (c < kCells68) {
      const int32_t s = window_slot68(nbr, c);
      double2 u = make_double2(0.0, 0.0), ab = make_double2(0.0, 0.0);
      if (FULL || (s >= 0 && a.kind[s] != 0)) {
        const double c2 = a.c2m[s];  // 0 when m <= 0
        u = make_double2(st.px[s], st.py[s]);
        ab = make_double2(lm * c2 * st.theta[s], mu4 * c2);
      }
      sh.fo.u[c] = u;
      sh.fo.ab[c] = ab;
    }
  }
  __syncthreads();
  const int32_t i = t * (kTile * kTile) + threadIdx.x;
  const int8_t kd = a.kind[i];
  if (kd == 0) return;
  const int32_t c = (threadIdx.x / kTile + 4) * kW68 + (threadIdx.x % kTile + 4);
  const double2 ui = sh.fo.u[c];
  const double2 abi = sh.fo.ab[c];
  double F0 = 0.0, F1 = 0.0;
  if (FULL || abi.y != 0.0) force_sum68<FULL>(sh.fo.u, sh.fo.ab, c, ui, abi, T, &F0, &F1);
  node_update(P, st, i, kd, ui, make_double2(st.vx[i], st.vy[i]), make_double2(st.ax[i], st.ay[i]),
              F0, F1);
  if (PRED && kd == 2) {
    const double dt = P.dt, h = 0.5 * dt;
    const double inv_rho = 1.0 / P.rho;
    const double a1x = F0 * inv_rho, a1y = F1 * inv_rho;
    st.pnx[i] = st.ux1[i] + dt * (st.vx1[i] + h * a1x);
    st.pny[i] = st.uy1[i] + dt * (st.vy1[i] + h * a1y);
  }
}

// ---------------------------------------------------------------------------------------------
// Register-blocked passes over the full tiles (every window cell a node with m == m_ref).
// A lane owns NR nodes in consecutive rows of one column of the tile (NR = 4: a wavefront per
// tile; NR = 2: two wavefronts per tile). The lane sweeps the window rows that these nodes reach:
// each row's cells are read from LDS once and feed the bonds of every node of the lane that
// reaches the row, so a node costs (9 NR + 60) / NR cell reads instead of 68.
// On a full tile m_j = m_i = m_ref for every bond, so the bond term collapses to class weights
// folded with the material (TileArgs::wt, wa, wb) and the theta_i part of the force,
// theta_i sum_k c1_k xi_hat_k, vanishes by the stencil's symmetry:
//   theta_i = sum_k C c1_k (u_j . o_k)
//   F_i     = sum_k (lambda - mu) C c1_k theta_j o_k + 8 mu C c2_k ((u_j - u_i) . o_k) o_k
// (o_k the integer offset, xi_hat = o_k dx / |xi|). The u_i term of the dilatation is dropped the
// same way (sum_k c1_k o_k = 0); the force keeps u_j - u_i, whose difference of nearby values is
// exact, since there the sum is a second difference.
// ---------------------------------------------------------------------------------------------

// The 24 x 24 window (halo 4) of a full tile in 9 wavefront-wide pieces: piece k of lane l is the
// pair (slot, window cell). Pieces 0-3 are the tile itself (64 consecutive slots each), 4 / 5 the
// rows above / below (the last / first 4 rows of the x-neighbours: 64 consecutive slots), 6 / 7 the
// columns left / right (4 slots of each of 16 rows), 8 the four 4 x 4 corners. nb: the 3 x 3
// neighbour tiles (wave-uniform; all present on a full tile, -1 where absent), nb[(dx + 1) * 3 +
// dy + 1].
struct WinCell {
  int32_t slot, cell, tile;  // tile < 0: the cell lies in an absent tile (slot is not valid)
};
template <int K>
__device__ __forceinline__ WinCell window_piece68(const int32_t *nb, int32_t l) {
  constexpr int kT = kTile * kTile;
  if constexpr (K < 4) {
    return WinCell{nb[4] * kT + l + 64 * K, ((l >> 4) + 4 * K + 4) * kW68 + (l & 15) + 4, nb[4]};
  } else if constexpr (K == 4) {
    return WinCell{nb[1] * kT + 192 + l, (l >> 4) * kW68 + (l & 15) + 4, nb[1]};
  } else if constexpr (K == 5) {
    return WinCell{nb[7] * kT + l, (20 + (l >> 4)) * kW68 + (l & 15) + 4, nb[7]};
  } else if constexpr (K == 6) {
    return WinCell{nb[3] * kT + (l >> 2) * kTile + 12 + (l & 3), ((l >> 2) + 4) * kW68 + (l & 3),
                   nb[3]};
  } else if constexpr (K == 7) {
    return WinCell{nb[5] * kT + (l >> 2) * kTile + (l & 3), ((l >> 2) + 4) * kW68 + 20 + (l & 3),
                   nb[5]};
  } else {
    const int32_t hx = (l >> 5) & 1, hy = (l >> 4) & 1, crx = (l >> 2) & 3, cry = l & 3;
    // corner tile nb[6 hx + 2 hy] as arithmetic on the four values (a select between them would be
    // turned into an indexed load from a stack copy of nb)
    const int32_t tq = nb[0] + hx * (nb[6] - nb[0]) + hy * (nb[2] - nb[0]) +
                       hx * hy * (nb[8] - nb[6] - nb[2] + nb[0]);
    return WinCell{tq * kT + ((hx ? 0 : 12) + crx) * kTile + (hy ? 0 : 12) + cry,
                   ((hx ? 20 : 0) + crx) * kW68 + (hy ? 20 : 0) + cry, tq};
  }
}

template <class F, int... K>
__device__ __forceinline__ void fill_window68(const int32_t *nb, int32_t l, F &&f,
                                              std::integer_sequence<int, K...>) {
  (f(window_piece68<K>(nb, l)), ...);
}

struct ClassRep {
  int a, b;
};
constexpr ClassRep class_rep(int cls) {
  for (int a = 0; a <= 4; a++)
    for (int b = 0; b <= a; b++)
      if (a * a + b * b == kStencilR2[cls]) return ClassRep{a, b};
  return ClassRep{-1, -1};
}
constexpr int iabs(int v) { return v < 0 ? -v : v; }
// Weight slot of the multiplier |component| of a bond of class cls: 0 for a, 1 for b.
constexpr int rep_slot(int cls, int m) { return class_rep(cls).a == m ? 0 : 1; }
// Weight slot of the product p in {a^2, a b, b^2}.
constexpr int prod_slot(int cls, int p) {
  return p == class_rep(cls).a * class_rep(cls).a ? 0
         : (p == class_rep(cls).a * class_rep(cls).b ? 1 : 2);
}

constexpr bool in_stencil68(int ox, int oy) {
  return ox * ox + oy * oy > 0 && ox * ox + oy * oy <= 20;
}

// Keep each row's reads next to its bonds (the accumulators are complete at the end of the row):
// hoisting every read of the sweep ahead of the arithmetic needs more registers than a wavefront
// has.
__device__ __forceinline__ void pin(double &v) { asm volatile("" : "+v"(v)); }

// ---------------------------------------------------------------------------------------------
// Row-symmetric sweeps. On a full tile every term is linear in the window values, with a weight
// that is even or odd in each component of the offset. The cells (D, +oy) and (D, -oy) of a sweep
// row therefore enter every node of the lane through their sum or their difference: a row forms
// these once for the lane's NR nodes, and each node then takes one FMA per offset (ox, |oy|) and
// term, instead of one per bond and term.
//   dilatation:  x part on the sums of u_x, y part on the differences of u_y;
//   force, x:    C c1 ox theta_j on the sums of theta, B_xx on the sums of ubar_x, B_xy on the
//                differences of ubar_y;
//   force, y:    C c1 oy theta_j on the differences of theta, B_yy on the sums of ubar_y, B_xy on
//                the differences of ubar_x;
// with B_pq = 8 C c2 o_p o_q (the bond-extension weights wb). The bond extension (u_j - u_i) . o
// enters as ubar_j - ubar_i, ubar = u - c with c the displacement of one node of the lane:
//   sum_k B_k (u_j - u_i) = sum_k B_k ubar_j - (sum_k B_k) ubar_i,
// the sums of B over the stencil being TileArgs::wbsum. c lies within 7 cells of every cell the
// lane reads, so ubar is a difference of nearby values and the sum stays the size of a first
// difference of u, as with per-bond differences, not the field's own size (whose cancellation in
// the second difference would cost (sigma / dx)^2 in rounding).
// Per node: 64 FMAs for the dilatation (120 per bond), 180 for the force (466 operations per bond),
// plus per row of the lane 8 (dilatation) or 42 (force) operations shared by its NR nodes.
// ---------------------------------------------------------------------------------------------
template <class F, int... I>
__device__ __forceinline__ void static_for(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}

// A cell (D, OY) of a sweep row is read when some node r < NR of the lane has the bond (D - r, OY).
constexpr bool row_cell_needed(int D, int OY, int NR) {
  for (int r = 0; r < NR; r++)
    if (in_stencil68(D - r, OY)) return true;
  return false;
}

// Class weights of the offset (OX, OY), OY >= 0.
template <int OX, int OY>
__device__ __forceinline__ double w_ax(const TileArgs &T) {  // C c1 ox
  constexpr int cls = stencil_class(OX * OX + OY * OY);
  constexpr int z = rep_slot(cls, iabs(OX));
  return OX > 0 ? T.wt[cls][z] : -T.wt[cls][z];
}
template <int OX, int OY>
__device__ __forceinline__ double w_ay(const TileArgs &T) {  // C c1 oy, OY > 0
  constexpr int cls = stencil_class(OX * OX + OY * OY);
  return T.wt[cls][rep_slot(cls, OY)];
}
template <int OX, int OY>
__device__ __forceinline__ double w_bxx(const TileArgs &T) {  // 8 C c2 ox^2, OX != 0
  constexpr int cls = stencil_class(OX * OX + OY * OY);
  return T.wb[cls][prod_slot(cls, OX * OX)];
}
template <int OX, int OY>
__device__ __forceinline__ double w_byy(const TileArgs &T) {  // 8 C c2 oy^2, OY > 0
  constexpr int cls = stencil_class(OX * OX + OY * OY);
  return T.wb[cls][prod_slot(cls, OY * OY)];
}
template <int OX, int OY>
__device__ __forceinline__ double w_bxy(const TileArgs &T) {  // 8 C c2 ox oy, OX != 0, OY > 0
  constexpr int cls = stencil_class(OX * OX + OY * OY);
  constexpr int z = prod_slot(cls, iabs(OX) * OY);
  return OX > 0 ? T.wb[cls][z] : -T.wb[cls][z];
}

// Full-tile dilatation, row-symmetric: theta_i = sum_k C c1_k (u_j . o_k); c0 is the window cell
// of the lane's first node, NR its nodes (consecutive rows).
template <int NR, int W = kW68>
struct ThetaSym {
  const double2 *su;
  int32_t c0;
  const TileArgs &T;
  double acc[NR];
  template <int D>
  __device__ __forceinline__ void row() {
    double2 u[9];
    static_for([&](auto k) {
      constexpr int OY = decltype(k)::value - 4;
      if constexpr (row_cell_needed(D, OY, NR)) u[OY + 4] = su[c0 + D * W + OY];
    }, std::make_integer_sequence<int, 9>{});
    double sx[5], dy[5];
    static_for([&](auto k) {
      constexpr int OY = decltype(k)::value;
      if constexpr (row_cell_needed(D, OY, NR)) {
        if constexpr (OY == 0) {
          sx[0] = u[4].x;
        } else {
          sx[OY] = u[4 + OY].x + u[4 - OY].x;
          dy[OY] = u[4 + OY].y - u[4 - OY].y;
        }
      }
    }, std::make_integer_sequence<int, 5>{});
    static_for([&](auto rr) {
      constexpr int R = decltype(rr)::value, OX = D - R;
      if constexpr (OX >= -4 && OX <= 4) {
        static_for([&](auto yy) {
          constexpr int OY = decltype(yy)::value;
          if constexpr (in_stencil68(OX, OY)) {
            if constexpr (OX != 0) acc[R] = fma(w_ax<OX, OY>(T), sx[OY], acc[R]);
            if constexpr (OY != 0) acc[R] = fma(w_ay<OX, OY>(T), dy[OY], acc[R]);
          }
        }, std::make_integer_sequence<int, 5>{});
      }
    }, std::make_integer_sequence<int, NR>{});
#pragma unroll
    for (int r = 0; r < NR; r++) pin(acc[r]);
    __builtin_amdgcn_sched_barrier(0);
  }
};

// Full-tile force, row-symmetric (see above):
//   F_i = sum_k (lambda - mu) C c1_k theta_j o_k + 8 mu C c2_k ((u_j - u_i) . o_k) o_k;
// the theta_i part vanishes by the stencil's symmetry. The window holds theta scaled by the
// instance's th_scale = (lambda - mu) / mu and the caller scales the sum by its f_scale = mu, so the
// class tables hold no material. cu: the lane's reference displacement c; fx, fy start at
// -(sum of B) ubar_i. NR: the lane's nodes (consecutive rows).
template <int NR, int WU = kW68, int WT = kW68>
struct ForceSym {
  const double2 *su;
  const double *sth;
  int32_t cu0, ct0;  // the lane's first node in the u window (row stride WU) and theta window (WT)
  const TileArgs &T;
  double2 cu;
  double fx[NR], fy[NR];
  template <int D>
  __device__ __forceinline__ void row() {
    double2 u[9];
    double th[9];
    static_for([&](auto k) {
      constexpr int OY = decltype(k)::value - 4;
      if constexpr (row_cell_needed(D, OY, NR)) {
        const double2 v = su[cu0 + D * WU + OY];
        u[OY + 4] = make_double2(v.x - cu.x, v.y - cu.y);
        th[OY + 4] = sth[ct0 + D * WT + OY];
      }
    }, std::make_integer_sequence<int, 9>{});
    double st[5], dt[5], sx[5], dx[5], sy[5], dy[5];
    static_for([&](auto k) {
      constexpr int OY = decltype(k)::value;
      if constexpr (row_cell_needed(D, OY, NR)) {
        if constexpr (OY == 0) {
          st[0] = th[4];
          sx[0] = u[4].x;
          sy[0] = u[4].y;
        } else {
          st[OY] = th[4 + OY] + th[4 - OY];
          dt[OY] = th[4 + OY] - th[4 - OY];
          sx[OY] = u[4 + OY].x + u[4 - OY].x;
          dx[OY] = u[4 + OY].x - u[4 - OY].x;
          sy[OY] = u[4 + OY].y + u[4 - OY].y;
          dy[OY] = u[4 + OY].y - u[4 - OY].y;
        }
      }
    }, std::make_integer_sequence<int, 5>{});
    static_for([&](auto rr) {
      constexpr int R = decltype(rr)::value, OX = D - R;
      if constexpr (OX >= -4 && OX <= 4) {
        static_for([&](auto yy) {
          constexpr int OY = decltype(yy)::value;
          if constexpr (in_stencil68(OX, OY)) {
            if constexpr (OX != 0) {
              fx[R] = fma(w_ax<OX, OY>(T), st[OY], fx[R]);
              fx[R] = fma(w_bxx<OX, OY>(T), sx[OY], fx[R]);
              if constexpr (OY != 0) fx[R] = fma(w_bxy<OX, OY>(T), dy[OY], fx[R]);
            }
            if constexpr (OY != 0) {
              fy[R] = fma(w_ay<OX, OY>(T), dt[OY], fy[R]);
              fy[R] = fma(w_byy<OX, OY>(T), sy[OY], fy[R]);
              if constexpr (OX != 0) fy[R] = fma(w_bxy<OX, OY>(T), dx[OY], fy[R]);
            }
          }
        }, std::make_integer_sequence<int, 5>{});
      }
    }, std::make_integer_sequence<int, NR>{});
#pragma unroll
    for (int r = 0; r < NR; r++) {
      pin(fx[r]);
      pin(fy[r]);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
};

template <class S, int... E>
__device__ __forceinline__ void sym_sweep(S &s, std::integer_sequence<int, E...>) {
  (s.template row<E - 4>(), ...);
}

// Dilatation of a quarter of a partial tile (tile rows 4 b .. 4 b + 3, one node per lane) with the
// masked per-bond sum of k_theta_t68 (the same arithmetic): window rows 4 b .. 4 b + 11 (12 x 24
// cells) in LDS, u in su[0, 288), the cell masks as bytes from su[288] on.
__device__ __forceinline__ void theta_part_quarter(const StepArgs &a, const StepState &st,
                                                   double2 *su, int32_t t, int32_t b) {
  const TileArgs &T = a.tile;
  const int32_t *nbr = T.tile_nbr + 9 * t;
  const int32_t lane = threadIdx.x & 63;
  int8_t *ok = reinterpret_cast<int8_t *>(su + 12 * kW68);
  for (int32_t c = lane; c < 12 * kW68; c += 64) {
    const int32_t s = window_slot68(nbr, 4 * b * kW68 + c);
    const bool on = s >= 0 && a.kind[s] != 0;
    su[c] = on ? make_double2(st.px[s], st.py[s]) : make_double2(0.0, 0.0);
    ok[c] = on;
  }
  __syncthreads();
  const int32_t i = t * (kTile * kTile) + (4 * b + lane / kTile) * kTile + lane % kTile;
  if (a.kind[i] != 1) return;
  const int32_t c = (lane / kTile + 4) * kW68 + (lane % kTile + 4);
  st.theta[i] = a.c2m[i] * theta_masked68(su, ok, c, su[c], T);
}

// Fast-path ABC rows: 16 lanes per row, four rows per wavefront, lane q of a row on its cloud
// members q, q + 16, ... (a cloud has at most 49), then a sum over the row's 16 lanes. One wavefront
// per row left 15 to 25 of its lanes idle on the (ragged) clouds and spent 12 cross-lane moves per
// value on the 64-lane sum.
// The row kernels give each XCD a contiguous range of rows (tile_of_block's mapping over row
// blocks): rows are in node order, so neighbouring rows share cloud nodes and their state, and the
// shared values are then read through one L2.
constexpr int kRowLanes = 16;
constexpr int kRowsPerBlock = kBlock / kRowLanes;

// Sum over the 16 lanes of a row group; every lane of the group gets it.
__device__ __forceinline__ double group16_sum(double v) {
  for (int o = kRowLanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kRowLanes);
  return v;
}

// ABC row `row`'s dilatation on its 16 lanes (q = lane in the row): G_theta (src/main.cpp:792-812),
// or for IDW the direct value (overridden by k_idw where a non-ABC neighbour exists,
// src/main.cpp:765-790). Reads only the theta history (or the prediction).
__device__ __forceinline__ void abc_theta_rows16(const StepArgs &a, const StepState &st,
                                                 const AbcPack &p, int32_t row, int32_t q) {
  const int32_t node = a.abc_slot[row];
  double th;
  if (a.theta_method == 1) {
    int64_t o;
    int32_t L;
    abc_row_span(a, row, o, L);
    const double *g = a.gth + 2 * (int64_t)a.abc_gofs[row];  // G_theta needs no lattice transform
    const int32_t *cl = p.idx + (int64_t)row * p.stride;
    double acc = 0.0;
    // four members per lane at a time (q, q + 16, q + 32, q + 48): their loads are unconditional
    // (clamped indices), so the four members' gathers are in flight together; a member past the
    // cloud or a pad adds nothing. The entries' loads need only the row (fixed stride), not its
    // span, so they go out together with the span's.
    for (int base = q; base < p.stride; base += 4 * kRowLanes) {
      int32_t k[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int jj = base + u * kRowLanes;
        k[u] = jj < p.stride ? cl[jj] : -1;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int jj = max(min(base + u * kRowLanes, L - 1), 0);
        const int32_t ks = k[u] >= 0 ? k[u] : 0;
        const double2 gm = *(const double2 *)(g + 2 * jj);  // p.idx is in block order
        const double2 h = p.th[ks];  // (theta^n, theta^{n-1}) of the member
        const double t = gtheta_term_exact(gm.x, gm.y, h.x, h.y);
        acc = k[u] >= 0 ? acc + t : acc;
      }
    }
    th = group16_sum(acc);
  } else {
    th = q == 0 ? theta_node_exact(a, st.px, st.py, node) : 0.0;
  }
  if (q == 0) st.theta[node] = th;
}

// ABC row `row` on its 16 lanes: U^{n+1} = G_u [U^n, U^{n-1}] and V^{n+1} = G_v [V^n, V^{n-1}]
// over the row's cloud (src/main.cpp:883-916, 933-974). The tile kernels compute the force at the
// ABC nodes like at any node. The members' state comes packed (p: U of an interior member from the
// prediction ring, of an ABC member from U), one 64-byte record per member.
__device__ __forceinline__ void abc_uv_rows16(const StepArgs &a, const StepState &st, const AbcPack &p,
                                              int32_t row, int32_t q) {
  int64_t o;
  int32_t L;
  abc_row_span(a, row, o, L);
  const int64_t gb = a.abc_gofs[row];
  const int xf = a.abc_xf[row];
  const int32_t *cl = p.idx + (int64_t)row * p.stride;
  double Ux = 0.0, Uy = 0.0, Vx = 0.0, Vy = 0.0;
  // four members per lane at a time, loads unconditional (see abc_theta_rows16)
  for (int base = q; base < p.stride; base += 4 * kRowLanes) {
    int32_t k[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int jj = base + u * kRowLanes;
      k[u] = jj < p.stride ? cl[jj] : -1;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int jj = max(min(base + u * kRowLanes, L - 1), 0);
      const int32_t ks = k[u] >= 0 ? k[u] : 0;
      // p.idx is in block order: member jj's coefficients are the block's jj-th
      const double *bu = a.gu + 8 * gb + 4 * jj, *bv = a.gv + 8 * gb + 4 * jj;
      double4 g0, g1, h0, h1;
      lattice_map(xf, *(const double4 *)bu, *(const double4 *)(bu + 4 * L), g0, g1);
      lattice_map(xf, *(const double4 *)bv, *(const double4 *)(bv + 4 * L), h0, h1);
      const double4 sn = p.cur[ks], sm = p.prev[ks];
      const double ux = sn.x, uy = sn.y, uxm = sm.x, uym = sm.y;
      const double vx = sn.z, vy = sn.w, vxm = sm.z, vym = sm.w;
      const double tux = g0.x * ux + g0.y * uy + g0.z * uxm + g0.w * uym;
      const double tuy = g1.x * ux + g1.y * uy + g1.z * uxm + g1.w * uym;
      const double tvx = h0.x * vx + h0.y * vy + h0.z * vxm + h0.w * vym;
      const double tvy = h1.x * vx + h1.y * vy + h1.z * vxm + h1.w * vym;
      if (k[u] >= 0) {
        Ux += tux;
        Uy += tuy;
        Vx += tvx;
        Vy += tvy;
      }
    }
  }
  Ux = group16_sum(Ux);
  Uy = group16_sum(Uy);
  Vx = group16_sum(Vx);
  Vy = group16_sum(Vy);
  if (q == 0) {
    const int32_t node = a.abc_slot[row];
    st.ux1[node] = Ux; st.uy1[node] = Uy;
    st.vx1[node] = Vx; st.vy1[node] = Vy;
  }
}

// Fast path, ABC rows' dilatation, kRowsPerBlock rows per block; may run beside the tile kernels.
// Launched on abc_blocks_xcd(a) blocks.
__global__ void __launch_bounds__(kBlock) k_abc_theta(StepArgs a, StepState st, AbcPack p) {
  const int32_t nrb = (a.n_abc + kRowsPerBlock - 1) / kRowsPerBlock;
  const int32_t row = tile_of_block(blockIdx.x, nrb) * kRowsPerBlock + (int32_t)threadIdx.x / kRowLanes;
  if (row < a.n_abc) abc_theta_rows16(a, st, p, row, threadIdx.x & (kRowLanes - 1));
}

// Fast path, ABC rows' G_u / G_v, kRowsPerBlock rows per block. Launched on abc_blocks_xcd(a)
// blocks.
__global__ void __launch_bounds__(kBlock) k_abc_uv(StepArgs a, StepState st, AbcPack p) {
  const int32_t nrb = (a.n_abc + kRowsPerBlock - 1) / kRowsPerBlock;
  const int32_t row = tile_of_block(blockIdx.x, nrb) * kRowsPerBlock + (int32_t)threadIdx.x / kRowLanes;
  if (row < a.n_abc) abc_uv_rows16(a, st, p, row, threadIdx.x & (kRowLanes - 1));
}

// The packed state of the ABC clouds' members (AbcPack), one thread per member.
__global__ void __launch_bounds__(kBlock) k_abc_pack(int32_t n, const int32_t *slot,
                                                     const int8_t *kind, StepState st, RingU r,
                                                     double4 *cur, double4 *prev, double2 *th) {
  const int32_t m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= n) return;
  const int32_t i = slot[m];
  const bool in = kind[i] == 1;
  cur[m] = make_double4((in ? r.ux : st.ux)[i], (in ? r.uy : st.uy)[i], st.vx[i], st.vy[i]);
  if (prev)
    prev[m] = make_double4((in ? r.uxm : st.uxm)[i], (in ? r.uym : st.uym)[i], st.vxm[i], st.vym[i]);
  if (th) th[m] = make_double2(st.thn[i], st.thnm1[i]);
}

// The register-blocked tile kernels take NR nodes per lane: NR = 4, one wavefront per tile; NR = 2,
// two wavefronts per tile sharing its window (half the LDS per wavefront, so twice the wavefronts per
// CU, for more rows swept per node). Lane t of the tile's 256 / NR threads owns the nodes
// (NR (t / 16) + r, t % 16), r < NR. Both fill the window with the same 9 pieces (the wavefronts of
// an NR = 2 block take alternate pieces), and their node arithmetic is the same, so they give the
// same bits.
template <int NR, class F>
__device__ __forceinline__ void fill_tile_window(const int32_t *nb, int32_t t, F &&f) {
  if constexpr (NR == 4) {
    fill_window68(nb, t, f, std::make_integer_sequence<int, 9>{});
  } else {
    if ((t >> 6) == 0) fill_window68(nb, t & 63, f, std::integer_sequence<int, 0, 2, 4, 6, 8>{});
    else fill_window68(nb, t & 63, f, std::integer_sequence<int, 1, 3, 5, 7>{});
  }
}

// LDS of theta_part_quarter: 12 x 24 cells of u, then their masks as bytes (in double2 units).
constexpr int kQuarterCells = 12 * kW68 + (12 * kW68 + 15) / 16;

// Dilatation of the full tiles, one tile per block (full_list order); blocks past the full tiles'
// (rounded up to a multiple of 8) take the partial tiles, a quarter of a tile per wavefront, so they
// fill the tail of the full tiles' launch instead of following it. With nrow_blocks > 0 (a multiple
// of 8) the first nrow_blocks blocks take the ABC rows instead, four per wavefront: G_theta and
// G_u / G_v read only the previous steps' state, so the step's ABC work runs in the same launch as
// the dilatation.
template <int NR>
__global__ void __launch_bounds__(256 / NR, NR == 4 ? 4 : 6) k_theta_rb(StepArgs a, StepState st,
                                                                        AbcPack pack,
                                                                        int32_t nrow_blocks,
                                                                        const int32_t *flist,
                                                                        int32_t nfull) {
  constexpr int kWaves = 4 / NR;
  __shared__ double2 su[kCells68 > kWaves * kQuarterCells ? kCells68 : kWaves * kQuarterCells];
  const TileArgs &T = a.tile;
  const int32_t t = threadIdx.x, wv = t >> 6;
  if ((int32_t)blockIdx.x < nrow_blocks) {
    const int32_t nrb = (a.n_abc + 4 * kWaves - 1) / (4 * kWaves);
    const int32_t row = tile_of_block(blockIdx.x, nrb) * 4 * kWaves + t / kRowLanes;
    if (row < a.n_abc) {
      abc_theta_rows16(a, st, pack, row, t & (kRowLanes - 1));
      abc_uv_rows16(a, st, pack, row, t & (kRowLanes - 1));
    }
    return;
  }
  const int32_t b = blockIdx.x - nrow_blocks;
  const int32_t nfb = (nfull + 7) / 8 * 8;
  if (b >= nfb) {
    const int32_t pq = (b - nfb) * kWaves + wv;  // quarter pq & 3 of partial tile pq >> 2
    if ((pq >> 2) < T.n_part)
      theta_part_quarter(a, st, su + wv * kQuarterCells, T.part_list[pq >> 2], pq & 3);
    return;
  }
  const int32_t q = tile_of_block(b, nfull);
  if (q >= nfull) return;
  const int32_t tl = flist[q];
  int32_t nb[9];
#pragma unroll
  for (int k = 0; k < 9; k++) nb[k] = T.tile_nbr[9 * tl + k];
  const int32_t ly = t & 15, lx0 = (t >> 4) * NR;
  int8_t kd[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) kd[r] = a.kind[tl * (kTile * kTile) + (lx0 + r) * kTile + ly];
  fill_tile_window<NR>(nb, t, [&](WinCell w) { su[w.cell] = make_double2(st.px[w.slot], st.py[w.slot]); });
  __syncthreads();
  ThetaSym<NR> P{su, (lx0 + 4) * kW68 + ly + 4, T, {}};
#pragma unroll
  for (int r = 0; r < NR; r++) P.acc[r] = 0.0;
  sym_sweep(P, std::make_integer_sequence<int, NR + 8>{});
#pragma unroll
  for (int r = 0; r < NR; r++)
    if (kd[r] == 1) st.theta[tl * (kTile * kTile) + (lx0 + r) * kTile + ly] = P.acc[r];
}

// Force pass of the full tiles with the fused Verlet corrector and predictor; one tile per block
// (full_list order), NR nodes per lane.
template <int NR>
__global__ void __launch_bounds__(256 / NR, NR == 4 ? 3 : 4) k_force_rb(StepArgs a, StepState st,
                                                                        const int32_t *flist,
                                                                        int32_t nfull) {
  __shared__ double2 su[kCells68];
  __shared__ double sth[kCells68];
  const TileArgs &T = a.tile;
  const int32_t q = tile_of_block(blockIdx.x, nfull);
  if (q >= nfull) return;
  const int32_t tl = flist[q];
  int32_t nb[9];
#pragma unroll
  for (int k = 0; k < 9; k++) nb[k] = T.tile_nbr[9 * tl + k];
  const int32_t t = threadIdx.x;
  const int32_t ly = t & 15, lx0 = (t >> 4) * NR;
  // the epilogue's inputs first: their loads overlap the window's
  int8_t kd[NR];
  double2 vi[NR], ai[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int32_t i = tl * (kTile * kTile) + (lx0 + r) * kTile + ly;
    kd[r] = a.kind[i];
    vi[r] = make_double2(st.vx[i], st.vy[i]);
    ai[r] = make_double2(st.ax[i], st.ay[i]);
  }
  const double ratio = a.inst[T.tile_inst[tl]].th_scale;
  fill_tile_window<NR>(nb, t, [&](WinCell w) {
    su[w.cell] = make_double2(st.px[w.slot], st.py[w.slot]);
    sth[w.cell] = ratio * st.theta[w.slot];
  });
  __syncthreads();
  const int32_t c0 = (lx0 + 4) * kW68 + ly + 4;
  double2 ui[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) ui[r] = su[c0 + r * kW68];
  // c: the displacement of row 4 floor(lx0 / 4) + 1 of the lane's column, the same node for NR = 4
  // and NR = 2 (the same bits)
  ForceSym<NR> P{su, sth, c0, c0, T, su[((lx0 & ~3) + 5) * kW68 + ly + 4], {}, {}};
#pragma unroll
  for (int r = 0; r < NR; r++) {
    P.fx[r] = -T.wbsum[0] * (ui[r].x - P.cu.x);
    P.fy[r] = -T.wbsum[1] * (ui[r].y - P.cu.y);
  }
  sym_sweep(P, std::make_integer_sequence<int, NR + 8>{});
  const InstParams &I = a.inst[T.tile_inst[tl]];
  const double scale = I.f_scale;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int32_t i = tl * (kTile * kTile) + (lx0 + r) * kTile + ly;
    if (kd[r] != 0)
      node_update(I, st, i, kd[r], ui[r], vi[r], ai[r], scale * P.fx[r], scale * P.fy[r]);
  }
}

constexpr int kW32 = kTile + 16;  // u window of k_force_deep: 32 x 32 cells (halo 8)

// Dilatation and force of a "deep" full tile in one launch. A deep tile's 3 x 3 neighbourhood is
// full and its 24 x 24 window holds interior nodes only (kind 1), so the theta of every window node
// is k_theta_rb's full-tile sum over u alone: the block computes the 576 values itself from a
// 32 x 32 window of u (lanes 0..95, 6 window rows each), stores the tile's own 256 for the other
// tiles' force pass, and runs k_force_rb<2>'s sweep and epilogue on them. The per-node arithmetic
// of both sums is that of k_theta_rb / k_force_rb<2> (the same FMA sequence per node), so the
// results are the same bits as the two-launch pass; the dilatation of the halo ring is computed
// twice (by each tile whose window holds it), 2.25 times the dilatation arithmetic for the deep
// tiles, in exchange for no dilatation launch in front of their force.
__global__ void __launch_bounds__(128, 3) k_force_deep(StepArgs a, StepState st,
                                                       const int32_t *list, int32_t count) {
  __shared__ double2 su[kW32 * kW32];
  __shared__ double sth[kCells68];
  const TileArgs &T = a.tile;
  const int32_t q = tile_of_block(blockIdx.x, count);
  if (q >= count) return;
  const int32_t tl = list[q];
  int32_t nb[9];
#pragma unroll
  for (int k = 0; k < 9; k++) nb[k] = T.tile_nbr[9 * tl + k];
  const int32_t t = threadIdx.x;
  const int32_t ly = t & 15, lx0 = (t >> 4) * 2;
  const double ratio = a.inst[T.tile_inst[tl]].th_scale;
  {
    // u window: column wc = t & 31 (tile column wc - 8), rows (t >> 5) + 4 k, k = 0..7 (tile rows
    // -8..-5 | 0..15 | 16..23: the x-neighbour -1, 0, +1 known per k at compile time)
    const int32_t wc = t & 31, y = wc - 8;
    const int32_t ay = y < 0 ? -1 : (y < kTile ? 0 : 1);
    const int32_t yy = y - ay * kTile;
    int32_t tcol[3];  // the tile of this column in x-neighbour row -1, 0, +1
#pragma unroll
    for (int ax = 0; ax < 3; ax++)
      tcol[ax] = ay < 0 ? nb[3 * ax] : (ay == 0 ? nb[3 * ax + 1] : nb[3 * ax + 2]);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int32_t wr = (t >> 5) + 4 * k, x = wr - 8;
      constexpr int kAx[8] = {-1, -1, 0, 0, 0, 0, 1, 1};
      const int32_t s = tcol[kAx[k] + 1] * (kTile * kTile) + (x - kAx[k] * kTile) * kTile + yy;
      su[wr * kW32 + wc] = make_double2(st.px[s], st.py[s]);
    }
  }
  __syncthreads();
  if (t < 96) {
    const int32_t tc = t % 24, tr0 = (t / 24) * 6;  // theta-window column and first row
    ThetaSym<6, kW32> P{su, (tr0 + 4) * kW32 + tc + 4, T, {}};
#pragma unroll
    for (int r = 0; r < 6; r++) P.acc[r] = 0.0;
    sym_sweep(P, std::make_integer_sequence<int, 14>{});
#pragma unroll
    for (int r = 0; r < 6; r++) {
      const int32_t cr = tr0 + r;
      sth[cr * kW68 + tc] = ratio * P.acc[r];
      const int32_t x = cr - 4, y = tc - 4;
      if (x >= 0 && x < kTile && y >= 0 && y < kTile)
        st.theta[tl * (kTile * kTile) + x * kTile + y] = P.acc[r];
    }
  }
  __syncthreads();
  // the epilogue's inputs (here, not in front of the dilatation sweep, which needs the registers):
  // their loads overlap the force sweep
  int8_t kd[2];
  double2 vi[2], ai[2];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int32_t i = tl * (kTile * kTile) + (lx0 + r) * kTile + ly;
    kd[r] = a.kind[i];
    vi[r] = make_double2(st.vx[i], st.vy[i]);
    ai[r] = make_double2(st.ax[i], st.ay[i]);
  }
  const int32_t cu0 = (lx0 + 8) * kW32 + ly + 8, ct0 = (lx0 + 4) * kW68 + ly + 4;
  double2 ui[2];
#pragma unroll
  for (int r = 0; r < 2; r++) ui[r] = su[cu0 + r * kW32];
  // c: the displacement of tile row 4 floor(lx0 / 4) + 1 of the lane's column, as k_force_rb
  ForceSym<2, kW32, kW68> P{su, sth, cu0, ct0, T, su[((lx0 & ~3) + 9) * kW32 + ly + 8], {}, {}};
#pragma unroll
  for (int r = 0; r < 2; r++) {
    P.fx[r] = -T.wbsum[0] * (ui[r].x - P.cu.x);
    P.fy[r] = -T.wbsum[1] * (ui[r].y - P.cu.y);
  }
  sym_sweep(P, std::make_integer_sequence<int, 10>{});
  const InstParams &I = a.inst[T.tile_inst[tl]];
  const double scale = I.f_scale;
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int32_t i = tl * (kTile * kTile) + (lx0 + r) * kTile + ly;
    if (kd[r] != 0)
      node_update(I, st, i, kd[r], ui[r], vi[r], ai[r], scale * P.fx[r], scale * P.fy[r]);
  }
}

// Fast path, after the force pass: the next prediction at the ABC nodes from G_u, G_v and the
// force kernels' acceleration (src/main.cpp:725-728).
__global__ void __launch_bounds__(kBlock) k_abc_predict(StepArgs a, StepState st) {
  const int32_t row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= a.n_abc) return;
  const int32_t i = a.abc_slot[row];
  const double dt = a.inst[a.tile.tile_inst[i / (kTile * kTile)]].dt, h = 0.5 * dt;
  st.pnx[i] = st.ux1[i] + dt * (st.vx1[i] + h * st.ax1[i]);
  st.pny[i] = st.uy1[i] + dt * (st.vy1[i] + h * st.ay1[i]);
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
  const InstParams &P = a.inst[T.tile_inst[t]];
  const double lm = P.lam - P.mu, mu4 = 4.0 * P.mu;
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
  const int8_t kd = a.kind[i];
  if (kd == 0) return;
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
  node_update(P, st, i, kd, ui, make_double2(st.vx[i], st.vy[i]), make_double2(st.ax[i], st.ay[i]),
              F0, F1);
}

// For every node slot, the stencil cells holding a node must be exactly the set of its CSR row
// (self excluded) for the tile path to reproduce the reference's horizons. Both sets as bit masks
// over the stencil's bonds (nt <= 128): bond_of maps an offset (ox, oy), |ox|, |oy| <= halo, to
// its bond, -1 off the stencil.
__global__ void __launch_bounds__(kBlock) k_tile_check(StepArgs a, const int16_t *bond_of,
                                                       int32_t *mismatch) {
  const TileArgs &T = a.tile;
  const int32_t t = blockIdx.x;
  const int32_t i = t * (kTile * kTile) + threadIdx.x;
  if (a.kind[i] == 0) return;
  const int32_t txy = T.tile_xy[t];
  const int32_t gx = (txy >> 16) * kTile + threadIdx.x / kTile;
  const int32_t gy = (txy & 0xffff) * kTile + threadIdx.x % kTile;
  uint64_t want[2] = {0, 0}, got[2] = {0, 0};
  for (int32_t k = 0; k < T.nt; k++) {
    // delta = ox * win + oy with |ox|, |oy| <= halo < win / 2
    const int32_t d = T.delta[k];
    const int32_t ox = (d + T.halo + T.halo * T.win) / T.win - T.halo;
    const int32_t oy = d - ox * T.win;
    const int32_t s = lattice_slot(T, gx + ox, gy + oy);
    if (s >= 0 && a.kind[s] != 0) want[k >> 6] |= 1ull << (k & 63);
  }
  const int32_t E = 2 * T.halo + 1;
  bool bad = false;
  for (int64_t q = a.row_ptr[i] + 1; q < a.row_ptr[i + 1] && !bad; q++) {
    const int32_t j = a.col[q];
    const int32_t tj = j / (kTile * kTile), lj = j % (kTile * kTile);
    const int32_t ox = (T.tile_xy[tj] >> 16) * kTile + lj / kTile - gx;
    const int32_t oy = (T.tile_xy[tj] & 0xffff) * kTile + lj % kTile - gy;
    const int32_t k = (ox < -T.halo || ox > T.halo || oy < -T.halo || oy > T.halo)
                          ? -1 : bond_of[(ox + T.halo) * E + oy + T.halo];
    const uint64_t bit = k >= 0 ? 1ull << (k & 63) : 0;
    bad = k < 0 || (got[k >> 6] & bit) != 0;
    if (!bad) got[k >> 6] |= bit;
  }
  if (bad || want[0] != got[0] || want[1] != got[1]) atomicAdd(mismatch, 1);
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

// Butterfly over the wavefront: each round merges a lane's partial with its partner's, lane l ^ off
// (sums commute, the maxima break ties by the lower node id), so a fixed pattern of merges.
__device__ __forceinline__ void wave_reduce_diag(DiagPartial &d) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    DiagPartial o;
    o.energy = __shfl_xor(d.energy, off);
    o.ek = __shfl_xor(d.ek, off);
    o.ep = __shfl_xor(d.ep, off);
    o.px = __shfl_xor(d.px, off);
    o.py = __shfl_xor(d.py, off);
    o.max_u = __shfl_xor(d.max_u, off);
    o.max_v = __shfl_xor(d.max_v, off);
    o.max_a = __shfl_xor(d.max_a, off);
    o.th_max = __shfl_xor(d.th_max, off);
    o.th_min = __shfl_xor(d.th_min, off);
    o.ub = __shfl_xor(d.ub, off);
    o.vb = __shfl_xor(d.vb, off);
    o.max_u_orig = __shfl_xor(d.max_u_orig, off);
    o.max_v_orig = __shfl_xor(d.max_v_orig, off);
    diag_merge(d, o);
  }
}

// Block reduction: the wavefronts' results (lane 0 of each) merged in wavefront order. Every
// thread returns the block's result. sh: kWavesPerBlock entries.
__device__ void block_reduce_diag(DiagPartial &d, DiagPartial *sh) {
  wave_reduce_diag(d);
  const int32_t wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[wid] = d;
  __syncthreads();
  d = sh[0];
  for (int w = 1; w < kWavesPerBlock; w++) diag_merge(d, sh[w]);
}

// Thread-strided merge of p[0, np) in a fixed order (thread j: j, j + kBlock, ...), four loads in
// flight per thread.
__device__ __forceinline__ void merge_strided(DiagPartial &d, const DiagPartial *p, int32_t np) {
  int32_t t = threadIdx.x;
  for (; t + 3 * kBlock < np; t += 4 * kBlock) {
    const DiagPartial p0 = p[t], p1 = p[t + kBlock], p2 = p[t + 2 * kBlock], p3 = p[t + 3 * kBlock];
    diag_merge(d, p0);
    diag_merge(d, p1);
    diag_merge(d, p2);
    diag_merge(d, p3);
  }
  for (; t < np; t += kBlock) diag_merge(d, p[t]);
}

__global__ void __launch_bounds__(kBlock) k_diag(StepArgs a, const double *ux, const double *uy,
                                                 const double *vx, const double *vy,
                                                 const double *ax, const double *ay,
                                                 const double *theta, const int32_t *perm,
                                                 double exclude_r, int32_t per_block,
                                                 DiagPartial *partials) {
  __shared__ DiagPartial sh[kWavesPerBlock];
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
                                                      const int32_t *perm,
                                                      DiagPartial *partials) {
  __shared__ DiagPartial sh[kWavesPerBlock];
  __shared__ double2 su[kMaxWin * kMaxWin];
  __shared__ int8_t sok[kMaxWin * kMaxWin];
  const TileArgs &T = a.tile;
  const int32_t t = blockIdx.x;
  const int32_t *nbr = T.tile_nbr + 9 * t;
  const InstParams &P = a.inst[T.tile_inst[t]];
  const double exclude_r = P.exclude_r;
  const int32_t W = T.win, K = T.halo;
  for (int32_t c = threadIdx.x; c < W * W; c += kBlock) {
    // window cell (rx, ry) relative to the tile's origin, |offset| <= halo < 16: one of the 3 x 3
    // neighbour tiles
    const int32_t rx = c / W - K, ry = c % W - K;
    const int32_t tx = rx < 0 ? 0 : (rx < kTile ? 1 : 2), ty = ry < 0 ? 0 : (ry < kTile ? 1 : 2);
    const int32_t tt = nbr[tx * 3 + ty];
    const int32_t s = tt < 0 ? -1
                             : tt * (kTile * kTile) + (rx - (tx - 1) * kTile) * kTile +
                                   (ry - (ty - 1) * kTile);
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
      d.px = P.rho * vxi * a.dx * a.dy;
      d.py = P.rho * vyi * a.dx * a.dy;
      const double ek = 0.5 * P.rho * (vxi * vxi + vyi * vyi) * a.dx * a.dy;
      d.ek = ek;
      const int32_t c = (threadIdx.x / kTile + T.halo) * W + (threadIdx.x % kTile + T.halo);
      double acc = 0.0;
      if (T.tile_full[t]) {  // every window cell holds a node: no mask
        const double thh = 0.5 * th;
        for (int32_t k = 0; k < T.nt; k++) {
          const double2 uj = su[c + T.delta[k]];
          const double ed = fma(uj.y - uyi, T.hy[k], (uj.x - uxi) * T.hx[k]) - thh * T.xi[k];
          acc = fma(ed * T.g2[k], ed, acc);
        }
      } else {
        for (int32_t k = 0; k < T.nt; k++) {
          const int32_t cj = c + T.delta[k];
          const double hx = T.hx[k], hy = T.hy[k], g2 = T.g2[k], xin = T.xi[k];
          const double2 uj = su[cj];
          const double ed = (uj.x - uxi) * hx + (uj.y - uyi) * hy - th * xin * 0.5;
          acc += sok[cj] ? ed * ed * g2 : 0.0;
        }
      }
      const double W_i = 0.5 * P.kappa * th * th + (4.0 * P.mu / a.m[i]) * acc;
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
  __shared__ DiagPartial sh[kWavesPerBlock];
  DiagPartial d;
  diag_identity(d);
  merge_strided(d, partials, np);
  block_reduce_diag(d, sh);
  if (threadIdx.x == 0) *out = d;
}

// One block per instance over its tiles' partials, in k_diag_final's order.
__global__ void __launch_bounds__(kBlock) k_diag_final_inst(const DiagPartial *partials,
                                                            const int32_t *tile_off,
                                                            DiagPartial *out) {
  __shared__ DiagPartial sh[kWavesPerBlock];
  const int32_t i = blockIdx.x, t0 = tile_off[i], np = tile_off[i + 1] - t0;
  DiagPartial d;
  diag_identity(d);
  merge_strided(d, partials + t0, np);
  block_reduce_diag(d, sh);
  if (threadIdx.x == 0) out[i] = d;
}

// ---------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------
static inline int32_t blocks_for(int64_t n) { return (int32_t)((n + kBlock - 1) / kBlock); }
static inline int32_t abc_blocks(const StepArgs &a) {
  return (a.n_abc + kWavesPerBlock - 1) / kWavesPerBlock;
}
// Blocks of the fast-path row kernels (kRowsPerBlock rows each) in 8 equal XCD ranges; the tail
// blocks exit.
static inline int32_t abc_blocks_xcd(const StepArgs &a) {
  return ((a.n_abc + kRowsPerBlock - 1) / kRowsPerBlock + 7) / 8 * 8;
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

static inline int32_t list_blocks(int32_t count) { return (count + 7) / 8 * 8; }

// The partial tiles of the register-blocked family: one node per thread (256 threads per tile).
// There are few of them, so the pass takes about one tile's latency, 4 times shorter than with
// 4 nodes per lane.
hipError_t launch_theta_part(const StepArgs &a, const StepState &st, hipStream_t s) {
  const TileArgs &T = a.tile;
  if (T.theta_kernel != kBlocked || T.n_part == 0) return hipSuccess;
  hipLaunchKernelGGL(k_theta_t68, dim3(list_blocks(T.n_part)), dim3(kBlock), 0, s, a, st,
                     T.part_list, T.n_part);
  return hipGetLastError();
}

hipError_t launch_theta(const StepArgs &a, const StepState &st, bool fast, hipStream_t s,
                        bool with_part) {
  if (fast) {
    const TileArgs &T = a.tile;
    if (T.theta_kernel == kBlocked) {
      // the partial tiles in the same launch (a quarter per wavefront) unless another stream takes
      // them
      const int32_t np = with_part ? T.n_part : 0;
      if (T.theta_nr == 2) {
        const int32_t nb = list_blocks(T.n_full) + 2 * np;
        if (nb > 0)
          hipLaunchKernelGGL(k_theta_rb<2>, dim3(nb), dim3(128), 0, s, a, st, AbcPack{}, 0, T.full_list,
                             T.n_full);
      } else {
        const int32_t nb = list_blocks(T.n_full) + 4 * np;
        if (nb > 0)
          hipLaunchKernelGGL(k_theta_rb<4>, dim3(nb), dim3(64), 0, s, a, st, AbcPack{}, 0, T.full_list,
                             T.n_full);
      }
      return hipGetLastError();
    } else if (T.theta_kernel == kUnrolled) {
      hipLaunchKernelGGL(k_theta_t68, dim3(tile_blocks(a)), dim3(kBlock), 0, s, a, st,
                         (const int32_t *)nullptr, T.n_tiles);
    } else {
      hipLaunchKernelGGL(k_theta_tile, dim3(tile_blocks(a)), dim3(kBlock), 0, s, a, st);
    }
  } else {
    const int32_t nb = blocks_for(a.n);
    hipLaunchKernelGGL(k_theta_exact, dim3(nb + abc_blocks(a)), dim3(kBlock), 0, s, a, st, nb);
  }
  return hipGetLastError();
}

hipError_t launch_theta_rows(const StepArgs &a, const StepState &st, const AbcPack &r, hipStream_t s) {
  const TileArgs &T = a.tile;
  if (T.theta_kernel != kBlocked) {  // the row kernels first, then the dilatation, on one stream
    hipError_t e;
    if ((e = launch_abc_theta(a, st, r, s)) != hipSuccess) return e;
    if ((e = launch_abc_uv(a, st, r, s)) != hipSuccess) return e;
    return launch_theta(a, st, true, s, true);
  }
  if (T.theta_nr == 2) {
    const int32_t nrow = ((a.n_abc + 7) / 8 + 7) / 8 * 8;
    const int32_t nb = nrow + list_blocks(T.n_full) + 2 * T.n_part;
    if (nb > 0)
      hipLaunchKernelGGL(k_theta_rb<2>, dim3(nb), dim3(128), 0, s, a, st, r, nrow, T.full_list, T.n_full);
  } else {
    const int32_t nrow = ((a.n_abc + 3) / 4 + 7) / 8 * 8;
    const int32_t nb = nrow + list_blocks(T.n_full) + 4 * T.n_part;
    if (nb > 0)
      hipLaunchKernelGGL(k_theta_rb<4>, dim3(nb), dim3(64), 0, s, a, st, r, nrow, T.full_list, T.n_full);
  }
  return hipGetLastError();
}

hipError_t launch_abc_theta(const StepArgs &a, const StepState &st, const AbcPack &p, hipStream_t s) {
  if (a.n_abc == 0) return hipSuccess;
  hipLaunchKernelGGL(k_abc_theta, dim3(abc_blocks_xcd(a)), dim3(kBlock), 0, s, a, st, p);
  return hipGetLastError();
}

hipError_t launch_abc_pack(int32_t n, const int32_t *slot, const int8_t *kind, const StepState &st,
                           const RingU &r, double4 *cur, double4 *prev, double2 *th, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_abc_pack, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, slot, kind,
                     st, r, cur, prev, th);
  return hipGetLastError();
}

hipError_t launch_idw(const StepArgs &a, const StepState &st, hipStream_t s) {
  if (abc_blocks(a) == 0) return hipSuccess;
  hipLaunchKernelGGL(k_idw, dim3(abc_blocks(a)), dim3(kBlock), 0, s, a, st);
  return hipGetLastError();
}

hipError_t launch_force(const StepArgs &a, const StepState &st, bool fast, hipStream_t s) {
  if (fast) {
    const int32_t nf = a.tile.n_full;
    if (a.tile.force_kernel == kBlocked) {
      if (nf > 0 && a.tile.force_nr == 2)
        hipLaunchKernelGGL(k_force_rb<2>, dim3(list_blocks(nf)), dim3(128), 0, s, a, st,
                           a.tile.full_list, nf);
      else if (nf > 0)
        hipLaunchKernelGGL(k_force_rb<4>, dim3(list_blocks(nf)), dim3(64), 0, s, a, st,
                           a.tile.full_list, nf);
    } else if (a.tile.force_kernel == kUnrolled) {
      if (nf > 0)
        hipLaunchKernelGGL(k_force_t68<true>, dim3(list_blocks(nf)), dim3(kBlock), 0, s, a, st,
                           a.tile.full_list, nf);
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

hipError_t launch_force_deep(const StepArgs &a, const StepState &st, const int32_t *list,
                             int32_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_force_deep, dim3(list_blocks(count)), dim3(128), 0, s, a, st, list, count);
  return hipGetLastError();
}

hipError_t launch_force_list(const StepArgs &a, const StepState &st, const int32_t *list,
                             int32_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  if (a.tile.force_nr == 2)
    hipLaunchKernelGGL(k_force_rb<2>, dim3(list_blocks(count)), dim3(128), 0, s, a, st, list, count);
  else
    hipLaunchKernelGGL(k_force_rb<4>, dim3(list_blocks(count)), dim3(64), 0, s, a, st, list, count);
  return hipGetLastError();
}

hipError_t launch_theta_list(const StepArgs &a, const StepState &st, const int32_t *list,
                             int32_t count, hipStream_t s) {
  const int32_t np = a.tile.n_part;
  if (a.tile.theta_nr == 2) {
    const int32_t nb = list_blocks(count) + 2 * np;
    if (nb > 0)
      hipLaunchKernelGGL(k_theta_rb<2>, dim3(nb), dim3(128), 0, s, a, st, AbcPack{}, 0, list, count);
  } else {
    const int32_t nb = list_blocks(count) + 4 * np;
    if (nb > 0)
      hipLaunchKernelGGL(k_theta_rb<4>, dim3(nb), dim3(64), 0, s, a, st, AbcPack{}, 0, list, count);
  }
  return hipGetLastError();
}

hipError_t launch_force_part(const StepArgs &a, const StepState &st, hipStream_t s,
                             bool predict_abc) {
  if (a.tile.force_kernel == kGeneric || a.tile.n_part == 0) return hipSuccess;
  // both families take the partial tiles one node per thread (see launch_theta_part)
  if (predict_abc)
    hipLaunchKernelGGL((k_force_t68<false, true>), dim3(list_blocks(a.tile.n_part)), dim3(kBlock),
                       0, s, a, st, a.tile.part_list, a.tile.n_part);
  else
    hipLaunchKernelGGL(k_force_t68<false>, dim3(list_blocks(a.tile.n_part)), dim3(kBlock), 0, s, a,
                       st, a.tile.part_list, a.tile.n_part);
  return hipGetLastError();
}

hipError_t launch_abc_uv(const StepArgs &a, const StepState &st, const AbcPack &p, hipStream_t s) {
  if (a.n_abc == 0) return hipSuccess;
  hipLaunchKernelGGL(k_abc_uv, dim3(abc_blocks_xcd(a)), dim3(kBlock), 0, s, a, st, p);
  return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) k_u_from_ring(int32_t n, const int8_t *kind,
                                                        const double *px, const double *py,
                                                        double *ux, double *uy) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || kind[i] != 1) return;
  ux[i] = px[i];
  uy[i] = py[i];
}

hipError_t launch_u_from_ring(const StepArgs &a, const double *px, const double *py, double *ux,
                              double *uy, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_u_from_ring, dim3(blocks_for(a.n)), dim3(kBlock), 0, s, a.n, a.kind, px, py,
                     ux, uy);
  return hipGetLastError();
}

hipError_t launch_abc_predict(const StepArgs &a, const StepState &st, hipStream_t s) {
  if (a.n_abc == 0) return hipSuccess;
  hipLaunchKernelGGL(k_abc_predict, dim3(blocks_for(a.n_abc)), dim3(kBlock), 0, s, a, st);
  return hipGetLastError();
}

hipError_t launch_tile_check(const StepArgs &a, const int16_t *bond_of, int32_t *d_mismatch,
                             hipStream_t s) {
  if (a.tile.n_tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tile_check, dim3(a.tile.n_tiles), dim3(kBlock), 0, s, a, bond_of,
                     d_mismatch);
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
                            const double *theta, const int32_t *perm, DiagPartial *partials,
                            DiagPartial *out, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_tile, dim3(a.tile.n_tiles), dim3(kBlock), 0, s, a, ux, uy, vx, vy, ax,
                     ay, theta, perm, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !out) return e;
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kBlock), 0, s, partials, a.tile.n_tiles, out);
  return hipGetLastError();
}

hipError_t launch_diag_final_inst(const DiagPartial *partials, const int32_t *tile_off,
                                  int32_t n_inst, DiagPartial *out, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_final_inst, dim3(n_inst), dim3(kBlock), 0, s, partials, tile_off, out);
  return hipGetLastError();
}

}  // namespace lps
