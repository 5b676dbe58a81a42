// HIP kernels of the LPS-ABC time loop for gfx950 (MI355X).
//
// One time step is two launches:
//   dilatation pass  theta_i for non-ABC nodes (thread per node, SELL-64 horizons) and, in the
//                    same grid, G_theta at ABC nodes (one wavefront per ABC row);
//   force pass       force density for every node plus the fused epilogue: velocity-Verlet
//                    corrector, G_u / G_v at ABC rows (one wavefront per row) and the next step's
//                    predictor, written to ping-pong buffers so that no thread overwrites a value
//                    another thread of the same launch still gathers.
// The output block (energy, momentum, maxima) is a third, every-30-steps pass with a deterministic
// two-level reduction.
//
// Two arithmetic variants share the kernels:
//   exact  per-bond arithmetic in the reference's operation order (src/main.cpp:736-874), with
//          floating-point contraction off, so results match a CPU build without FMA bit for bit
//          given the same neighbour order;
//   fast   bond geometry from a lattice-offset table (unit vector, |xi| beta dV, beta dV),
//          per-node coefficients hoisted (2/m_i, 8 mu/m_i, 2 (lambda-mu) theta_i/m_i), FMA on.
//          No square root or division per bond.
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

// ---------------------------------------------------------------------------------------------
// Neighbour access
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

__device__ __forceinline__ int geo_index(int32_t li, int32_t lj, int K) {
  const int dxi = (lj >> 16) - (li >> 16);
  const int dyi = (lj & 0xffff) - (li & 0xffff);
  return (dxi + K) * (2 * K + 1) + (dyi + K);
}

// Stage the lattice-offset table in LDS (at most (2*6+1)^2 entries).
constexpr int kMaxGeo = 169;  // K <= 6

__device__ __forceinline__ void load_geo(const StepArgs &a, Geo *sgeo) {
  const int ng = (2 * a.geo_k + 1) * (2 * a.geo_k + 1);
  for (int t = threadIdx.x; t < ng; t += blockDim.x) sgeo[t] = a.geo[t];
  __syncthreads();
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
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

// ---------------------------------------------------------------------------------------------
// Dilatation of one node from a field (exact), sequential over the sorted horizon.
// ---------------------------------------------------------------------------------------------
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

__device__ __forceinline__ double theta_node_fast(const StepArgs &a, const Geo *sgeo,
                                                  const double *fx, const double *fy, int32_t i) {
  const int32_t li = a.lat[i];
  const double ui = fx[i], vi = fy[i];
  double sum = 0.0;
  const SellRow r = sell_row(a, i);
  for (int k = 0; k < r.w; k++) {
    const int32_t j = r[k];
    if (j < 0) break;
    const Geo g = sgeo[geo_index(li, a.lat[j], a.geo_k)];
    sum += g.g1 * ((fx[j] - ui) * g.hx + (fy[j] - vi) * g.hy);
  }
  return a.c2m[i] * sum;
}

// theta of a whole field (exact order), used for the theta-history initialisation
// (src/main.cpp:664-682).
__global__ void __launch_bounds__(kBlock) k_theta_field(StepArgs a, const double *fx,
                                                        const double *fy, double *theta) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  theta[i] = theta_node_exact(a, fx, fy, i);
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
// Dilatation pass (src/main.cpp:736-814)
// ---------------------------------------------------------------------------------------------
template <bool EXACT>
__global__ void __launch_bounds__(kBlock) k_theta(StepArgs a, StepState st, int32_t n_int_blocks) {
  __shared__ Geo sgeo[EXACT ? 1 : kMaxGeo];
  __shared__ double sterm[kWavesPerBlock][EXACT ? 64 : 1];
  __shared__ int8_t sflag[kWavesPerBlock][EXACT ? 64 : 1];
  if (!EXACT) load_geo(a, sgeo);
  const int32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if ((int32_t)blockIdx.x < n_int_blocks) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n_int) return;
    const double th = EXACT ? theta_node_exact(a, st.px, st.py, i)
                            : theta_node_fast(a, sgeo, st.px, st.py, i);
    st.theta[i] = th;
    if (!EXACT) st.acoef[i] = (a.lam - a.mu) * a.c2m[i] * th;
    return;
  }
  // one wavefront per ABC row
  const int32_t row = (blockIdx.x - n_int_blocks) * kWavesPerBlock + wid;
  if (row >= a.n_abc) return;
  const int32_t node = a.n_int + row;
  double th;
  if (a.theta_method == 1) {
    // G_theta: theta_a = sum_k G[2k] ThN[c_k] + G[2k+1] ThNm1[c_k] (src/main.cpp:795-812)
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
        sterm[wid][lane] = t;
        sflag[wid][lane] = (k >= 0);
        __threadfence_block();
        if (lane == 0) {
          const int lim = min(64, a.M - base);
          for (int q = 0; q < lim; q++)
            if (sflag[wid][q]) acc = add_exact(acc, sterm[wid][q]);
        }
        __threadfence_block();
      } else {
        acc += wave_sum(t);
      }
    }
    th = acc;
  } else {
    // IDW: the direct value stays where no non-ABC neighbour exists (src/main.cpp:787-789);
    // the override itself runs in k_idw once every interior theta is final.
    th = (lane == 0) ? theta_node_exact(a, st.px, st.py, node) : 0.0;
  }
  if (lane == 0) {
    st.theta[node] = th;
    if (!EXACT) st.acoef[node] = (a.lam - a.mu) * a.c2m[node] * th;
  }
}

// IDW dilatation at ABC rows (src/main.cpp:769-790), one wavefront per row: per-bond weights in
// parallel, sums in horizon order by lane 0.
__global__ void __launch_bounds__(kBlock) k_idw(StepArgs a, StepState st, int32_t fast) {
  __shared__ double swt[kWavesPerBlock][kMaxHorizon];
  __shared__ double sval[kWavesPerBlock][kMaxHorizon];
  const int32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int32_t row = blockIdx.x * kWavesPerBlock + wid;
  if (row >= a.n_abc) return;
  const int32_t node = a.n_int + row;
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
      double th;
      {
#pragma clang fp contract(off)
        th = theta_avg / w_sum;
      }
      st.theta[node] = th;
      if (fast) st.acoef[node] = (a.lam - a.mu) * a.c2m[node] * th;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Force pass + epilogue (src/main.cpp:821-974 and the next predictor :725-728)
// ---------------------------------------------------------------------------------------------
template <bool EXACT>
__device__ __forceinline__ void force_node(const StepArgs &a, const StepState &st, const Geo *sgeo,
                                           int32_t i, double *fx_out, double *fy_out) {
  const double pxi = st.px[i], pyi = st.py[i];
  const SellRow r = sell_row(a, i);
  if (EXACT) {
    const double m_i = a.m[i];
    if (m_i <= 0.0) { *fx_out = 0.0; *fy_out = 0.0; return; }
    const double xi0 = a.x[i], yi0 = a.y[i], th_i = st.theta[i];
    double F0 = 0.0, F1 = 0.0;
    for (int k = 0; k < r.w; k++) {
      const int32_t j = r[k];
      if (j < 0) break;
      double fx, fy;
      if (force_term_exact(a.x[j] - xi0, a.y[j] - yi0, st.px[j] - pxi, st.py[j] - pyi, m_i,
                           a.m[j], th_i, st.theta[j], a.lam, a.mu, a.delta, a.dx, a.dy, &fx,
                           &fy)) {
        F0 = add_exact(F0, fx);
        F1 = add_exact(F1, fy);
      }
    }
    {
#pragma clang fp contract(off)
      *fx_out = F0 / a.rho;
      *fy_out = F1 / a.rho;
    }
  } else {
    const double bi = a.bcoef[i];
    if (bi < 0.0) { *fx_out = 0.0; *fy_out = 0.0; return; }
    const double ai = st.acoef[i];
    const int32_t li = a.lat[i];
    double F0 = 0.0, F1 = 0.0;
    for (int k = 0; k < r.w; k++) {
      const int32_t j = r[k];
      if (j < 0) break;
      const double bj = a.bcoef[j];
      if (bj < 0.0) continue;
      const Geo g = sgeo[geo_index(li, a.lat[j], a.geo_k)];
      const double e = (st.px[j] - pxi) * g.hx + (st.py[j] - pyi) * g.hy;
      const double s = (ai + st.acoef[j]) * g.g1 + (bi + bj) * e * g.g2;
      F0 += s * g.hx;
      F1 += s * g.hy;
    }
    *fx_out = F0 / a.rho;
    *fy_out = F1 / a.rho;
  }
}

template <bool EXACT>
__global__ void __launch_bounds__(kBlock) k_force(StepArgs a, StepState st, int32_t n_int_blocks) {
  __shared__ Geo sgeo[EXACT ? 1 : kMaxGeo];
  __shared__ double sf[kWavesPerBlock][2][EXACT ? kMaxHorizon : 1];
  __shared__ double sc[kWavesPerBlock][4][EXACT ? 64 : 1];
  __shared__ int8_t sflag[kWavesPerBlock][EXACT ? kMaxHorizon : 1];
  if (!EXACT) load_geo(a, sgeo);
  const int32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const double dt = a.dt;
  if ((int32_t)blockIdx.x < n_int_blocks) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n_int) return;
    double a1x, a1y;
    force_node<EXACT>(a, st, sgeo, i, &a1x, &a1y);
    const double u1x = st.px[i], u1y = st.py[i];
    double v1x, v1y, p1x, p1y;
    if (EXACT) {
      verlet_exact(u1x, st.vx[i], st.ax[i], a1x, dt, &v1x, &p1x);
      verlet_exact(u1y, st.vy[i], st.ay[i], a1y, dt, &v1y, &p1y);
    } else {
      const double h = 0.5 * dt;
      v1x = st.vx[i] + h * (st.ax[i] + a1x);
      v1y = st.vy[i] + h * (st.ay[i] + a1y);
      p1x = u1x + dt * (v1x + h * a1x);
      p1y = u1y + dt * (v1y + h * a1y);
    }
    st.ax1[i] = a1x; st.ay1[i] = a1y;
    st.ux1[i] = u1x; st.uy1[i] = u1y;
    st.vx1[i] = v1x; st.vy1[i] = v1y;
    st.pnx[i] = p1x; st.pny[i] = p1y;
    return;
  }
  // one wavefront per ABC row
  const int32_t row = (blockIdx.x - n_int_blocks) * kWavesPerBlock + wid;
  if (row >= a.n_abc) return;
  const int32_t node = a.n_int + row;
  // force on the ABC node: bonds in parallel, summed in horizon order
  double F0 = 0.0, F1 = 0.0;
  const double m_i = a.m[node];
  {
    const SellRow r = sell_row(a, node);
    const double pxi = st.px[node], pyi = st.py[node];
    if (EXACT) {
      const double xi0 = a.x[node], yi0 = a.y[node], th_i = st.theta[node];
      for (int k = lane; k < r.w; k += 64) {
        const int32_t j = r[k];
        double fx = 0.0, fy = 0.0;
        bool ok = false;
        if (j >= 0 && m_i > 0.0)
          ok = force_term_exact(a.x[j] - xi0, a.y[j] - yi0, st.px[j] - pxi, st.py[j] - pyi, m_i,
                                a.m[j], th_i, st.theta[j], a.lam, a.mu, a.delta, a.dx, a.dy, &fx,
                                &fy);
        sf[wid][0][k] = fx;
        sf[wid][1][k] = fy;
        sflag[wid][k] = ok;
      }
      __threadfence_block();
      if (lane == 0 && m_i > 0.0) {
        for (int k = 0; k < r.w; k++) {
          if (r[k] < 0) break;
          if (!sflag[wid][k]) continue;
          F0 = add_exact(F0, sf[wid][0][k]);
          F1 = add_exact(F1, sf[wid][1][k]);
        }
      }
      __threadfence_block();
    } else {
      const double bi = a.bcoef[node];
      if (bi >= 0.0) {
        const double ai = st.acoef[node];
        const int32_t li = a.lat[node];
        for (int k = lane; k < r.w; k += 64) {
          const int32_t j = r[k];
          if (j < 0) continue;
          const double bj = a.bcoef[j];
          if (bj < 0.0) continue;
          const Geo g = sgeo[geo_index(li, a.lat[j], a.geo_k)];
          const double e = (st.px[j] - pxi) * g.hx + (st.py[j] - pyi) * g.hy;
          const double s = (ai + st.acoef[j]) * g.g1 + (bi + bj) * e * g.g2;
          F0 += s * g.hx;
          F1 += s * g.hy;
        }
        F0 = wave_sum(F0);
        F1 = wave_sum(F1);
      }
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
      sc[wid][0][lane] = t0;
      sc[wid][1][lane] = t1;
      sc[wid][2][lane] = t2;
      sc[wid][3][lane] = t3;
      sflag[wid][lane] = (k >= 0);
      __threadfence_block();
      if (lane == 0) {
        const int lim = min(64, a.M - base);
        for (int q = 0; q < lim; q++) {
          if (!sflag[wid][q]) continue;
          Ux = add_exact(Ux, sc[wid][0][q]);
          Uy = add_exact(Uy, sc[wid][1][q]);
          Vx = add_exact(Vx, sc[wid][2][q]);
          Vy = add_exact(Vy, sc[wid][3][q]);
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

// ---------------------------------------------------------------------------------------------
// Output block (src/main.cpp:986-1062): per-node terms, block partials, final reduction.
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

template <bool EXACT>
__global__ void __launch_bounds__(kBlock) k_diag(StepArgs a, const double *ux, const double *uy,
                                                 const double *vx, const double *vy,
                                                 const double *ax, const double *ay,
                                                 const double *theta, const int32_t *perm,
                                                 double exclude_r, int32_t per_block,
                                                 DiagPartial *partials) {
  __shared__ DiagPartial sh[kBlock];
  __shared__ Geo sgeo[EXACT ? 1 : kMaxGeo];
  if (!EXACT) load_geo(a, sgeo);
  DiagPartial d;
  diag_identity(d);
  const int32_t base = blockIdx.x * per_block;
  const int32_t end = min(a.n, base + per_block);
  for (int32_t i = base + threadIdx.x; i < end; i += kBlock) {
#pragma clang fp contract(off)
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
    if (i >= a.n_int) {  // ABC rows: boundary norms (src/main.cpp:1050-1060)
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
    if (EXACT) {
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
    } else {
      const int32_t li = a.lat[i];
      const double c = 4.0 * a.mu / m_i;
      double acc = 0.0;
      for (int k = 0; k < row.w; k++) {
        const int32_t n = row[k];
        if (n < 0) break;
        const Geo g = sgeo[geo_index(li, a.lat[n], a.geo_k)];
        const double e = (ux[n] - ux[i]) * g.hx + (uy[n] - uy[i]) * g.hy;
        // |xi| = g1 / g2 where beta > 0; bonds with beta == 0 carry g2 == 0 and add nothing
        const double xin = (g.g2 != 0.0) ? g.g1 / g.g2 : 0.0;
        const double ed = e - 0.5 * th * xin;
        acc += ed * ed * g.g2;
      }
      W_i += c * acc;
    }
    const double ep = W_i * a.dx * a.dy;
    d.ep += ep;
    d.energy += ek + ep;
  }
  block_reduce_diag(d, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = d;
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

hipError_t launch_predict(const StepArgs &a, const double *ux, const double *uy, const double *vx,
                          const double *vy, const double *ax, const double *ay, double *px,
                          double *py, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_predict, dim3(blocks_for(a.n)), dim3(kBlock), 0, s, a.n, a.dt, ux, uy, vx,
                     vy, ax, ay, px, py);
  return hipGetLastError();
}

hipError_t launch_theta(const StepArgs &a, const StepState &st, bool exact, hipStream_t s) {
  const int32_t nib = blocks_for(a.n_int);
  const int32_t nab = (a.n_abc + kWavesPerBlock - 1) / kWavesPerBlock;
  if (nib + nab == 0) return hipSuccess;
  if (exact)
    hipLaunchKernelGGL(k_theta<true>, dim3(nib + nab), dim3(kBlock), 0, s, a, st, nib);
  else
    hipLaunchKernelGGL(k_theta<false>, dim3(nib + nab), dim3(kBlock), 0, s, a, st, nib);
  return hipGetLastError();
}

hipError_t launch_idw(const StepArgs &a, const StepState &st, bool exact, hipStream_t s) {
  const int32_t nab = (a.n_abc + kWavesPerBlock - 1) / kWavesPerBlock;
  if (nab == 0) return hipSuccess;
  hipLaunchKernelGGL(k_idw, dim3(nab), dim3(kBlock), 0, s, a, st, exact ? 0 : 1);
  return hipGetLastError();
}

hipError_t launch_force(const StepArgs &a, const StepState &st, bool exact, hipStream_t s) {
  const int32_t nib = blocks_for(a.n_int);
  const int32_t nab = (a.n_abc + kWavesPerBlock - 1) / kWavesPerBlock;
  if (nib + nab == 0) return hipSuccess;
  if (exact)
    hipLaunchKernelGGL(k_force<true>, dim3(nib + nab), dim3(kBlock), 0, s, a, st, nib);
  else
    hipLaunchKernelGGL(k_force<false>, dim3(nib + nab), dim3(kBlock), 0, s, a, st, nib);
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

// Partials: one per block of kDiagPerBlock nodes.
constexpr int32_t kDiagPerBlock = 4096;

int32_t diag_partials(int32_t n) { return std::max<int32_t>(1, (n + kDiagPerBlock - 1) / kDiagPerBlock); }

hipError_t launch_diag(const StepArgs &a, const double *ux, const double *uy, const double *vx,
                       const double *vy, const double *ax, const double *ay, const double *theta,
                       const int32_t *perm, double exclude_r, bool exact, DiagPartial *partials,
                       int32_t n_partials, DiagPartial *out, hipStream_t s) {
  if (exact)
    hipLaunchKernelGGL(k_diag<true>, dim3(n_partials), dim3(kBlock), 0, s, a, ux, uy, vx, vy, ax,
                       ay, theta, perm, exclude_r, kDiagPerBlock, partials);
  else
    hipLaunchKernelGGL(k_diag<false>, dim3(n_partials), dim3(kBlock), 0, s, a, ux, uy, vx, vy, ax,
                       ay, theta, perm, exclude_r, kDiagPerBlock, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kBlock), 0, s, partials, n_partials, out);
  return hipGetLastError();
}

}  // namespace lps
