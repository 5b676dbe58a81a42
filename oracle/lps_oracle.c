/* CPU oracle of the LPS-ABC solver's time loop (restatement of src/main.cpp of the reference).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP implementation. Only tests, the smoke entry and
 * the bench's CPU-baseline leg load it; the product path never does.
 *
 * Built with -ffp-contract=off: the reference is compiled for baseline x86-64 (CMakeLists.txt:15,
 * no -march), so it has no fused multiply-add, and every expression below keeps the reference's
 * association order.
 *
 * How each function is pinned (the checks live in tests/):
 *   oracle_build_horizons  the reference kd-tree build (src/helpers/kd_tree_interface.h:41-126,
 *                          compiled from the reference sources) as golden neighbour sets; brute
 *                          force on tiny inputs; lattice counts 69 = 1 + 68 and symmetry.
 *   m_i, m_ref             closed form over the 68-offset template, m_ref = 25.95427624970138.
 *   dilatation             affine fields (theta = 2 eps for u = eps x), rigid motions (theta = 0).
 *   force                  affine fields and rigid motions give zero interior force; sum F = 0;
 *                          plane-wave known answer F = M^LPS u (preprocessor/main.cpp:238-246).
 *   time loop, ABC, energy the reference solver binary (src/main.cpp compiled from its sources)
 *                          on generated cases: .energy/.verify series and final VTK fields.
 *   prenotch visibility    the reference solver on a notched case (series and fields); the
 *                          intersection test against exact rational orientation on random and
 *                          degenerate (collinear, touching) segments.
 */
#include "lps_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SUBSTEP_NO 30 /* src/main.cpp:34 */

static char g_err[512];
const char *oracle_error(void) { return g_err; }

struct oracle_ctx {
  oracle_problem p;
  /* owned copies of the inputs */
  double *nodes, *ic, *icd, *Gu, *Gv, *Gth;
  int32_t *bctypes, *abc, *clouds;
  /* horizons, CSR with self first */
  int64_t *hp;
  int32_t *hc;
  double *m;
  double m_ref, Lambda, Mu, Kappa, exclude_r, E0;
  /* state: the reference's Un1/Un/Unm1, Udn1/Udn/Udnm1, Uddn1/Uddn and theta history */
  double *Un1, *Un, *Unm1, *Udn1, *Udn, *Udnm1, *Uddn1, *Uddn;
  double *Theta, *ThetaN, *ThetaNm1;
  int64_t s;
  int threads;
};

/* ------------------------------------------------------------------------------------------- */
/* Horizons (src/helpers/kd_tree_interface.h:41-126 with the libkdtree++ box search of
 * libraries/libkdtree++/kdtree++/region.hpp:44-53, 86-95).                                    */
/* ------------------------------------------------------------------------------------------- */
typedef struct { double d; int32_t j; } cand_t;

static int cand_cmp(const void *a, const void *b) {
  const cand_t *x = (const cand_t *)a, *y = (const cand_t *)b;
  if (x->d < y->d) return -1;
  if (x->d > y->d) return 1;
  return (x->j > y->j) - (x->j < y->j);
}

/* Euclidean distance as Node::distance_euclidean computes it (kd_tree_interface.h:30-35). */
static inline double node_dist(double xj, double yj, double xi, double yi) {
  double x = xj - xi;
  double y = yj - yi;
  return sqrt(x * x + y * y);
}

int64_t oracle_build_horizons(int32_t N, const double *nodes, double r, int32_t max_len,
                              int64_t *row_ptr, int32_t *cols) {
  if (N <= 0) { if (row_ptr) row_ptr[0] = 0; return 0; }
  double xmin = nodes[0], ymin = nodes[1], xmax = nodes[0], ymax = nodes[1];
  for (int32_t i = 1; i < N; i++) {
    double x = nodes[2 * i], y = nodes[2 * i + 1];
    if (x < xmin) xmin = x;
    if (x > xmax) xmax = x;
    if (y < ymin) ymin = y;
    if (y > ymax) ymax = y;
  }
  /* cell grid of side r; a candidate scan of +-2 cells covers the disk with margin */
  const double h = r > 0 ? r : 1.0;
  int64_t ncx = (int64_t)floor((xmax - xmin) / h) + 1, ncy = (int64_t)floor((ymax - ymin) / h) + 1;
  int64_t ncell = ncx * ncy;
  int64_t *cstart = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
  int32_t *cidx = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
  int64_t *ccell = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
  for (int32_t i = 0; i < N; i++) {
    int64_t cx = (int64_t)floor((nodes[2 * i] - xmin) / h), cy = (int64_t)floor((nodes[2 * i + 1] - ymin) / h);
    ccell[i] = cx * ncy + cy;
    cstart[ccell[i] + 1]++;
  }
  for (int64_t c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)ncell);
  memcpy(fill, cstart, sizeof(int64_t) * (size_t)ncell);
  for (int32_t i = 0; i < N; i++) cidx[fill[ccell[i]]++] = i; /* ascending index per cell */
  free(fill);

  int64_t *cnt = (int64_t *)calloc((size_t)N, sizeof(int64_t));
  int overflow = 0;
  for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel
    {
      int cap = 1024;
      cand_t *buf = (cand_t *)malloc(sizeof(cand_t) * (size_t)cap);
#pragma omp for schedule(static)
      for (int32_t i = 0; i < N; i++) {
        const double xi = nodes[2 * i], yi = nodes[2 * i + 1];
        /* box of the kd-tree range query: [x - r, x + r] inclusive */
        const double lox = xi - r, hix = xi + r, loy = yi - r, hiy = yi + r;
        int64_t cx = (int64_t)floor((xi - xmin) / h), cy = (int64_t)floor((yi - ymin) / h);
        int n = 0;
        for (int64_t ax = cx - 2; ax <= cx + 2; ax++) {
          if (ax < 0 || ax >= ncx) continue;
          for (int64_t ay = cy - 2; ay <= cy + 2; ay++) {
            if (ay < 0 || ay >= ncy) continue;
            int64_t c = ax * ncy + ay;
            for (int64_t q = cstart[c]; q < cstart[c + 1]; q++) {
              int32_t j = cidx[q];
              double xj = nodes[2 * j], yj = nodes[2 * j + 1];
              if (xj < lox || hix < xj || yj < loy || hiy < yj) continue;
              double d = node_dist(xj, yj, xi, yi);
              if (d <= r) {
                if (n == cap) { cap *= 2; buf = (cand_t *)realloc(buf, sizeof(cand_t) * (size_t)cap); }
                buf[n].d = d;
                buf[n].j = j;
                n++;
              }
            }
          }
        }
        if (pass == 0) {
          cnt[i] = n;
          if (n > max_len) {
#pragma omp atomic write
            overflow = 1;
          }
        } else if (cols) {
          qsort(buf, (size_t)n, sizeof(cand_t), cand_cmp);
          for (int k = 0; k < n; k++) cols[row_ptr[i] + k] = buf[k].j;
        }
      }
      free(buf);
    }
    if (pass == 0) {
      if (overflow) break;
      if (row_ptr) {
        row_ptr[0] = 0;
        for (int32_t i = 0; i < N; i++) row_ptr[i + 1] = row_ptr[i] + cnt[i];
      }
      if (!cols || !row_ptr) break;
    }
  }
  int64_t nnz = 0;
  for (int32_t i = 0; i < N; i++) nnz += cnt[i];
  free(cnt); free(cstart); free(cidx); free(ccell);
  if (overflow) {
    snprintf(g_err, sizeof g_err, "found_node_no > max_node_no; increase max_node_no.");
    return -1;
  }
  return nnz;
}

/* ------------------------------------------------------------------------------------------- */
/* Pointwise pieces of src/main.cpp                                                            */
/* ------------------------------------------------------------------------------------------- */
static inline double beta_fn(double xi, double R, double dx) { /* src/main.cpp:100-107 */
  if (xi <= R - dx / 2.0) return 1.0;
  else if (xi <= R + dx / 2.0) return (R + dx / 2.0 - xi) / dx;
  else return 0.0;
}

#define X(i) c->nodes[2 * (i)]
#define Y(i) c->nodes[2 * (i) + 1]

/* Dilatation of node i from field u (src/main.cpp:736-761; also the theta-history init :664-682). */
static inline double theta_node(const oracle_ctx *c, const double *u, int32_t i) {
  const double R = c->p.delta, dx = c->p.dx, dy = c->p.dy;
  double theta_sum = 0.0;
  for (int64_t q = c->hp[i] + 1; q < c->hp[i + 1]; q++) {
    int32_t n = c->hc[q];
    const double xi_x = X(n) - X(i);
    const double xi_y = Y(n) - Y(i);
    const double xi_norm = sqrt(xi_x * xi_x + xi_y * xi_y);
    const double eta_x = u[2 * n] - u[2 * i];
    const double eta_y = u[2 * n + 1] - u[2 * i + 1];
    const double e = (eta_x * xi_x + eta_y * xi_y) / xi_norm;
    const double w = 1.0;
    const double b = beta_fn(xi_norm, R, dx);
    theta_sum += w * xi_norm * e * b * dx * dy;
  }
  return (c->m[i] > 0.0) ? (2.0 / c->m[i]) * theta_sum : 0.0;
}

/* Force density of node i (src/main.cpp:821-874). */
static inline void force_node(const oracle_ctx *c, const double *u, const double *th, int32_t i,
                              double *acc) {
  const double R = c->p.delta, dx = c->p.dx, dy = c->p.dy;
  const double Lambda = c->Lambda, Mu = c->Mu;
  double F0 = 0.0, F1 = 0.0;
  const double m_i = c->m[i];
  if (m_i <= 0.0) {
    acc[2 * i] = 0.0;
    acc[2 * i + 1] = 0.0;
    return;
  }
  for (int64_t q = c->hp[i] + 1; q < c->hp[i + 1]; q++) {
    int32_t n = c->hc[q];
    const double xi_x = X(n) - X(i);
    const double xi_y = Y(n) - Y(i);
    const double xi_norm = sqrt(xi_x * xi_x + xi_y * xi_y);
    const double eta_x = u[2 * n] - u[2 * i];
    const double eta_y = u[2 * n + 1] - u[2 * i + 1];
    const double e = (eta_x * xi_x + eta_y * xi_y) / xi_norm;
    const double w = 1.0;
    const double b = beta_fn(xi_norm, R, dx);
    const double m_j = c->m[n];
    if (m_j <= 0.0) continue;
    const double t_i = (2.0 * (Lambda - Mu) / m_i) * th[i] * xi_norm * w + (8.0 * Mu / m_i) * w * e;
    const double t_j = (2.0 * (Lambda - Mu) / m_j) * th[n] * xi_norm * w + (8.0 * Mu / m_j) * w * e;
    const double scale = (t_i + t_j) * b * dx * dy;
    F0 += scale * xi_x / xi_norm;
    F1 += scale * xi_y / xi_norm;
  }
  acc[2 * i] = F0 / c->p.rho;
  acc[2 * i + 1] = F1 / c->p.rho;
}

/* ------------------------------------------------------------------------------------------- */
/* Prenotch visibility (src/main.cpp:119-140, 386-402)                                          */
/* ------------------------------------------------------------------------------------------- */
/* Side of point (px, py) relative to the directed segment (x1, y1) -> (x2, y2): the sign of the
 * cross product; on the segment's line, -1 before its start, +1 past its end, 0 on it
 * (src/main.cpp:119-132). */
static int side_of(double x1, double y1, double x2, double y2, double px, double py) {
  const double dx = x2 - x1, dy = y2 - y1;
  double qx = px - x1, qy = py - y1;
  double c = qx * dy - qy * dx;
  if (c == 0.0) {
    c = qx * dx + qy * dy;
    if (c > 0.0) {
      qx -= dx;
      qy -= dy;
      c = qx * dx + qy * dy;
      if (c < 0.0) c = 0.0;
    }
  }
  return (c > 0.0) - (c < 0.0);
}

/* src/main.cpp:134-140: each segment's end points lie on both sides of (or on) the other's. */
int oracle_segments_intersect(double x1, double y1, double x2, double y2, double x3, double y3,
                              double x4, double y4) {
  return side_of(x1, y1, x2, y2, x3, y3) * side_of(x1, y1, x2, y2, x4, y4) <= 0 &&
         side_of(x3, y3, x4, y4, x1, y1) * side_of(x3, y3, x4, y4, x2, y2) <= 0;
}

/* Drops every bond i -> j (self excluded) that a prenotch segment crosses, keeping the order of
 * the rest (src/main.cpp:386-402 marks them -1; every later loop skips -1 entries, which is the
 * same as removing them). Returns the new nnz. */
static int64_t cut_prenotch_bonds(int32_t N, const double *nodes, int32_t n_notch,
                                  const double *notch, int64_t *hp, int32_t *hc) {
  int64_t out = 0;
  for (int32_t i = 0; i < N; i++) {
    const int64_t b = hp[i], e = hp[i + 1];
    hp[i] = out;
    for (int64_t q = b; q < e; q++) {
      const int32_t j = hc[q];
      int cut = 0;
      if (q > b)
        for (int32_t k = 0; k < n_notch && !cut; k++)
          cut = oracle_segments_intersect(nodes[2 * (int64_t)i], nodes[2 * (int64_t)i + 1],
                                          nodes[2 * (int64_t)j], nodes[2 * (int64_t)j + 1],
                                          notch[4 * k], notch[4 * k + 1], notch[4 * k + 2],
                                          notch[4 * k + 3]);
      if (!cut) hc[out++] = j;
    }
  }
  hp[N] = out;
  return out;
}

/* ------------------------------------------------------------------------------------------- */
/* Setup (src/main.cpp:346-356, 408-486, 518-535, 598-684)                                     */
/* ------------------------------------------------------------------------------------------- */
static void *dup(const void *src, size_t bytes) {
  if (!src || !bytes) return NULL;
  void *d = malloc(bytes);
  memcpy(d, src, bytes);
  return d;
}

oracle_ctx *oracle_create(const oracle_problem *p, int32_t threads) {
  oracle_ctx *c = (oracle_ctx *)calloc(1, sizeof(oracle_ctx));
  const int32_t N = p->N, C = p->C, M = p->M;
  c->p = *p;
  c->threads = threads;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  c->nodes = (double *)dup(p->nodes, sizeof(double) * 2 * (size_t)N);
  c->bctypes = (int32_t *)dup(p->bctypes, sizeof(int32_t) * 2 * (size_t)N);
  c->ic = (double *)dup(p->ic, sizeof(double) * 2 * (size_t)N);
  c->icd = (double *)dup(p->icd, sizeof(double) * 2 * (size_t)N);
  if (C > 0) {
    c->abc = (int32_t *)dup(p->abc, sizeof(int32_t) * (size_t)C);
    c->clouds = (int32_t *)dup(p->clouds, sizeof(int32_t) * (size_t)C * M);
    c->Gu = (double *)dup(p->Gu, sizeof(double) * 2 * (size_t)C * 4 * M);
    c->Gv = (double *)dup(p->Gv, sizeof(double) * 2 * (size_t)C * 4 * M);
    if (p->Gth) c->Gth = (double *)dup(p->Gth, sizeof(double) * (size_t)C * 2 * M);
  }
  /* horizons, radius delta + dx/2 + 1e-5, at most 128 entries (src/main.cpp:35, 352) */
  const double search_r = p->delta + p->dx / 2.0 + 1e-5;
  c->hp = (int64_t *)malloc(sizeof(int64_t) * ((size_t)N + 1));
  int64_t nnz = oracle_build_horizons(N, c->nodes, search_r, 128, c->hp, NULL);
  if (nnz < 0) { oracle_destroy(c); return NULL; }
  c->hc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  oracle_build_horizons(N, c->nodes, search_r, 128, c->hp, c->hc);
  if (p->n_notch > 0 && p->notch)
    cut_prenotch_bonds(N, c->nodes, p->n_notch, p->notch, c->hp, c->hc);

  /* Lame constants (src/main.cpp:215-217) */
  c->Lambda = p->E * p->nu / ((1.0 + p->nu) * (1.0 - 2.0 * p->nu));
  c->Mu = p->E / (2.0 * (1.0 + p->nu));
  c->Kappa = c->Lambda + c->Mu;

  /* m_i (src/main.cpp:412-426) */
  c->m = (double *)malloc(sizeof(double) * (size_t)N);
  const double R = p->delta, dx = p->dx, dy = p->dy;
#pragma omp parallel for
  for (int32_t i = 0; i < N; i++) {
    double m_i = 0.0;
    for (int64_t q = c->hp[i] + 1; q < c->hp[i + 1]; q++) {
      int32_t n = c->hc[q];
      const double xi_x = X(n) - X(i);
      const double xi_y = Y(n) - Y(i);
      const double xi_norm = sqrt(xi_x * xi_x + xi_y * xi_y);
      const double w = 1.0;
      const double b = beta_fn(xi_norm, R, dx);
      m_i += w * xi_norm * xi_norm * b * dx * dy;
    }
    c->m[i] = m_i;
  }
  /* m_ref = max m_i over m_i > 0 (src/main.cpp:430-442) */
  double m_max = 0.0;
  for (int32_t i = 0; i < N; i++)
    if (c->m[i] > 0.0 && c->m[i] > m_max) m_max = c->m[i];
  c->m_ref = m_max;
  /* m_i := m_ref at ABC nodes (src/main.cpp:469-480) */
  if (p->abc_enabled && C > 0)
    for (int32_t i = 0; i < C; i++) c->m[c->abc[i]] = c->m_ref;

  /* energy/output exclusion radius (src/main.cpp:518-530) */
  double domainRadius = 0.0;
  for (int32_t i = 0; i < N; i++) {
    double r = sqrt(X(i) * X(i) + Y(i) * Y(i));
    if (r > domainRadius) domainRadius = r;
  }
  c->exclude_r = (p->abc_enabled && C > 0) ? domainRadius - R - dx / 100.0 : -1.0;

  /* state and initial conditions (src/main.cpp:598-634) */
  size_t n2 = sizeof(double) * 2 * (size_t)N;
  c->Un1 = (double *)calloc(1, n2); c->Un = (double *)calloc(1, n2); c->Unm1 = (double *)calloc(1, n2);
  c->Udn1 = (double *)calloc(1, n2); c->Udn = (double *)calloc(1, n2); c->Udnm1 = (double *)calloc(1, n2);
  c->Uddn1 = (double *)calloc(1, n2); c->Uddn = (double *)calloc(1, n2);
  c->Theta = (double *)calloc((size_t)N, sizeof(double));
  c->ThetaN = (double *)calloc((size_t)N, sizeof(double));
  c->ThetaNm1 = (double *)calloc((size_t)N, sizeof(double));
  for (int64_t i = 0; i < 2 * (int64_t)N; i++) {
    c->Udn[i] = c->icd[i];
    c->Un[i] = c->ic[i];
  }
  /* theta history from the IC, with m_i already overridden (src/main.cpp:664-682) */
  if (p->theta_method == 1) {
#pragma omp parallel for
    for (int32_t i = 0; i < N; i++) {
      c->ThetaN[i] = theta_node(c, c->Un, i);
      c->ThetaNm1[i] = 0.0;
    }
  }
  c->E0 = -1.0;
  c->s = 0;
  return c;
}

void oracle_destroy(oracle_ctx *c) {
  if (!c) return;
  free(c->nodes); free(c->bctypes); free(c->ic); free(c->icd); free(c->Gu); free(c->Gv);
  free(c->Gth); free(c->abc); free(c->clouds); free(c->hp); free(c->hc); free(c->m);
  free(c->Un1); free(c->Un); free(c->Unm1); free(c->Udn1); free(c->Udn); free(c->Udnm1);
  free(c->Uddn1); free(c->Uddn); free(c->Theta); free(c->ThetaN); free(c->ThetaNm1);
  free(c);
}

/* ------------------------------------------------------------------------------------------- */
/* Output block (src/main.cpp:986-1062)                                                        */
/* ------------------------------------------------------------------------------------------- */
static void diagnostics(oracle_ctx *c, const double *Un1, const double *Udn1, const double *Uddn1,
                        const double *Theta, oracle_diag *o) {
  const int32_t N = c->p.N;
  const double R = c->p.delta, dx = c->p.dx, dy = c->p.dy, Rho = c->p.rho;
  const double Kappa = c->Kappa, Mu = c->Mu;
  double energy = 0.0, energyK = 0.0, energyP = 0.0, Px = 0.0, Py = 0.0;
  double max_u = 0.0, max_v = 0.0, max_a = 0.0, theta_min = 1e30, theta_max = -1e30;
  int32_t max_u_node = -1, max_v_node = -1;
  for (int32_t i = 0; i < N; i++) {
    double x = X(i), y = Y(i);
    double r = sqrt(x * x + y * y);
    double u_mag = sqrt(Un1[2 * i] * Un1[2 * i] + Un1[2 * i + 1] * Un1[2 * i + 1]);
    double v_mag = sqrt(Udn1[2 * i] * Udn1[2 * i] + Udn1[2 * i + 1] * Udn1[2 * i + 1]);
    double a_mag = sqrt(Uddn1[2 * i] * Uddn1[2 * i] + Uddn1[2 * i + 1] * Uddn1[2 * i + 1]);
    if (u_mag > max_u) { max_u = u_mag; max_u_node = i; }
    if (v_mag > max_v) { max_v = v_mag; max_v_node = i; }
    max_a = (max_a < a_mag) ? a_mag : max_a; /* std::max(max_a, a_mag) */
    theta_min = (Theta[i] < theta_min) ? Theta[i] : theta_min;
    theta_max = (theta_max < Theta[i]) ? Theta[i] : theta_max;
    if (c->exclude_r > 0.0 && r > c->exclude_r) continue;
    Px += Rho * Udn1[2 * i] * dx * dy;
    Py += Rho * Udn1[2 * i + 1] * dx * dy;
    double vx = Udn1[2 * i], vy = Udn1[2 * i + 1];
    double ek = 0.5 * Rho * (vx * vx + vy * vy) * dx * dy;
    energyK += ek;
    const double m_i = c->m[i];
    double W_i = 0.5 * Kappa * Theta[i] * Theta[i];
    for (int64_t q = c->hp[i] + 1; q < c->hp[i + 1]; q++) {
      int32_t n = c->hc[q];
      const double xi_x = X(n) - X(i);
      const double xi_y = Y(n) - Y(i);
      const double xi_norm = sqrt(xi_x * xi_x + xi_y * xi_y);
      const double eta_x = Un1[2 * n] - Un1[2 * i];
      const double eta_y = Un1[2 * n + 1] - Un1[2 * i + 1];
      const double e = (eta_x * xi_x + eta_y * xi_y) / xi_norm;
      const double ed = e - Theta[i] * xi_norm / 2.0;
      const double w = 1.0;
      const double b = beta_fn(xi_norm, R, dx);
      W_i += (4.0 * Mu / m_i) * w * ed * ed * b * dx * dy;
    }
    double ep = W_i * dx * dy;
    energyP += ep;
    energy += ek + ep;
  }
  double uB_sq = 0.0, vB_sq = 0.0;
  int32_t nActive = 0;
  if (c->p.abc_enabled) {
    for (int32_t i = 0; i < c->p.C; i++) {
      int32_t node = c->abc[i];
      if (c->bctypes[2 * node] != 1) continue;
      nActive++;
      uB_sq += Un1[2 * node] * Un1[2 * node] + Un1[2 * node + 1] * Un1[2 * node + 1];
      vB_sq += Udn1[2 * node] * Udn1[2 * node] + Udn1[2 * node + 1] * Udn1[2 * node + 1];
    }
  }
  o->step = c->s;
  o->t = (double)c->s * c->p.dt;
  o->energy = energy; o->energy_k = energyK; o->energy_p = energyP;
  o->px = Px; o->py = Py;
  o->max_u = max_u; o->max_v = max_v; o->max_a = max_a;
  o->theta_max = theta_max; o->theta_min = theta_min;
  o->ub_rms = (nActive > 0) ? sqrt(uB_sq / nActive) : 0.0;
  o->vb_rms = (nActive > 0) ? sqrt(vB_sq / nActive) : 0.0;
  o->max_u_node = max_u_node; o->max_v_node = max_v_node;
  o->de_rel = 0.0;
}

/* Between steps the current state sits in Un/Udn/Uddn (after the rotation). */
void oracle_diagnostics(oracle_ctx *c, oracle_diag *out) {
  diagnostics(c, c->Un, c->Udn, c->Uddn, c->Theta, out);
}

/* ------------------------------------------------------------------------------------------- */
/* Time loop (src/main.cpp:718-1164)                                                           */
/* ------------------------------------------------------------------------------------------- */
#define SWAP(a, b) do { double *t_ = (a); (a) = (b); (b) = t_; } while (0)

int32_t oracle_run(oracle_ctx *c, int32_t nsteps, oracle_diag *recs, int32_t max_recs) {
  const int32_t N = c->p.N, C = c->p.C, M = c->p.M;
  const double dt = c->p.dt;
  const int abc_on = c->p.abc_enabled;
  int32_t nrec = 0;
  for (int32_t it = 0; it < nsteps; it++) {
    c->s++;
    double *Un1 = c->Un1, *Un = c->Un, *Unm1 = c->Unm1;
    double *Udn1 = c->Udn1, *Udn = c->Udn, *Udnm1 = c->Udnm1;
    double *Uddn1 = c->Uddn1, *Uddn = c->Uddn, *Theta = c->Theta;

    /* predictor (:725-728) */
#pragma omp parallel for
    for (int64_t i = 0; i < 2 * (int64_t)N; i++) Un1[i] = Un[i] + dt * (Udn[i] + (dt / 2.0) * Uddn[i]);

    /* dilatation (:736-761) */
#pragma omp parallel for
    for (int32_t i = 0; i < N; i++) Theta[i] = theta_node(c, Un1, i);

    /* ABC dilatation override (:764-814) */
    if (abc_on) {
      if (c->p.theta_method == 0) {
        for (int32_t a = 0; a < C; a++) { /* IDW (:769-790) */
          int32_t node = c->abc[a];
          if (c->bctypes[2 * node] != 1) continue;
          double theta_avg = 0.0, w_sum = 0.0;
          for (int64_t q = c->hp[node] + 1; q < c->hp[node + 1]; q++) {
            int32_t nb = c->hc[q];
            if (c->bctypes[2 * nb] == 1) continue;
            double rx = X(nb) - X(node);
            double ry = Y(nb) - Y(node);
            double dist = sqrt(rx * rx + ry * ry);
            double wt = 1.0 / (dist + 1e-30);
            if (c->p.plant == 1) wt = wt * wt; /* planted mistake (tests only) */
            theta_avg += wt * Theta[nb];
            w_sum += wt;
          }
          if (w_sum > 0.0) Theta[node] = theta_avg / w_sum;
        }
      } else {
        for (int32_t a = 0; a < C; a++) { /* G_theta (:795-812) */
          int32_t node = c->abc[a];
          if (c->bctypes[2 * node] != 1) continue;
          double theta_abc = 0.0;
          const double *g = c->Gth + (size_t)a * 2 * M;
          for (int32_t j = 0; j < M; j++) {
            int32_t k = c->clouds[(size_t)a * M + j];
            if (k == -1) continue;
            theta_abc += g[2 * j] * c->ThetaN[k] + g[2 * j + 1] * c->ThetaNm1[k];
          }
          Theta[node] = theta_abc;
        }
      }
    }

    /* force (:821-874) */
#pragma omp parallel for
    for (int32_t i = 0; i < N; i++) force_node(c, Un1, Theta, i, Uddn1);

    /* G_u (:883-916) */
    if (abc_on) {
      for (int32_t a = 0; a < C; a++) {
        int32_t node = c->abc[a];
        if (c->bctypes[2 * node] != 1) continue;
        const double *g0 = c->Gu + (size_t)(2 * a) * 4 * M, *g1 = g0 + 4 * M;
        Un1[2 * node] = 0.0;
        Un1[2 * node + 1] = 0.0;
        for (int32_t j = 0; j < M; j++) {
          int32_t k = c->clouds[(size_t)a * M + j];
          if (k == -1) continue;
          Un1[2 * node] += g0[4 * j] * Un[2 * k] + g0[4 * j + 1] * Un[2 * k + 1] +
                           g0[4 * j + 2] * Unm1[2 * k] + g0[4 * j + 3] * Unm1[2 * k + 1];
          Un1[2 * node + 1] += g1[4 * j] * Un[2 * k] + g1[4 * j + 1] * Un[2 * k + 1] +
                               g1[4 * j + 2] * Unm1[2 * k] + g1[4 * j + 3] * Unm1[2 * k + 1];
        }
      }
    }

    /* corrector (:922-925) */
#pragma omp parallel for
    for (int64_t i = 0; i < 2 * (int64_t)N; i++) Udn1[i] = Udn[i] + (dt / 2.0) * Uddn[i] + (dt / 2.0) * Uddn1[i];

    /* G_v (:933-974) */
    if (abc_on) {
      for (int32_t a = 0; a < C; a++) {
        int32_t node = c->abc[a];
        if (c->bctypes[2 * node] != 1) continue;
        const double *g0 = c->Gv + (size_t)(2 * a) * 4 * M, *g1 = g0 + 4 * M;
        Udn1[2 * node] = 0.0;
        Udn1[2 * node + 1] = 0.0;
        for (int32_t j = 0; j < M; j++) {
          int32_t k = c->clouds[(size_t)a * M + j];
          if (k == -1) continue;
          Udn1[2 * node] += g0[4 * j] * Udn[2 * k] + g0[4 * j + 1] * Udn[2 * k + 1] +
                            g0[4 * j + 2] * Udnm1[2 * k] + g0[4 * j + 3] * Udnm1[2 * k + 1];
          Udn1[2 * node + 1] += g1[4 * j] * Udn[2 * k] + g1[4 * j + 1] * Udn[2 * k + 1] +
                                g1[4 * j + 2] * Udnm1[2 * k] + g1[4 * j + 3] * Udnm1[2 * k + 1];
        }
      }
    }

    /* output (:981-1133) */
    int bad = 0;
    if (c->s % SUBSTEP_NO == 0) {
      oracle_diag d;
      diagnostics(c, Un1, Udn1, Uddn1, Theta, &d);
      if (c->E0 < 0.0) c->E0 = d.energy;
      d.de_rel = (c->E0 > 0.0) ? (d.energy - c->E0) / c->E0 : 0.0;
      if (recs && nrec < max_recs) recs[nrec] = d;
      nrec++;
      if (isnan(d.energy) || isinf(d.energy) || isnan(d.max_u) || isinf(d.max_u)) bad = 1;
    }

    /* rotation (:1145-1163) */
    SWAP(c->Unm1, c->Un);   /* Unm1 <- Un, then Un <- Un1 */
    SWAP(c->Un, c->Un1);
    SWAP(c->Udnm1, c->Udn);
    SWAP(c->Udn, c->Udn1);
    SWAP(c->Uddn, c->Uddn1);
    if (c->p.theta_method == 1) {
      SWAP(c->ThetaNm1, c->ThetaN);
      for (int32_t i = 0; i < N; i++) c->ThetaN[i] = c->Theta[i];
    }
    if (bad) return -1;
  }
  return nrec;
}

/* After a step the reference's Un holds U^{n+1}; oracle_fetch returns that state. */
void oracle_fetch(oracle_ctx *c, double *u, double *v, double *a, double *theta, double *m) {
  const size_t n2 = sizeof(double) * 2 * (size_t)c->p.N;
  if (u) memcpy(u, c->Un, n2);
  if (v) memcpy(v, c->Udn, n2);
  if (a) memcpy(a, c->Uddn, n2);
  if (theta) memcpy(theta, c->Theta, sizeof(double) * (size_t)c->p.N);
  if (m) memcpy(m, c->m, sizeof(double) * (size_t)c->p.N);
}

int64_t oracle_step_index(oracle_ctx *c) { return c->s; }
double oracle_m_ref(oracle_ctx *c) { return c->m_ref; }

int64_t oracle_horizons(oracle_ctx *c, int64_t *row_ptr, int32_t *cols) {
  int64_t nnz = c->hp[c->p.N];
  if (row_ptr) memcpy(row_ptr, c->hp, sizeof(int64_t) * ((size_t)c->p.N + 1));
  if (cols) memcpy(cols, c->hc, sizeof(int32_t) * (size_t)nnz);
  return nnz;
}

void oracle_dilatation(oracle_ctx *c, const double *u, double *theta) {
#pragma omp parallel for
  for (int32_t i = 0; i < c->p.N; i++) theta[i] = theta_node(c, u, i);
}

void oracle_force(oracle_ctx *c, const double *u, const double *theta, double *acc) {
#pragma omp parallel for
  for (int32_t i = 0; i < c->p.N; i++) force_node(c, u, theta, i, acc);
}
