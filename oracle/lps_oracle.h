/* CPU oracle of the LPS-ABC solver's time loop.
 *
 * TEST INFRASTRUCTURE ONLY. This restatement of src/main.cpp (the reference solver) is the checker
 * for the HIP implementation: only tests, the smoke entry and the CPU-baseline leg of the bench may
 * load it. The product path never calls it.
 *
 * Every function names the reference lines it follows. Pinning of each function (what checks it
 * against the reference or the mathematics) is listed at the top of lps_oracle.c.
 */
#ifndef LPS_ORACLE_H
#define LPS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_problem {
  int32_t N, C, M;            /* nodes, ABC nodes, max cloud size */
  double E, nu, rho, dt, delta, dx, dy;
  const double *nodes;        /* N x 2 */
  const int32_t *bctypes;     /* 2N */
  const double *ic, *icd;     /* 2N each, interleaved (x, y) */
  const int32_t *abc;         /* C */
  const int32_t *clouds;      /* C x M, after the reader's (int)(x + 0.5) rounding */
  const double *Gu, *Gv;      /* 2C x 4M */
  const double *Gth;          /* C x 2M (NULL when theta_method == 0) */
  int32_t theta_method;       /* 1: G_theta (src/main.cpp:38 default), 0: IDW */
  int32_t abc_enabled;        /* 1: ABCS_STEP == 2, 0: ABCS_STEP == -1 */
  int32_t n_notch;            /* prenotch segments (.data1[2]) */
  const double *notch;        /* n_notch x 4: x0 y0 x1 y1 (.prenotch) */
  int32_t plant;              /* tests only: 0 the reference's algorithm; 1 plants a wrong IDW
                                 weight (1/d^2 for 1/d), so a test can show its fixture catches it */
} oracle_problem;

/* One output record, the quantities of src/main.cpp:986-1062. */
typedef struct oracle_diag {
  int64_t step;
  double t, energy, energy_k, energy_p, px, py, max_u, max_v, max_a, theta_max, theta_min;
  double ub_rms, vb_rms, de_rel;
  int32_t max_u_node, max_v_node;
} oracle_diag;

typedef struct oracle_ctx oracle_ctx;

/* Horizons: for every node the nodes within sqrt(dx^2+dy^2) <= r (self included, first), sorted by
 * that distance, ties by ascending index. Returns nnz (self included) or -1 when a row holds more
 * than max_len entries. row_ptr has N+1 entries; cols is filled only when non-NULL. */
int64_t oracle_build_horizons(int32_t N, const double *nodes, double r, int32_t max_len,
                              int64_t *row_ptr, int32_t *cols);

/* Prenotch visibility (src/main.cpp:119-140, 386-402): 1 when the segment (x1,y1)-(x2,y2)
 * intersects or touches the segment (x3,y3)-(x4,y4). */
int oracle_segments_intersect(double x1, double y1, double x2, double y2, double x3, double y3,
                              double x4, double y4);

oracle_ctx *oracle_create(const oracle_problem *p, int32_t threads);
void oracle_destroy(oracle_ctx *c);
const char *oracle_error(void);

/* Advance nsteps time steps. Records produced at steps s % 30 == 0 are written to recs (if not
 * NULL, up to max_recs); returns the number of records, or -1 on NaN/Inf (src/main.cpp:1126-1133). */
int32_t oracle_run(oracle_ctx *c, int32_t nsteps, oracle_diag *recs, int32_t max_recs);

/* Diagnostics of the current state (the output block of src/main.cpp:986-1062), without
 * touching E0. */
void oracle_diagnostics(oracle_ctx *c, oracle_diag *out);

/* Current fields, original node order: u, v, a interleaved (2N), theta (N), m (N).
 * Any pointer may be NULL. */
void oracle_fetch(oracle_ctx *c, double *u, double *v, double *a, double *theta, double *m);
int64_t oracle_step_index(oracle_ctx *c);
double oracle_m_ref(oracle_ctx *c);
int64_t oracle_horizons(oracle_ctx *c, int64_t *row_ptr, int32_t *cols);

/* Single passes on caller arrays, for pass-level tests (original order, interleaved fields). */
void oracle_dilatation(oracle_ctx *c, const double *u, double *theta);
void oracle_force(oracle_ctx *c, const double *u, const double *theta, double *acc);

#ifdef __cplusplus
}
#endif
#endif
